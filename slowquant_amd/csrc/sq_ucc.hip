// Unfactorised UCC on the device (replaces construct_ucc_state /
// get_ucc_T / build_operator_matrix + scipy expm_multiply,
// operator_state_algebra.py:413-469,870-960, and the finite-difference theta
// gradient of ucc_wavefunction.py:1025-1098).
//
// T(theta) = sum_k theta_k T_k with every gate generator T_k taken from the
// program's raw terms; exp(t T) v is applied matrix-free by a scaled Taylor
// series: s steps of exp(t T / s) with ||t T / s||_1 <= 1/2 (bound
// sum_terms |t theta_k coef|, each raw string has at most one unit entry per
// column), K terms per step until (1/2)^K / K! < 2^-60.
// The analytic gradient uses the block-triangular generator
// M = [[T, T_i], [0, T]]:  exp(M) [0; v] = [d/dtheta_i exp(T) v; exp(T) v].

#include <cmath>

#include "sq_internal.h"

namespace sq {

constexpr int kUB = 256;

struct SV {
  int nsa, nsb;
  const uint32_t* astr;
  const uint32_t* bstr;
  const int32_t* arank;
  const int32_t* brank;
};

__device__ __forceinline__ bool ops_adjoint(const SpinOps& o, uint32_t& m, int& par) {
  for (int j = 0; j < o.n; ++j) {
    uint32_t bit = 1u << o.orb[j];
    par += __popc(m & (bit - 1u));
    if (!o.cre[j]) {
      if (m & bit) return false;
      m |= bit;
    } else {
      if (!(m & bit)) return false;
      m &= ~bit;
    }
  }
  return true;
}

// nxt[I] = f * sum_t w[t] <I|O_t|J> cur[J] (+ f2 * sum over terms [u0,u1) with
// unit weights applied to cur2); acc[I] += nxt[I] when acc != nullptr.
__global__ void k_tapply_w(SV v, const Term* __restrict__ terms, const double* __restrict__ w, int t0, int t1,
                           const double* __restrict__ cur, int u0, int u1, const double* __restrict__ cur2,
                           double f, double* __restrict__ nxt, double* __restrict__ acc) {
  int ia = blockIdx.y;
  int ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= v.nsb) return;
  uint32_t a0 = v.astr[ia], b0 = v.bstr[ib];
  double val = 0.0;
  for (int t = t0; t < t1; ++t) {
    const Term& T = terms[t];
    uint32_t a = a0, b = b0;
    int par = 0;
    if (!ops_adjoint(T.a, a, par) || !ops_adjoint(T.b, b, par)) continue;
    int ra = v.arank[a], rb = v.brank[b];
    if (ra < 0 || rb < 0) continue;
    double c = w[t] * T.coef;
    val += ((par & 1) ? -c : c) * cur[(size_t)ra * v.nsb + rb];
  }
  if (cur2) {
    for (int t = u0; t < u1; ++t) {
      const Term& T = terms[t];
      uint32_t a = a0, b = b0;
      int par = 0;
      if (!ops_adjoint(T.a, a, par) || !ops_adjoint(T.b, b, par)) continue;
      int ra = v.arank[a], rb = v.brank[b];
      if (ra < 0 || rb < 0) continue;
      val += ((par & 1) ? -T.coef : T.coef) * cur2[(size_t)ra * v.nsb + rb];
    }
  }
  val *= f;
  size_t I = (size_t)ia * v.nsb + ib;
  nxt[I] = val;
  if (acc) acc[I] += val;
}

// w[t] = theta[gate of term t]
__global__ void k_term_weights(const GateDev* __restrict__ gates, int ngates, const double* __restrict__ theta,
                               double* __restrict__ w) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngates) return;
  for (int t = gates[g].term_begin; t < gates[g].term_end; ++t) w[t] = theta[g];
}

__global__ void k_zero(double* __restrict__ x, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

static SV sv(const Space& s) { return SV{s.nsa, s.nsb, s.d_astr, s.d_bstr, s.d_arank, s.d_brank}; }

struct TaylorPlan {
  int steps, K;
  double a;  // t / steps
};

static TaylorPlan taylor_plan(const Program& p, const std::vector<double>& theta, double t) {
  double b = 0.0;
  for (size_t g = 0; g < p.gates.size(); ++g)
    for (int k = p.gates[g].term_begin; k < p.gates[g].term_end; ++k) b += std::fabs(t * theta[g] * p.terms[k].coef);
  TaylorPlan P;
  P.steps = std::max(1, (int)std::ceil(b / 0.5));
  const double r = b / P.steps;
  P.K = 1;
  double term = r;
  while (term > std::ldexp(1.0, -60) && P.K < 60) {
    ++P.K;
    term *= r / P.K;
  }
  P.a = t / P.steps;
  return P;
}

// out = exp(t T(theta)) in.  work >= 3*dim + nterms doubles.
void ucc_expm_impl(const Space& s, const Program& p, const double* theta_dev, double t, const double* in,
                   double* out, double* work, hipStream_t st) {
  const int ng = (int)p.gates.size();
  std::vector<double> th(ng);
  if (ng) SQ_HIP(hipMemcpyAsync(th.data(), theta_dev, sizeof(double) * ng, hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  TaylorPlan P = taylor_plan(p, th, t);
  double* A = work;
  double* B = A + s.dim;
  double* w = B + s.dim;
  const int nt = (int)p.terms.size();
  if (ng) k_term_weights<<<(ng + 255) / 256, 256, 0, st>>>(p.d_gates, ng, theta_dev, w);
  SQ_HIP(hipGetLastError());
  copy_vec(in, out, s.dim, st);
  dim3 grid((s.nsb + kUB - 1) / kUB, s.nsa);
  for (int step = 0; step < P.steps; ++step) {
    copy_vec(out, A, s.dim, st);
    double* cur = A;
    double* nxt = B;
    for (int k = 1; k <= P.K; ++k) {
      SQ_PROF(PC_TAPPLY, st, 24.0 * (double)s.dim, 0.0,
              k_tapply_w<<<grid, kUB, 0, st>>>(sv(s), p.d_terms, w, 0, nt, cur, 0, 0, nullptr, P.a / k, nxt, out));
      std::swap(cur, nxt);
    }
  }
}

// Analytic dE/dtheta for E = <psi|H|psi>, psi = exp(T(theta)) psi0, given
// lam = H psi: grad_dev[i] = 2 <lam| d_i psi>.  work >= 6*dim + nterms + partials.
void ucc_grad_impl(const Space& s, const Program& p, const double* theta_dev, const double* psi0, const double* lam,
                   double* grad_dev, double* work, double* partial, hipStream_t st) {
  const int ng = (int)p.gates.size();
  std::vector<double> th(ng);
  if (ng) SQ_HIP(hipMemcpyAsync(th.data(), theta_dev, sizeof(double) * ng, hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  TaylorPlan P = taylor_plan(p, th, 1.0);
  // the T_i block adds at most ||T_i|| to the bound; refine per gate
  double* top = work;
  double* bot = top + s.dim;
  double* tA = bot + s.dim;
  double* tB = tA + s.dim;
  double* bA = tB + s.dim;
  double* bB = bA + s.dim;
  double* w = bB + s.dim;
  const int nt = (int)p.terms.size();
  if (ng) k_term_weights<<<(ng + 255) / 256, 256, 0, st>>>(p.d_gates, ng, theta_dev, w);
  SQ_HIP(hipGetLastError());
  dim3 grid((s.nsb + kUB - 1) / kUB, s.nsa);
  for (int i = 0; i < ng; ++i) {
    const GateDev& gi = p.gates[i];
    double bi = 0.0;
    for (int k = gi.term_begin; k < gi.term_end; ++k) bi += std::fabs(p.terms[k].coef);
    double b = 0.0;
    for (size_t g = 0; g < p.gates.size(); ++g)
      for (int k = p.gates[g].term_begin; k < p.gates[g].term_end; ++k) b += std::fabs(th[g] * p.terms[k].coef);
    b += bi;
    int steps = std::max(1, (int)std::ceil(b / 0.5));
    double r = b / steps;
    int K = 1;
    double term = r;
    while (term > std::ldexp(1.0, -60) && K < 60) {
      ++K;
      term *= r / K;
    }
    const double a = 1.0 / steps;
    // [top; bot] = [0; psi0]
    k_zero<<<(unsigned)((s.dim + kUB - 1) / kUB), kUB, 0, st>>>(top, s.dim);
    copy_vec(psi0, bot, s.dim, st);
    for (int step = 0; step < steps; ++step) {
      copy_vec(top, tA, s.dim, st);
      copy_vec(bot, bA, s.dim, st);
      double *tc = tA, *tn = tB, *bc = bA, *bn = bB;
      for (int k = 1; k <= K; ++k) {
        // new top = a/k (T top + T_i bot); new bot = a/k T bot
        k_tapply_w<<<grid, kUB, 0, st>>>(sv(s), p.d_terms, w, 0, nt, tc, gi.term_begin, gi.term_end, bc, a / k, tn,
                                         top);
        k_tapply_w<<<grid, kUB, 0, st>>>(sv(s), p.d_terms, w, 0, nt, bc, 0, 0, nullptr, a / k, bn, bot);
        SQ_HIP(hipGetLastError());
        std::swap(tc, tn);
        std::swap(bc, bn);
      }
    }
    dot_dev(lam, top, s.dim, grad_dev, i, 2.0, partial, st);
  }
}

}  // namespace sq

using namespace sq;

extern "C" {

int sq_ucc_expm(const sq_space* sp, const sq_program* pr, const double* theta, double t, const double* in,
                double* out, void* work_v, void* stream) {
  double* work = static_cast<double*>(work_v);
  return guarded([&] {
    SQ_CHECK(sp && pr && in && out && work && (theta || pr->p.gates.empty()), SQ_ERR_ARG, "NULL argument");
    SQ_CHECK(in != out, SQ_ERR_ARG, "in and out must differ");
    ucc_expm_impl(sp->s, pr->p, theta, t, in, out, work, as_stream(stream));
  });
}

int sq_ucc_energy_grad(const sq_space* sp, const sq_program* pr, const sq_hamiltonian* H, const double* theta,
                       const double* psi0, double* energy, double* grad, void* work_v, void* stream) {
  double* work = static_cast<double*>(work_v);
  return guarded([&] {
    SQ_CHECK(sp && pr && H && psi0 && energy && work && (theta || pr->p.gates.empty()), SQ_ERR_ARG, "NULL argument");
    const Space& s = sp->s;
    const int ng = (int)pr->p.gates.size();
    hipStream_t st = as_stream(stream);
    double* psi = work;
    double* lam = psi + s.dim;
    double* scal = lam + s.dim;            // 64
    double* gdev = scal + 64;              // ng
    double* partial = gdev + std::max(ng, 1) + 1024;  // 1024
    double* rest = partial + 1024;
    ucc_expm_impl(s, pr->p, theta, 1.0, psi0, psi, rest, st);
    sigma_energy_dev(s, H->h, psi, lam, scal, rest, st);
    if (grad && ng) {
      ucc_grad_impl(s, pr->p, theta, psi0, lam, gdev, rest, partial, st);
      SQ_HIP(hipMemcpyAsync(grad, gdev, sizeof(double) * ng, hipMemcpyDeviceToHost, st));
    }
    SQ_HIP(hipMemcpyAsync(energy, scal, sizeof(double), hipMemcpyDeviceToHost, st));
    SQ_HIP(hipStreamSynchronize(st));
  });
}

int64_t sq_ucc_work_bytes(const sq_space* sp, const sq_program* pr) {
  if (!sp || !pr) return -1;
  const Space& s = sp->s;
  int64_t nt = (int64_t)pr->p.terms.size() + 64;
  int64_t ng = (int64_t)pr->p.gates.size() + 1;
  int64_t body = std::max<int64_t>(6 * s.dim + nt, sigma_work_bytes(s) / 8);
  return 8 * (2 * s.dim + 64 + ng + 2048 + body);
}

}  // extern "C"
