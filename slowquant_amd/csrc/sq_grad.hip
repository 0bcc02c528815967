// Energy + theta gradient drivers (replaces ups_wavefunction.py:1019-1142
// for the theta block) and workspace sizing.
//
//   adjoint          forward sweep psi = U psi0, lambda = H psi, reverse sweep
//                    g_i = 2 <lambda|T_i psi>, un-applying gate i on psi and
//                    lambda in place (no stored intermediates).
//   plain            per-parameter derivative propagation from scratch,
//                    d_i psi = U_N..U_{i+1} U_i T_i U_{i-1}..U_1 psi0 (O(N^2)).
//   reference sweep  the reference's product-rule loop (ups_wavefunction.py:1090-1141).

#include <cstring>
#include <mutex>
#include <vector>

#include "sq_internal.h"

namespace sq {

// Gradient workspace (doubles): psi | lam | t1 | t2 | s1 | s2 | partials | scalars(64) | grad(N) | sigma work
static int64_t grad_doubles(const Space& s, int n_gates) {
  return 6 * s.dim + nblocks2d(s) + 1024 + 64 + n_gates + sigma_work_bytes(s) / 8;
}

struct GradBufs {
  double *psi, *lam, *t1, *t2, *s1, *s2, *partial, *scal, *grad, *sigw;
};

static GradBufs carve(const Space& s, int n_gates, double* work) {
  GradBufs b;
  b.psi = work;
  b.lam = b.psi + s.dim;
  b.t1 = b.lam + s.dim;
  b.t2 = b.t1 + s.dim;
  b.s1 = b.t2 + s.dim;
  b.s2 = b.s1 + s.dim;
  b.partial = b.s2 + s.dim;
  b.scal = b.partial + nblocks2d(s) + 1024;
  b.grad = b.scal + 64;
  b.sigw = b.grad + n_gates;
  return b;
}

// Device -> host copy of n doubles through a pinned staging buffer: one DMA
// and no pageable bounce buffer; synchronizes the stream.
static void copy_to_host_pinned(double* host, const double* dev, size_t n, hipStream_t st) {
  static std::mutex mu;
  static double* buf = nullptr;
  static size_t cap = 0;
  std::lock_guard<std::mutex> lock(mu);
  if (n > cap) {
    if (buf) (void)hipHostFree(buf);
    buf = nullptr;
    cap = 0;
    SQ_HIP(hipHostMalloc(reinterpret_cast<void**>(&buf), sizeof(double) * n, hipHostMallocDefault));
    cap = n;
  }
  SQ_HIP(hipMemcpyAsync(buf, dev, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  std::memcpy(host, buf, sizeof(double) * n);
}

// Energy (scal[0]) and gradient (scal + 64) to the host in one copy.
static void finish(const GradBufs& b, int n, double* energy, double* grad, hipStream_t st) {
  std::vector<double> h(64 + (size_t)n);
  copy_to_host_pinned(h.data(), b.scal, h.size(), st);
  *energy = h[0];
  if (n > 0) std::memcpy(grad, h.data() + 64, sizeof(double) * n);
}

// psi = U psi0 (t1, t2: per-gate work; s1: tile scratch).  deriv_gate >= 0
// inserts T_g after gate g.
static void forward(const Space& s, const Program& p, const double* theta, const double* psi0, double* psi,
                    const GradBufs& b, int deriv_gate, hipStream_t st) {
  const TilePlan* T = tiles_for(s, p);
  if (T) {
    tiles_forward(s, p, *T, theta, psi0, psi, b.s1, b.t1, deriv_gate, false, st);
    return;
  }
  copy_vec(psi0, psi, s.dim, st);
  for (int g = 0; g < (int)p.gates.size(); ++g) {
    apply_gate_impl(s, p, g, theta, false, psi, b.t1, st);
    if (g == deriv_gate) {
      grad_action_impl(s, p, g, psi, b.t2, st);
      copy_vec(b.t2, psi, s.dim, st);
    }
  }
}

void energy_grad_adjoint(const Space& s, const Program& p, const Hamiltonian& H, const double* theta,
                         const double* psi0, double* energy, double* grad, double* work, hipStream_t st) {
  const int n = (int)p.gates.size();
  GradBufs b = carve(s, n, work);
  forward(s, p, theta, psi0, b.psi, b, -1, st);
  sigma_energy_dev(s, H, b.psi, b.lam, b.scal, b.sigw, st);
  const TilePlan* T = tiles_for(s, p);
  if (T) {
    tiles_reverse(s, p, *T, theta, b.psi, b.lam, b.t1, b.partial, b.grad, true, st);  // matrices from forward()
  } else {
    for (int g = n - 1; g >= 0; --g) reverse_step_impl(s, p, g, theta, b.psi, b.lam, b.grad, b.t1, b.partial, st);
  }
  finish(b, n, energy, grad, st);
}

// Reverse half of the adjoint on caller-owned psi = U psi0 and lam = H psi
// (consumed in place); dE/dtheta to host.
void adjoint_reverse(const Space& s, const Program& p, const double* theta, double* psi, double* lam, double* grad,
                     double* work, hipStream_t st) {
  const int n = (int)p.gates.size();
  GradBufs b = carve(s, n, work);
  const TilePlan* T = tiles_for(s, p);
  if (T) {
    tiles_reverse(s, p, *T, theta, psi, lam, b.t1, b.partial, b.grad, false, st);
  } else {
    for (int g = n - 1; g >= 0; --g) reverse_step_impl(s, p, g, theta, psi, lam, b.grad, b.t1, b.partial, st);
  }
  if (n > 0) copy_to_host_pinned(grad, b.grad, (size_t)n, st);
  else SQ_HIP(hipStreamSynchronize(st));
}

void energy_grad_plain(const Space& s, const Program& p, const Hamiltonian& H, const double* theta,
                       const double* psi0, double* energy, double* grad, double* work, hipStream_t st) {
  const int n = (int)p.gates.size();
  GradBufs b = carve(s, n, work);
  forward(s, p, theta, psi0, b.psi, b, -1, st);
  sigma_energy_dev(s, H, b.psi, b.lam, b.scal, b.sigw, st);
  // b.psi is free now: it holds each derivative vector d_i psi in turn.
  double* v = b.psi;
  for (int i = 0; i < n; ++i) {
    forward(s, p, theta, psi0, v, b, i, st);
    dot_dev(b.lam, v, s.dim, b.grad, i, 2.0, b.partial, st);
  }
  finish(b, n, energy, grad, st);
}

void energy_grad_reference_sweep(const Space& s, const Program& p, const Hamiltonian& H, const double* theta,
                                 const double* psi0, double* energy, double* grad, double* work, hipStream_t st) {
  const int n = (int)p.gates.size();
  GradBufs b = carve(s, n, work);
  forward(s, p, theta, psi0, b.psi, b, -1, st);
  sigma_energy_dev(s, H, b.psi, b.lam, b.scal, b.sigw, st);
  // bra = U^+ H psi
  const TilePlan* T = tiles_for(s, p);
  if (T) {
    tiles_forward(s, p, *T, theta, b.lam, b.lam, b.s1, b.t1, -1, true, st);
  } else {
    for (int g = n - 1; g >= 0; --g) apply_gate_impl(s, p, g, theta, true, b.lam, b.t1, st);
  }
  double* ket = b.psi;
  copy_vec(psi0, ket, s.dim, st);
  for (int i = 0; i < n; ++i) {
    int64_t nb = tdot_impl(s, p, i, ket, b.lam, b.partial, st);
    reduce_impl(b.partial, nb, b.grad, i, false, 2.0, st);
    apply_gate_impl(s, p, i, theta, false, b.lam, b.t1, st);
    apply_gate_impl(s, p, i, theta, false, ket, b.t1, st);
  }
  finish(b, n, energy, grad, st);
}

}  // namespace sq

using namespace sq;

extern "C" {

int64_t sq_work_bytes(const sq_space* sp, int what) {
  if (!sp) return -1;
  const Space& s = sp->s;
  switch (what) {
    case 0:  // state preparation / single gates
      return 8 * (3 * s.dim + 64);
    case 1:  // sigma
      return sigma_work_bytes(s);
    case 2:  // gradients, for programs of up to 65536 gates
      return 8 * grad_doubles(s, 65536);
    case 3:  // RDMs
      return rdm_work_bytes(s, nullptr, nullptr, nullptr);
    default:
      return -1;
  }
}

#define SQ_GRAD_ENTRY(name, impl)                                                                               \
  int name(const sq_space* sp, const sq_program* pr, const sq_hamiltonian* H, const double* theta,            \
           const double* psi0, double* energy, double* grad, double* work, void* stream) {                    \
    return guarded([&] {                                                                                        \
      SQ_CHECK(sp && pr && H && psi0 && energy && work, SQ_ERR_ARG, "NULL argument");                          \
      SQ_CHECK(pr->p.space == &sp->s && H->h.space == &sp->s, SQ_ERR_ARG,                                       \
               "program / Hamiltonian built for another space");                                                \
      SQ_CHECK(theta || pr->p.gates.empty(), SQ_ERR_ARG, "theta is NULL");                                      \
      SQ_CHECK(grad || pr->p.gates.empty(), SQ_ERR_ARG, "grad is NULL");                                        \
      SQ_CHECK((int)pr->p.gates.size() <= 65536, SQ_ERR_ARG, "too many gates");                                 \
      impl(sp->s, pr->p, H->h, theta, psi0, energy, grad, work, as_stream(stream));                             \
    });                                                                                                         \
  }

SQ_GRAD_ENTRY(sq_energy_grad_adjoint, energy_grad_adjoint)
SQ_GRAD_ENTRY(sq_energy_grad_plain, energy_grad_plain)
SQ_GRAD_ENTRY(sq_energy_grad_reference_sweep, energy_grad_reference_sweep)

int sq_adjoint_reverse(const sq_space* sp, const sq_program* pr, const double* theta, double* psi, double* lam,
                       double* grad, double* work, void* stream) {
  return guarded([&] {
    SQ_CHECK(sp && pr && psi && lam && work, SQ_ERR_ARG, "NULL argument");
    SQ_CHECK(pr->p.space == &sp->s, SQ_ERR_ARG, "program built for another space");
    SQ_CHECK((theta && grad) || pr->p.gates.empty(), SQ_ERR_ARG, "theta or grad is NULL");
    SQ_CHECK((int)pr->p.gates.size() <= 65536, SQ_ERR_ARG, "too many gates");
    adjoint_reverse(sp->s, pr->p, theta, psi, lam, grad, work, as_stream(stream));
  });
}

}  // extern "C"
