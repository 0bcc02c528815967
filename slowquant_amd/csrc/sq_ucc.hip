// Unfactorised UCC on the device (replaces construct_ucc_state /
// get_ucc_T / build_operator_matrix + scipy expm_multiply,
// operator_state_algebra.py:413-469,870-960, and the finite-difference theta
// gradient of ucc_wavefunction.py:1025-1098).
//
// T(theta) = sum_k theta_k T_k with every gate generator T_k taken from the
// program's raw terms; exp(t T) v is applied matrix-free by a scaled Taylor
// series: s steps of exp(t T / s) with |t| nu / s <= 2 and K terms per step
// until (|t| nu / s)^K / K! < 2^-60, where nu bounds ||T||_2 (a power-iteration
// estimate with margin, capped by the triangle-inequality bound).
// The analytic gradient integrates <e^{-sT} lam | T_i | e^{-sT} psi> over
// s in [0, 1] by composite Gauss-Legendre (lam = H psi), all gates per node.
// T's sparsity pattern does not depend on theta: it is built once per program
// (rows of (source determinant, term, sign)) and every T application is then
// one weighted sparse matrix-vector product; patterns above 2^29 entries fall
// back to kernels that rescan the terms per determinant.

#include <cmath>
#include <cstdlib>
#include <vector>

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

// Adjoint of term t on strings (a0, b0): images (a, b) and sign parity;
// false if it does not apply.
__device__ __forceinline__ bool term_adjoint(const TermFast& F, const Term* __restrict__ terms, int t, uint32_t a0,
                                             uint32_t b0, uint32_t& a, uint32_t& b, int& par) {
  if (F.flags & 1u) {
    if ((a0 & F.sa) != F.sa || (a0 & F.ca) || (b0 & F.sb) != F.sb || (b0 & F.cb)) return false;
    a = a0 ^ F.sa ^ F.ca;
    b = b0 ^ F.sb ^ F.cb;
    par = __popc(a0 & F.xa) + __popc(b0 & F.xb) + (int)(F.flags >> 1);
    return true;
  }
  const Term& T = terms[t];
  a = a0;
  b = b0;
  par = 0;
  return ops_adjoint(T.a, a, par) && ops_adjoint(T.b, b, par);
}

// part[y][I] = sum over the terms t = g, g + G, ... of w[t] <I|O_t|J> cur[J]
// summed over the groups g = y*nw + wave of one block (w = nullptr: unit
// weights).  A block is nw = min(G, kTG) waves over the same 64 determinants
// (lane = determinant, flattened across rows), one term group per wave; the
// waves' sums are added in LDS in wave order.  Splitting the terms over G groups
// keeps small spaces' grids full; k_tapply_sum adds the part rows in order.
constexpr int kTG = 8;

__global__ __launch_bounds__(64 * kTG) void k_tapply_part(SV v, const Term* __restrict__ terms,
                                                         const TermFast* __restrict__ fast,
                                                         const double* __restrict__ w, int nt, int G,
                                                         const double* __restrict__ cur, double* __restrict__ part) {
  __shared__ double red[kTG][64];
  const size_t dim = (size_t)v.nsa * v.nsb;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const size_t I = (size_t)blockIdx.x * 64 + lane;
  const int g = blockIdx.y * nw + wave;
  double val = 0.0;
  if (I < dim && g < G) {
    const int ia = (int)(I / (uint32_t)v.nsb), ib = (int)(I - (size_t)ia * v.nsb);
    const uint32_t a0 = v.astr[ia], b0 = v.bstr[ib];
    for (int t = g; t < nt; t += G) {
      const TermFast F = fast[t];
      uint32_t a, b;
      int par;
      if (!term_adjoint(F, terms, t, a0, b0, a, b, par)) continue;
      const int ra = v.arank[a], rb = v.brank[b];
      if (ra < 0 || rb < 0) continue;
      const double c = (w ? w[t] : 1.0) * F.coef;
      val += ((par & 1) ? -c : c) * cur[(size_t)ra * v.nsb + rb];
    }
  }
  red[wave][lane] = val;
  __syncthreads();
  if (wave == 0 && I < dim) {
    double s = 0.0;
    for (int q = 0; q < nw; ++q) s += red[q][lane];
    part[(size_t)blockIdx.y * dim + I] = s;
  }
}

// part[g][I] for one term group g = blockIdx.y per block (all waves of a
// block on the same term stream, so they share its scalar-cache lines); used
// when there are too few groups to fold (large spaces).
__global__ __launch_bounds__(kUB) void k_tapply_group(SV v, const Term* __restrict__ terms,
                                                     const TermFast* __restrict__ fast,
                                                     const double* __restrict__ w, int nt, int G,
                                                     const double* __restrict__ cur, double* __restrict__ part) {
  const size_t dim = (size_t)v.nsa * v.nsb;
  const size_t I = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= dim) return;
  const int ia = (int)(I / (uint32_t)v.nsb), ib = (int)(I - (size_t)ia * v.nsb);
  const uint32_t a0 = v.astr[ia], b0 = v.bstr[ib];
  const int g = blockIdx.y;
  double val = 0.0;
  for (int t = g; t < nt; t += G) {
    const TermFast F = fast[t];
    uint32_t a, b;
    int par;
    if (!term_adjoint(F, terms, t, a0, b0, a, b, par)) continue;
    const int ra = v.arank[a], rb = v.brank[b];
    if (ra < 0 || rb < 0) continue;
    const double c = (w ? w[t] : 1.0) * F.coef;
    val += ((par & 1) ? -c : c) * cur[(size_t)ra * v.nsb + rb];
  }
  part[(size_t)g * dim + I] = val;
}

// T's sparsity pattern, row I = the terms whose adjoint maps I to a
// determinant J of the space, in term order: count, then fill the entries
// (J, t | parity << 31) between rowp[I] and rowp[I + 1].
__device__ __forceinline__ bool term_entry(const SV& v, const Term* __restrict__ terms,
                                           const TermFast* __restrict__ fast, int t, uint32_t a0, uint32_t b0,
                                           uint32_t& J, int& par) {
  uint32_t a, b;
  if (!term_adjoint(fast[t], terms, t, a0, b0, a, b, par)) return false;
  const int ra = v.arank[a], rb = v.brank[b];
  if (ra < 0 || rb < 0) return false;
  J = (uint32_t)ra * (uint32_t)v.nsb + (uint32_t)rb;
  return true;
}

__global__ void k_csr_count(SV v, const Term* __restrict__ terms, const TermFast* __restrict__ fast, int nt,
                            int32_t* __restrict__ cnt) {
  const size_t dim = (size_t)v.nsa * v.nsb;
  const size_t I = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= dim) return;
  const int ia = (int)(I / (uint32_t)v.nsb), ib = (int)(I - (size_t)ia * v.nsb);
  const uint32_t a0 = v.astr[ia], b0 = v.bstr[ib];
  int c = 0;
  for (int t = 0; t < nt; ++t) {
    uint32_t J;
    int par;
    if (term_entry(v, terms, fast, t, a0, b0, J, par)) ++c;
  }
  cnt[I] = c;
}

__global__ void k_csr_fill(SV v, const Term* __restrict__ terms, const TermFast* __restrict__ fast, int nt,
                           const int64_t* __restrict__ rowp, uint2* __restrict__ ent) {
  const size_t dim = (size_t)v.nsa * v.nsb;
  const size_t I = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= dim) return;
  const int ia = (int)(I / (uint32_t)v.nsb), ib = (int)(I - (size_t)ia * v.nsb);
  const uint32_t a0 = v.astr[ia], b0 = v.bstr[ib];
  int64_t k = rowp[I];
  const int64_t e = rowp[I + 1];
  for (int t = 0; t < nt && k < e; ++t) {
    uint32_t J;
    int par;
    if (term_entry(v, terms, fast, t, a0, b0, J, par)) ent[k++] = make_uint2(J, (uint32_t)t | ((uint32_t)(par & 1) << 31));
  }
}

// nxt[I] = f sum_{entries (J, t, sign) of row I} sign w[t] coef[t] cur[J] and
// acc[I] += nxt[I] (acc != nullptr), one wave per row: lane l takes entries
// l, l + 64, ... and the wave sums in a fixed shuffle order.
constexpr int kCsrRows = 4;  // rows (waves) per block

__global__ __launch_bounds__(64 * kCsrRows) void k_tapply_csr(int64_t dim, const int64_t* __restrict__ rowp,
                                                              const uint2* __restrict__ ent,
                                                              const double* __restrict__ coef,
                                                              const double* __restrict__ w,
                                                              const double* __restrict__ cur, double f,
                                                              double* __restrict__ nxt, double* __restrict__ acc) {
  const int lane = threadIdx.x & 63;
  const int64_t I = (int64_t)blockIdx.x * kCsrRows + (threadIdx.x >> 6);
  if (I >= dim) return;  // uniform per wave; no block barrier below
  const int64_t b = rowp[I], e = rowp[I + 1];
  double val = 0.0;
  for (int64_t k = b + lane; k < e; k += 64) {
    const uint2 x = ent[k];
    const uint32_t t = x.y & 0x7fffffffu;
    const double c = (w ? w[t] : 1.0) * coef[t] * cur[x.x];
    val += (x.y >> 31) ? -c : c;
  }
  for (int off = 32; off > 0; off >>= 1) val += __shfl_down(val, off, 64);
  if (lane == 0) {
    const double r = f * val;
    nxt[I] = r;
    if (acc) acc[I] += r;
  }
}

// Node outputs of one multi-node Taylor series: per vector, kNodes states
// (the 8 Gauss-Legendre nodes of a subinterval and its end) at out + q * dim,
// and the weight of the current Taylor term for each.
constexpr int kNodes = 9;
struct NodeW {
  double r[kNodes];
};

// Two vectors through the same T in one pass over the pattern:
// nxt_v = f T cur_v (v = 0, 1) and out_v[q][I] += r_q nxt_v[I] for q < kNodes.
// One wave per row as k_tapply_csr (same summation order); after the wave
// sum, lanes q and 32 + q (q < kNodes) update node q of vector 0 and 1.
__global__ __launch_bounds__(64 * kCsrRows) void k_tapply_csr2(
    int64_t dim, const int64_t* __restrict__ rowp, const uint2* __restrict__ ent, const double* __restrict__ coef,
    const double* __restrict__ w, const double* __restrict__ cur0, const double* __restrict__ cur1, double f,
    double* __restrict__ nxt0, double* __restrict__ nxt1, double* __restrict__ out0, double* __restrict__ out1,
    NodeW nw) {
  const int lane = threadIdx.x & 63;
  const int64_t I = (int64_t)blockIdx.x * kCsrRows + (threadIdx.x >> 6);
  if (I >= dim) return;  // uniform per wave; no block barrier below
  const int64_t b = rowp[I], e = rowp[I + 1];
  double v0 = 0.0, v1 = 0.0;
  for (int64_t k = b + lane; k < e; k += 64) {
    const uint2 x = ent[k];
    const uint32_t t = x.y & 0x7fffffffu;
    const double wc = w[t] * coef[t];
    const double c0 = wc * cur0[x.x], c1 = wc * cur1[x.x];
    v0 += (x.y >> 31) ? -c0 : c0;
    v1 += (x.y >> 31) ? -c1 : c1;
  }
  for (int off = 32; off > 0; off >>= 1) {
    v0 += __shfl_down(v0, off, 64);
    v1 += __shfl_down(v1, off, 64);
  }
  v0 = f * __shfl(v0, 0, 64);
  v1 = f * __shfl(v1, 0, 64);
  if (lane == 0) nxt0[I] = v0;
  if (lane == 32) nxt1[I] = v1;
  const int q = lane & 31;
  if (q < kNodes) {
    double* o = (lane < 32 ? out0 : out1) + (size_t)q * dim + I;
    *o += nw.r[q] * (lane < 32 ? v0 : v1);
  }
}

// out_v[q][I] = base_v[I] for q < kNodes (the k = 0 term of every node).
__global__ void k_nodes_init(int64_t dim, const double* __restrict__ b0, const double* __restrict__ b1,
                             double* __restrict__ out0, double* __restrict__ out1) {
  const int64_t I = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (I >= dim) return;
  const double x0 = b0[I], x1 = b1[I];
  for (int q = 0; q < kNodes; ++q) {
    out0[(size_t)q * dim + I] = x0;
    out1[(size_t)q * dim + I] = x1;
  }
}

// nxt[I] = f * sum_g part[g][I]; acc[I] += nxt[I] when acc != nullptr.
__global__ void k_tapply_sum(const double* __restrict__ part, int G, int64_t dim, double f, double* __restrict__ nxt,
                             double* __restrict__ acc) {
  const int64_t I = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (I >= dim) return;
  double v = 0.0;
  for (int g = 0; g < G; ++g) v += part[(size_t)g * dim + I];
  v *= f;
  nxt[I] = v;
  if (acc) acc[I] += v;
}

// Per gate i, per block b of amplitudes: dpart[b*ng + i] = sum over the
// block's I of v[I] (T_i u)[I] with T_i the gate's generator (unit weight).
// Block (b, y) covers determinants [256 b, 256 b + 256) and takes the gates
// i = y, y + gridDim.y, ...; every (b, i) is written by exactly one block and
// summed over b in a fixed order later.
__global__ void k_gate_dots(SV sv, const Term* __restrict__ terms, const TermFast* __restrict__ fast,
                            const GateDev* __restrict__ gates, int ng, const double* __restrict__ u,
                            const double* __restrict__ v, double* __restrict__ dpart) {
  __shared__ double red[kUB / 64];
  const size_t dim = (size_t)sv.nsa * sv.nsb;
  const size_t I = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // determinants flattened
  const bool ok = I < dim;
  const int ia = ok ? (int)(I / (uint32_t)sv.nsb) : 0;
  const int ib = ok ? (int)(I - (size_t)ia * sv.nsb) : 0;
  const int b = blockIdx.x;
  const uint32_t a0 = sv.astr[ia], b0 = sv.bstr[ib];
  const double vI = ok ? v[I] : 0.0;
  for (int i = blockIdx.y; i < ng; i += gridDim.y) {
    double val = 0.0;
    if (ok && vI != 0.0) {
      for (int t = gates[i].term_begin; t < gates[i].term_end; ++t) {
        const TermFast F = fast[t];
        uint32_t a, bb;
        int par;
        if (!term_adjoint(F, terms, t, a0, b0, a, bb, par)) continue;
        const int ra = sv.arank[a], rb = sv.brank[bb];
        if (ra < 0 || rb < 0) continue;
        val += ((par & 1) ? -F.coef : F.coef) * u[(size_t)ra * sv.nsb + rb];
      }
    }
    double c = vI * val;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += red[w];
      dpart[(size_t)b * ng + i] = tot;
    }
    __syncthreads();  // red is reused by the next gate
  }
}

// gsum[i] += f sum over gate i's pattern entries (I, J, t, sign) of
// sign coef[t] v[I] u[J], i.e. f <v|T_i u>; one block per gate, its entries
// (rows ascending) strided over the threads and summed in a fixed order.
__global__ __launch_bounds__(kUB) void k_gate_dots_csr(const int64_t* __restrict__ gptr,
                                                       const uint4* __restrict__ gent,
                                                       const double* __restrict__ coef,
                                                       const double* __restrict__ u,
                                                       const double* __restrict__ v, double f,
                                                       double* __restrict__ gsum) {
  __shared__ double red[kUB / 64];
  const int i = blockIdx.x;
  const int64_t e = gptr[i + 1];
  double val = 0.0;
  for (int64_t k = gptr[i] + threadIdx.x; k < e; k += blockDim.x) {
    const uint4 x = gent[k];
    const double c = coef[x.z & 0x7fffffffu] * v[x.x] * u[x.y];
    val += (x.z >> 31) ? -c : c;
  }
  for (int off = 32; off > 0; off >>= 1) val += __shfl_down(val, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = val;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += red[w];
    gsum[i] += f * tot;
  }
}

// Gate dots over the 8 nodes of one subinterval: block (s, i) takes split s
// of gate i's pattern entries (I, J, t, sign) and writes
// dpart[s * ng + i] = sum sign coef[t] sum_q wq[q] v_q[I] u_q[J]
// (v_q = v + q * dim, u_q likewise), the entries strided over the threads and
// summed in a fixed order; k_gate_accum adds the splits in order.
__global__ __launch_bounds__(kUB) void k_gate_dots_nodes(const int64_t* __restrict__ gptr,
                                                         const uint4* __restrict__ gent,
                                                         const double* __restrict__ coef,
                                                         const double* __restrict__ u, const double* __restrict__ v,
                                                         int64_t dim, NodeW wq, int ng, double* __restrict__ dpart) {
  __shared__ double red[kUB / 64];
  const int i = blockIdx.x, sp = blockIdx.y, S = gridDim.y;
  const int64_t g0 = gptr[i], g1 = gptr[i + 1];
  const int64_t chunk = (g1 - g0 + S - 1) / S;
  const int64_t b = g0 + sp * chunk, e = min(g1, b + chunk);
  double val = 0.0;
  for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
    const uint4 x = gent[k];
    double d = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) d += wq.r[q] * v[(size_t)q * dim + x.x] * u[(size_t)q * dim + x.y];
    const double c = coef[x.z & 0x7fffffffu] * d;
    val += (x.z >> 31) ? -c : c;
  }
  for (int off = 32; off > 0; off >>= 1) val += __shfl_down(val, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = val;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += red[w];
    dpart[(size_t)sp * ng + i] = tot;
  }
}

// gsum[i] += f * sum_b dpart[b*ng + i] (fixed order).
__global__ void k_gate_accum(const double* __restrict__ dpart, int nb, int ng, double f, double* __restrict__ gsum) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ng) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += dpart[(size_t)b * ng + i];
  gsum[i] += f * s;
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

// A fixed pseudo-random start vector in [-1, 1) (integer hash of the index).
__global__ void k_fill_pattern(double* __restrict__ x, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  h ^= h >> 31;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 29;
  x[i] = (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

__global__ void k_scale(double* __restrict__ x, int64_t n, const double* __restrict__ s, int slot) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] *= s[slot];
}

__global__ void k_inv_sqrt_slot(double* __restrict__ s, int from, int to) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s[to] = s[from] > 0.0 ? 1.0 / sqrt(s[from]) : 0.0;
}

static SV sv(const Space& s) { return SV{s.nsa, s.nsb, s.d_astr, s.d_bstr, s.d_arank, s.d_brank}; }

// Mask form of one spin part's adjoint (ops applied in order as in
// ops_adjoint: a creator needs its orbital occupied and clears it, an
// annihilator needs it empty and sets it).  Op j adds popc(m_j & below_j)
// with m_j = m ^ (bits of the ops before j); mod 2 that is popc(m & below_j)
// + popc(flipped_before_j & below_j), so the parity is popc(m & x) + c0 with
// x = XOR of the below_j masks and c0 a constant.  False if an orbital
// repeats.
static bool part_fast(const SpinOps& o, uint32_t& set, uint32_t& clear, uint32_t& x, int& c0) {
  set = clear = x = 0;
  c0 = 0;
  uint32_t flipped = 0;
  for (int j = 0; j < o.n; ++j) {
    const uint32_t bit = 1u << o.orb[j];
    if (flipped & bit) return false;
    x ^= bit - 1u;
    c0 += __builtin_popcount(flipped & (bit - 1u));
    flipped |= bit;
    (o.cre[j] ? set : clear) |= bit;
  }
  return true;
}

// Terms in mask form where their spin parts allow it (uploaded).
static TermFast* upload_term_fast(const std::vector<Term>& terms) {
  std::vector<TermFast> f(std::max<size_t>(terms.size(), 1), TermFast{0, 0, 0, 0, 0, 0, 0u, 0u, 0.0});
  for (size_t t = 0; t < terms.size(); ++t) {
    TermFast& F = f[t];
    F.coef = terms[t].coef;
    int ca0 = 0, cb0 = 0;
    if (part_fast(terms[t].a, F.sa, F.ca, F.xa, ca0) && part_fast(terms[t].b, F.sb, F.cb, F.xb, cb0))
      F.flags = 1u | ((uint32_t)((ca0 + cb0) & 1) << 1);
    else
      F = TermFast{0, 0, 0, 0, 0, 0, 0u, 0u, terms[t].coef};
  }
  TermFast* d = nullptr;
  SQ_HIP(hipMalloc(&d, sizeof(TermFast) * f.size()));
  const hipError_t e = hipMemcpy(d, f.data(), sizeof(TermFast) * f.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    SQ_HIP(e);
  }
  return d;
}

static const TermFast* term_fast(const Program& p) {
  if (!p.d_term_fast) p.d_term_fast = upload_term_fast(p.terms);
  return p.d_term_fast;
}

struct UccCsr {
  int64_t nnz = 0;
  int64_t* rowp = nullptr;  // [dim + 1]
  uint2* ent = nullptr;     // [nnz]
  double* coef = nullptr;   // [terms]
  // The same entries grouped by gate (rows ascending inside a gate) as
  // (I, J, t | sign << 31, 0), for the gradient's per-gate dots; nullptr
  // above kGateMaxEntries.
  int64_t* gptr = nullptr;  // [gates + 1]
  uint4* gent = nullptr;    // [nnz]
  std::vector<int64_t> host_rowp;  // row offsets on the host while the pattern is built
};

void ucc_csr_free(UccCsr* c) {
  if (!c) return;
  if (c->rowp) (void)hipFree(c->rowp);
  if (c->ent) (void)hipFree(c->ent);
  if (c->coef) (void)hipFree(c->coef);
  if (c->gptr) (void)hipFree(c->gptr);
  if (c->gent) (void)hipFree(c->gent);
  delete c;
}

constexpr int64_t kGateMaxEntries = (int64_t)1 << 28;

// Regroup the row pattern by gate on the host (stable: rows stay ascending).
static void build_gate_major(const Space& s, const Program& p, const std::vector<int64_t>& rowp, UccCsr* c) {
  const int ng = (int)p.gates.size();
  if (ng == 0 || c->nnz > kGateMaxEntries) return;
  std::vector<uint2> ent((size_t)c->nnz);
  SQ_HIP(hipMemcpy(ent.data(), c->ent, sizeof(uint2) * ent.size(), hipMemcpyDeviceToHost));
  std::vector<int32_t> t2g(p.terms.size(), -1);
  for (int g = 0; g < ng; ++g)
    for (int t = p.gates[g].term_begin; t < p.gates[g].term_end; ++t) t2g[t] = g;
  std::vector<int64_t> gptr((size_t)ng + 1, 0);
  for (const uint2& x : ent) {
    const int g = t2g[x.y & 0x7fffffffu];
    if (g >= 0) ++gptr[(size_t)g + 1];
  }
  for (int g = 0; g < ng; ++g) gptr[(size_t)g + 1] += gptr[g];
  std::vector<uint4> gent((size_t)gptr[ng]);
  std::vector<int64_t> pos(gptr.begin(), gptr.end() - 1);
  for (int64_t I = 0; I < s.dim; ++I)
    for (int64_t k = rowp[I]; k < rowp[I + 1]; ++k) {
      const uint2 x = ent[(size_t)k];
      const int g = t2g[x.y & 0x7fffffffu];
      if (g >= 0) gent[(size_t)pos[g]++] = make_uint4((uint32_t)I, x.x, x.y, 0u);
    }
  hipError_t e = hipMalloc(&c->gptr, sizeof(int64_t) * gptr.size());
  if (e == hipSuccess) e = hipMalloc(&c->gent, sizeof(uint4) * std::max<size_t>(gent.size(), 1));
  if (e == hipSuccess) e = hipMemcpy(c->gptr, gptr.data(), sizeof(int64_t) * gptr.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !gent.empty())
    e = hipMemcpy(c->gent, gent.data(), sizeof(uint4) * gent.size(), hipMemcpyHostToDevice);
  SQ_HIP(e);
}

// Largest pattern kept (entries of 8 bytes); above it, and with SQ_UCC_CSR=0,
// T is applied by the term-scan kernels below.
constexpr int64_t kCsrMaxEntries = (int64_t)1 << 29;
// Largest pattern kept for a compiled operator (256 MB); larger operators
// (a Hamiltonian over a big space) keep the term scan.
constexpr int64_t kOpMaxEntries = (int64_t)1 << 25;

// Sparsity pattern of sum_t coef_t O_t (terms on the device, with their mask
// forms) on space s: per determinant I, the entries (J, t | parity << 31) of
// the terms whose adjoint maps I into the space (synchronous); nullptr when
// it is empty, above max_nnz entries, or the one-wave-per-row product would
// not fit one grid dimension.
static UccCsr* build_pattern(const Space& s, const std::vector<Term>& terms, const Term* d_terms,
                             const TermFast* fast, int64_t max_nnz, hipStream_t st) {
  const int nt = (int)terms.size();
  // k_tapply_csr runs one 64-lane wave per determinant: 64 * dim work-items
  // must stay below the 2^32 limit of one grid dimension
  if (s.dim == 0 || nt == 0 || 64 * s.dim >= ((int64_t)1 << 32)) return nullptr;
  const unsigned nb = (unsigned)((s.dim + kUB - 1) / kUB);
  int32_t* cnt = nullptr;
  SQ_HIP(hipMalloc(&cnt, sizeof(int32_t) * (size_t)s.dim));
  std::vector<int32_t> hc((size_t)s.dim);
  k_csr_count<<<nb, kUB, 0, st>>>(sv(s), d_terms, fast, nt, cnt);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(hc.data(), cnt, sizeof(int32_t) * hc.size(), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(cnt);
  SQ_HIP(e);
  std::vector<int64_t> rowp((size_t)s.dim + 1, 0);
  for (size_t i = 0; i < hc.size(); ++i) rowp[i + 1] = rowp[i] + hc[i];
  const int64_t nnz = rowp.back();
  if (nnz == 0 || nnz > max_nnz) return nullptr;
  UccCsr* c = new UccCsr;
  c->nnz = nnz;
  std::vector<double> coef((size_t)nt);
  for (int t = 0; t < nt; ++t) coef[t] = terms[t].coef;
  e = hipMalloc(&c->rowp, sizeof(int64_t) * rowp.size());
  if (e == hipSuccess) e = hipMalloc(&c->ent, sizeof(uint2) * (size_t)nnz);
  if (e == hipSuccess) e = hipMalloc(&c->coef, sizeof(double) * coef.size());
  if (e == hipSuccess) e = hipMemcpy(c->rowp, rowp.data(), sizeof(int64_t) * rowp.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->coef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    k_csr_fill<<<nb, kUB, 0, st>>>(sv(s), d_terms, fast, nt, c->rowp, c->ent);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    ucc_csr_free(c);
    SQ_HIP(e);
  }
  c->host_rowp = std::move(rowp);
  return c;
}

static bool csr_enabled() {
  const char* env = getenv("SQ_UCC_CSR");
  return !(env && env[0] == '0');
}

// The program's sparsity pattern of T on space s, built on first use
// (synchronous); nullptr where it is not used.
static const UccCsr* ucc_csr(const Space& s, const Program& p, hipStream_t st) {
  if (p.ucc_csr_tried) return p.ucc_csr;
  p.ucc_csr_tried = true;
  if (!csr_enabled()) return nullptr;
  UccCsr* c = build_pattern(s, p.terms, p.d_terms, term_fast(p), kCsrMaxEntries, st);
  if (!c) return nullptr;
  try {
    build_gate_major(s, p, c->host_rowp, c);
  } catch (...) {
    ucc_csr_free(c);
    throw;
  }
  c->host_rowp.clear();
  c->host_rowp.shrink_to_fit();
  p.ucc_csr = c;
  return c;
}

bool operator_apply_pattern(const Space& s, sq_operator& op, const double* in, double* out, hipStream_t st) {
  if (!op.csr_tried) {
    op.csr_tried = true;
    if (!csr_enabled()) return false;
    op.d_fast = upload_term_fast(op.terms);
    op.csr = build_pattern(s, op.terms, op.d_terms, op.d_fast, kOpMaxEntries, st);
    if (op.csr) {
      op.csr->host_rowp.clear();
      op.csr->host_rowp.shrink_to_fit();
    }
  }
  const UccCsr* c = op.csr;
  if (!c) return false;
  SQ_PROF(PC_TAPPLY, st, 16.0 * (double)s.dim + 8.0 * (double)c->nnz, 0.0,
          k_tapply_csr<<<(unsigned)((s.dim + kCsrRows - 1) / kCsrRows), 64 * kCsrRows, 0, st>>>(
              s.dim, c->rowp, c->ent, c->coef, nullptr, in, 1.0, out, nullptr));
  return true;
}

void operator_pattern_free(sq_operator& op) {
  if (op.csr) ucc_csr_free(op.csr);
  op.csr = nullptr;
  if (op.d_fast) (void)hipFree(op.d_fast);
  op.d_fast = nullptr;
}

// Term groups of the split T application: enough (determinant, group)
// threads to fill the GPU for small spaces, at most 64 groups.
static int term_groups(const Space& s) {
  const int64_t G = 262144 / std::max<int64_t>(s.dim, 1);
  return (int)std::max<int64_t>(1, std::min<int64_t>(64, G));
}

// Work of the UCC routines (doubles), carved in this order; A2, B2 and the
// node states NC, ND (kNodes vectors each) serve the gradient's
// multi-node series.
struct UccWork {
  double *A, *B, *C, *D, *w, *part, *dpart, *partial, *scal, *A2, *B2, *NC, *ND;
};

static int64_t dot_blocks(const Space& s) { return (s.dim + kUB - 1) / kUB; }

static int gate_groups(const Space& s, int ng) {
  const int64_t nb = dot_blocks(s);
  return (int)std::max<int64_t>(1, std::min<int64_t>(ng, std::max<int64_t>(1, 512 / nb)));
}

int64_t ucc_work_doubles(const Space& s, const Program& p) {
  const int64_t ng = (int64_t)std::max<size_t>(p.gates.size(), 1);
  return 4 * s.dim + (int64_t)p.terms.size() + 64 + (int64_t)term_groups(s) * s.dim + dot_blocks(s) * ng + 1024 +
         64 + 64 + (2 + 2 * kNodes) * s.dim;
}

static UccWork carve_ucc(const Space& s, const Program& p, double* work) {
  UccWork u;
  const int64_t ng = (int64_t)std::max<size_t>(p.gates.size(), 1);
  u.A = work;
  u.B = u.A + s.dim;
  u.C = u.B + s.dim;
  u.D = u.C + s.dim;
  u.w = u.D + s.dim;
  u.part = u.w + p.terms.size() + 64;
  u.dpart = u.part + (size_t)term_groups(s) * s.dim;
  u.partial = u.dpart + (size_t)dot_blocks(s) * ng;
  u.scal = u.partial + 1024;
  u.A2 = u.scal + 64;
  u.B2 = u.A2 + s.dim;
  u.NC = u.B2 + s.dim;
  u.ND = u.NC + (size_t)kNodes * s.dim;
  return u;
}

// nxt = f T cur (w: term weights, nullptr for unit weights); acc += nxt.
static void tapply(const Space& s, const Program& p, const double* w, const double* cur, double f, double* nxt,
                   double* acc, double* part, hipStream_t st) {
  if (const UccCsr* c = ucc_csr(s, p, st)) {
    SQ_PROF(PC_TAPPLY, st, 16.0 * (double)s.dim + 8.0 * (double)c->nnz, 0.0,
            k_tapply_csr<<<(unsigned)((s.dim + kCsrRows - 1) / kCsrRows), 64 * kCsrRows, 0, st>>>(
                s.dim, c->rowp, c->ent, c->coef, w, cur, f, nxt, acc));
    SQ_HIP(hipGetLastError());
    return;
  }
  const int G = term_groups(s);
  int rows = G;  // part rows
  if (G >= kTG) {  // small space: fold kTG groups per block in LDS
    rows = (G + kTG - 1) / kTG;
    const dim3 grid((unsigned)((s.dim + 63) / 64), rows);
    SQ_PROF(PC_TAPPLY, st, 24.0 * (double)s.dim, 0.0,
            k_tapply_part<<<grid, 64 * kTG, 0, st>>>(sv(s), p.d_terms, term_fast(p), w, (int)p.terms.size(), G,
                                                      cur, part));
  } else {
    const dim3 grid((unsigned)((s.dim + kUB - 1) / kUB), G);
    SQ_PROF(PC_TAPPLY, st, 24.0 * (double)s.dim, 0.0,
            k_tapply_group<<<grid, kUB, 0, st>>>(sv(s), p.d_terms, term_fast(p), w, (int)p.terms.size(), G, cur,
                                                 part));
  }
  k_tapply_sum<<<(unsigned)((s.dim + kUB - 1) / kUB), kUB, 0, st>>>(part, rows, s.dim, f, nxt, acc);
  SQ_HIP(hipGetLastError());
}

static double read_slot(const double* scal, int slot, hipStream_t st) {
  double h = 0.0;
  SQ_HIP(hipMemcpyAsync(&h, scal + slot, sizeof(double), hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  return h;
}

// Terms of a Taylor series of exp(t T) with |t| ||T||_2 <= r: the smallest K
// with r^K / K! < 2^-60 (at most 60).
static int taylor_terms(double r) {
  int K = 1;
  double term = r;
  while (term > std::ldexp(1.0, -60) && K < 60) {
    ++K;
    term *= r / K;
  }
  return K;
}

// Taylor plan for exp(t T) with ||T||_2 <= nu: s steps of exp(t T / s) with
// |t| nu / s <= 2 and taylor_terms(|t| nu / s) terms per step (a radius of 2
// costs one bit to cancellation: the largest term is 2).
struct TaylorPlan {
  int steps, K;
  double a;  // t / steps
};

static TaylorPlan taylor_plan(double nu, double t) {
  const double b = std::fabs(t) * nu;
  TaylorPlan P;
  P.steps = std::max(1, (int)std::ceil(b / 2.0));
  P.K = taylor_terms(b / P.steps);
  P.a = t / P.steps;
  return P;
}

// ||T||_2 for the weights w: T is real antisymmetric (normal), so its
// 2-norm is its spectral radius.  Power iteration (16 steps from a fixed
// pseudo-random vector, in A and B) estimates it from below as
// ||T x_15|| / ||x_15||; the returned value is 1.5 x the estimate + 0.05,
// capped by the triangle-inequality bound b = sum_t |w_t coef_t| (each raw
// string has at most one unit entry per column).  The iterates are taken of
// T / b (norm <= 1), so they need no renormalisation between steps and the
// 16 products run back to back; two dots and one read-back finish it.  If
// they underflow (||T|| below 2^-60 b), the steps are rerun with a
// renormalisation each.
static double t_norm(const Space& s, const Program& p, const std::vector<double>& th, const UccWork& u,
                     hipStream_t st) {
  double bound = 0.0;
  for (size_t g = 0; g < p.gates.size(); ++g)
    for (int k = p.gates[g].term_begin; k < p.gates[g].term_end; ++k) bound += std::fabs(th[g] * p.terms[k].coef);
  if (bound == 0.0 || s.dim == 0) return 0.0;
  const unsigned nb = (unsigned)((s.dim + kUB - 1) / kUB);
  double* x = u.A;
  double* y = u.B;
  k_fill_pattern<<<nb, kUB, 0, st>>>(x, s.dim);
  SQ_HIP(hipGetLastError());
  for (int it = 0; it < 15; ++it) {
    tapply(s, p, u.w, x, 1.0 / bound, y, nullptr, u.part, st);  // y = (T / b) x
    std::swap(x, y);
  }
  tapply(s, p, u.w, x, 1.0 / bound, y, nullptr, u.part, st);
  dot_dev(x, x, s.dim, u.scal, 0, 1.0, u.partial, st);
  dot_dev(y, y, s.dim, u.scal, 1, 1.0, u.partial, st);
  double d[2];
  SQ_HIP(hipMemcpyAsync(d, u.scal, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  double est;
  if (d[0] > 1e-200 && std::isfinite(d[1])) {
    est = bound * std::sqrt(std::max(0.0, d[1] / d[0]));
  } else {
    k_fill_pattern<<<nb, kUB, 0, st>>>(x, s.dim);
    SQ_HIP(hipGetLastError());
    for (int it = 0; it < 16; ++it) {
      dot_dev(x, x, s.dim, u.scal, 0, 1.0, u.partial, st);
      k_inv_sqrt_slot<<<1, 64, 0, st>>>(u.scal, 0, 1);
      k_scale<<<nb, kUB, 0, st>>>(x, s.dim, u.scal, 1);  // |x| = 1
      tapply(s, p, u.w, x, 1.0, y, nullptr, u.part, st);  // y = T x
      std::swap(x, y);
    }
    dot_dev(x, x, s.dim, u.scal, 0, 1.0, u.partial, st);
    est = std::sqrt(std::max(0.0, read_slot(u.scal, 0, st)));
  }
  return std::min(bound, 1.5 * est + 0.05);
}

// out = exp(t T) in over the plan for norm nu (A, B: Taylor terms).
static void expm_apply(const Space& s, const Program& p, const UccWork& u, double nu, double t, const double* in,
                       double* out, hipStream_t st) {
  const TaylorPlan P = taylor_plan(nu, t);
  if (in != out) copy_vec(in, out, s.dim, st);
  if (nu == 0.0) return;
  for (int step = 0; step < P.steps; ++step) {
    copy_vec(out, u.A, s.dim, st);
    double* cur = u.A;
    double* nxt = u.B;
    for (int k = 1; k <= P.K; ++k) {
      tapply(s, p, u.w, cur, P.a / k, nxt, out, u.part, st);
      std::swap(cur, nxt);
    }
  }
}

static std::vector<double> theta_host(const Program& p, const double* theta_dev, hipStream_t st) {
  const int ng = (int)p.gates.size();
  std::vector<double> th(ng);
  if (ng) SQ_HIP(hipMemcpyAsync(th.data(), theta_dev, sizeof(double) * ng, hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  return th;
}

static void term_weights(const Program& p, const double* theta_dev, double* w, hipStream_t st) {
  const int ng = (int)p.gates.size();
  if (ng) k_term_weights<<<(ng + 255) / 256, 256, 0, st>>>(p.d_gates, ng, theta_dev, w);
  SQ_HIP(hipGetLastError());
}

// out = exp(t T(theta)) in.  work: ucc_work_doubles.
void ucc_expm_impl(const Space& s, const Program& p, const double* theta_dev, double t, const double* in,
                   double* out, double* work, hipStream_t st) {
  const UccWork u = carve_ucc(s, p, work);
  const std::vector<double> th = theta_host(p, theta_dev, st);
  term_weights(p, theta_dev, u.w, st);
  const double nu = t_norm(s, p, th, u, st);
  expm_apply(s, p, u, nu, t, in, out, st);
}

// 8-point Gauss-Legendre nodes and weights on [0, 1].
static const double kGLx[8] = {0.019855071751231856, 0.10166676129318664, 0.2372337950418355,  0.4082826787521751,
                               0.5917173212478249,   0.7627662049581645,  0.8983332387068134,  0.9801449282487681};
static const double kGLw[8] = {0.05061426814518813, 0.11119051722668724, 0.15685332293894363, 0.18134189168918100,
                               0.18134189168918100, 0.15685332293894363, 0.11119051722668724, 0.05061426814518813};

// Analytic dE/dtheta for E = <psi|H|psi>, psi = exp(T) psi0, lam = H psi:
//   d psi / d theta_i = int_0^1 e^{sT} T_i e^{(1-s)T} psi0 ds, and with
//   e^{sT}^T = e^{-sT} (T real antisymmetric)
//   grad_i = 2 int_0^1 <e^{-sT} lam | T_i | e^{-sT} psi> ds,
// by composite 8-point Gauss-Legendre over m subintervals of length h with
// h * 2 nu <= 4 (the integrand is entire with frequencies bounded by
// 2 ||T||_2 = omega; the 8-point rule's error on one subinterval is below
// h (8!)^4 / (17 (16!)^3) (h omega)^16 max|f|, so over [0, 1] it is below
// 1.7e-23 * 4^16 max|f| < 1e-13 max|f|).
// Without the gate-major pattern, psi and lam are propagated from node to
// node (C, D) and every node adds all gates' dots in one pass.
//
// With the gate-major pattern, each subinterval [a, a + h] runs one Taylor
// series of exp(-h T) on psi and lam together (k_tapply_csr2: the pattern read
// once for both), whose terms z_k = (-h)^k / k! T^k C also give every node:
// e^{-h x_q T} C = sum_k x_q^k z_k, accumulated into the node states as the
// terms are made (and the end of the subinterval, x = 1, starts the next);
// h nu <= 2 bounds the series' radius.  The 8 nodes' gate dots then run in
// one launch, split over blocks per gate.
static void nodes_apply(const Space& s, const UccCsr* c, const UccWork& u, double nu, double h, hipStream_t st) {
  const int K = taylor_terms(h * nu);
  const unsigned nb = (unsigned)((s.dim + kUB - 1) / kUB);
  k_nodes_init<<<nb, kUB, 0, st>>>(s.dim, u.C, u.D, u.NC, u.ND);
  SQ_HIP(hipGetLastError());
  double x[kNodes], pw[kNodes];
  for (int q = 0; q < 8; ++q) x[q] = kGLx[q];
  x[8] = 1.0;
  for (int q = 0; q < kNodes; ++q) pw[q] = 1.0;
  const double* cur0 = u.C;
  const double* cur1 = u.D;
  double* b0[2] = {u.A, u.B};
  double* b1[2] = {u.A2, u.B2};
  const unsigned rows = (unsigned)((s.dim + kCsrRows - 1) / kCsrRows);
  for (int k = 1; k <= K; ++k) {
    NodeW r;
    for (int q = 0; q < kNodes; ++q) r.r[q] = pw[q] *= x[q];
    double* n0 = b0[k & 1];
    double* n1 = b1[k & 1];
    SQ_PROF(PC_TAPPLY, st, 32.0 * (double)s.dim + 8.0 * (double)c->nnz + 16.0 * kNodes * (double)s.dim, 0.0,
            k_tapply_csr2<<<rows, 64 * kCsrRows, 0, st>>>(s.dim, c->rowp, c->ent, c->coef, u.w, cur0, cur1, -h / k,
                                                          n0, n1, u.NC, u.ND, r));
    SQ_HIP(hipGetLastError());
    cur0 = n0;
    cur1 = n1;
  }
}

static void ucc_grad_impl(const Space& s, const Program& p, const UccWork& u, double nu, const double* psi,
                          const double* lam, double* grad_dev, hipStream_t st) {
  const int ng = (int)p.gates.size();
  const int m = std::max(1, (int)std::ceil(0.5 * nu));
  copy_vec(psi, u.C, s.dim, st);
  copy_vec(lam, u.D, s.dim, st);
  k_zero<<<(unsigned)((ng + kUB - 1) / kUB), kUB, 0, st>>>(grad_dev, ng);
  SQ_HIP(hipGetLastError());
  const int nb = (int)dot_blocks(s);
  const dim3 dgrid(nb, gate_groups(s, ng));
  const UccCsr* csr = ucc_csr(s, p, st);
  static const int nodes_env = [] {
    const char* e = getenv("SQ_UCC_NODES");
    return (e && e[0] == '0') ? 0 : 1;
  }();
  if (csr && csr->gent && nodes_env) {
    const double h = 1.0 / m;
    const int S = (int)std::max<int64_t>(1, std::min<int64_t>(dot_blocks(s), std::max(1, 1024 / ng)));
    NodeW wq{};
    for (int q = 0; q < 8; ++q) wq.r[q] = 2.0 * kGLw[q] / m;
    for (int j = 0; j < m; ++j) {
      nodes_apply(s, csr, u, nu, h, st);
      k_gate_dots_nodes<<<dim3((unsigned)ng, (unsigned)S), kUB, 0, st>>>(csr->gptr, csr->gent, csr->coef, u.NC,
                                                                        u.ND, s.dim, wq, ng, u.dpart);
      k_gate_accum<<<(unsigned)((ng + kUB - 1) / kUB), kUB, 0, st>>>(u.dpart, S, ng, 1.0, grad_dev);
      SQ_HIP(hipGetLastError());
      if (j + 1 < m) {
        copy_vec(u.NC + (size_t)8 * s.dim, u.C, s.dim, st);
        copy_vec(u.ND + (size_t)8 * s.dim, u.D, s.dim, st);
      }
    }
    return;
  }
  double prev = 0.0;
  for (int j = 0; j < m; ++j) {
    for (int q = 0; q < 8; ++q) {
      const double node = (j + kGLx[q]) / m;
      const double ds = node - prev;
      expm_apply(s, p, u, nu, -ds, u.C, u.C, st);
      expm_apply(s, p, u, nu, -ds, u.D, u.D, st);
      prev = node;
      if (csr && csr->gent) {
        k_gate_dots_csr<<<(unsigned)ng, kUB, 0, st>>>(csr->gptr, csr->gent, csr->coef, u.C, u.D,
                                                     2.0 * kGLw[q] / m, grad_dev);
      } else {
        k_gate_dots<<<dgrid, kUB, 0, st>>>(sv(s), p.d_terms, term_fast(p), p.d_gates, ng, u.C, u.D, u.dpart);
        k_gate_accum<<<(unsigned)((ng + kUB - 1) / kUB), kUB, 0, st>>>(u.dpart, nb, ng, 2.0 * kGLw[q] / m,
                                                                       grad_dev);
      }
      SQ_HIP(hipGetLastError());
    }
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
    SQ_CHECK(pr->p.space == &sp->s, SQ_ERR_ARG, "program built for another space");
    ucc_expm_impl(sp->s, pr->p, theta, t, in, out, work, as_stream(stream));
  });
}

int sq_ucc_energy_grad(const sq_space* sp, const sq_program* pr, const sq_hamiltonian* H, const double* theta,
                       const double* psi0, double* energy, double* grad, void* work_v, void* stream) {
  double* work = static_cast<double*>(work_v);
  return guarded([&] {
    SQ_CHECK(sp && pr && H && psi0 && energy && work && (theta || pr->p.gates.empty()), SQ_ERR_ARG, "NULL argument");
    SQ_CHECK(pr->p.space == &sp->s && H->h.space == &sp->s, SQ_ERR_ARG,
             "program / Hamiltonian built for another space");
    const Space& s = sp->s;
    const Program& p = pr->p;
    const int ng = (int)p.gates.size();
    hipStream_t st = as_stream(stream);
    double* psi = work;
    double* lam = psi + s.dim;
    double* scal = lam + s.dim;            // 64
    double* gdev = scal + 64;              // ng
    double* rest = gdev + std::max(ng, 1) + 64;
    const UccWork u = carve_ucc(s, p, rest);
    const std::vector<double> th = theta_host(p, theta, st);
    term_weights(p, theta, u.w, st);
    const double nu = t_norm(s, p, th, u, st);
    expm_apply(s, p, u, nu, 1.0, psi0, psi, st);
    sigma_energy_dev(s, H->h, psi, lam, scal, rest, st);  // reuses the UCC work area
    if (grad && ng) {
      term_weights(p, theta, u.w, st);
      ucc_grad_impl(s, p, u, nu, psi, lam, gdev, st);
      SQ_HIP(hipMemcpyAsync(grad, gdev, sizeof(double) * ng, hipMemcpyDeviceToHost, st));
    }
    SQ_HIP(hipMemcpyAsync(energy, scal, sizeof(double), hipMemcpyDeviceToHost, st));
    SQ_HIP(hipStreamSynchronize(st));
  });
}

int64_t sq_ucc_work_bytes(const sq_space* sp, const sq_program* pr) {
  if (!sp || !pr) return -1;
  const Space& s = sp->s;
  const int64_t ng = (int64_t)pr->p.gates.size() + 1;
  const int64_t body = std::max<int64_t>(ucc_work_doubles(s, pr->p), sigma_work_bytes(s) / 8);
  return 8 * (2 * s.dim + 64 + ng + 64 + body);
}

}  // extern "C"
