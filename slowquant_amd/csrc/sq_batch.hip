// Batched energy + adjoint gradient for many independent UPS instances of
// uneven active spaces in ONE persistent launch (the reference's batching
// primitive is the _SA family, operator_state_algebra.py:282-410,633-781,
// which needs equal dimensions; here every instance has its own space,
// Hamiltonian, gate list and theta).
//
// One workgroup owns a list of instances (longest-processing-time order,
// assigned on the host, no atomics).  Per instance, inside the workgroup:
//   psi = HF;  forward rotations;  D build;  G = A.D with FP64 MFMA;
//   sigma gather + energy;  reverse sweep with fused <lambda|T_i psi> dots.
// State vectors and D/G live in a per-workgroup slot of global scratch
// (L2-resident for the small spaces), so instances never interact.
// The launch can be captured in a hipGraph and replayed.

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <string>

#include "sq_internal.h"

namespace sq {

constexpr int kBT = 512;    // threads per batch workgroup (8 waves)
constexpr int kBatchMT = 5;  // row tiles of the sigma GEMM (kpad <= 80: up to 12 orbitals)

// k -> whether packed pair k = p(p+1)/2 + q is diagonal (p == q), k < 91
__constant__ uint8_t c_bdiag[91];

struct BatchSpace {
  int n, nsa, nsb, npair, K4, kpad;
  const uint32_t* astr;
  const uint32_t* bstr;
  const int32_t* arank;
  const int32_t* brank;
  const int32_t* aexc;  // packed, k-major
  const int32_t* bexc;
};

struct BatchInstance {
  int space;        // index into spaces
  int gate0, ngate; // rotation records [gate0, gate0 + ngate)
  int theta0;       // offset into theta array (one theta per layout gate)
  int nparam;
  int A_off;        // offset into A storage (K4 * kpad doubles)
  double e_core;
  int out0;         // offset of the gradient in the output array
};

// One rotation half: string G = sgn * Ga * Gb with distinct indices; the
// layout gate's theta index and angle scale.
// A rotation string touches exactly the (alpha, beta) pairs where both spin
// parts apply: precomputed lists of (string, image | parity << 31) per spin
// (shared between rotations with the same spin part on the same space); a
// spin part without operators acts on every string (list offset -1).
struct BatchRot {
  double sgn, scale;
  int theta_idx;    // index relative to the instance's theta0
  int grad_idx;     // same, for the gradient accumulation
  int aoff, na;     // alpha list [aoff, aoff + na), or aoff = -1: every alpha string
  int boff, nb;     // beta list, likewise
};

struct BatchDev {
  const BatchSpace* spaces;
  const BatchInstance* inst;
  const BatchRot* rots;
  const int2* lists;       // (string, image | parity << 31) entries of the spin-part lists
  const double* theta;
  const double* A;
  const int32_t* wg_list;  // concatenated instance lists per workgroup
  const int32_t* wg_off;   // [nwg + 1]
  double* scratch;         // per-workgroup slots
  int64_t slot_doubles;
  int64_t dmax;
  double* energy;          // [ninst]
  double* grad;            // concatenated gradients
  const double* const* psi0;  // [ninst] start states (engine convention), nullptr: HF determinant
};

// Visit every (alpha row, beta column) of a space with 32-bit index math:
// a block-wide 2-D walk, rows_per_step rows of min(nsb, blockDim) columns.
template <class F>
__device__ __forceinline__ void for_each_det(int nsa, int nsb, F&& f) {
  const int cols = nsb < (int)blockDim.x ? nsb : (int)blockDim.x;
  const int rps = (int)blockDim.x / cols;
  const int r0 = (int)threadIdx.x / cols, c0 = (int)threadIdx.x - r0 * cols;
  if (r0 >= rps) return;
  for (int ia = r0; ia < nsa; ia += rps)
    for (int ib = c0; ib < nsb; ib += cols) f(ia, ib, ia * nsb + ib);
}

__device__ __forceinline__ int2 list_entry(const int2* lists, int off, int i) {
  return off < 0 ? int2{i, i} : lists[off + i];
}

__device__ double block_reduce(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
  return t;  // every thread gets the sum
}

// One rotation half of the forward (REV = false) or reverse sweep over its
// spin-part lists, U pairs of determinants per thread in flight: the loads of
// all U pairs are issued before any of them is updated (the pairs of one
// rotation are disjoint).  Returns this thread's part of
// scale * <lam|T psi> (reverse sweep).
template <int U, bool REV>
__device__ __forceinline__ double rotate_lists(const int2* __restrict__ lists, const BatchRot& R, int nsb, double c,
                                               double s, bool apply, double* psi, double* lam) {
  const int na = R.na, nb = R.nb;
  const int cols = nb < (int)blockDim.x ? (nb > 0 ? nb : 1) : (int)blockDim.x;
  const int rps = (int)blockDim.x / cols;
  const int r0 = (int)threadIdx.x / cols, c0 = (int)threadIdx.x - r0 * cols;
  double dot = 0.0;
  if (r0 >= rps) return 0.0;
  for (int bi = c0; bi < nb; bi += cols) {
    const int2 Bq = list_entry(lists, R.boff, bi);
    for (int a0 = r0; a0 < na; a0 += rps * U) {
      int x[U], J[U];
      double sg[U], pd[U], pj[U], ld[U], lj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int ai = a0 + u * rps;
        x[u] = -1;
        if (ai < na) {
          const int2 A = list_entry(lists, R.aoff, ai);
          x[u] = A.x * nsb + Bq.x;
          J[u] = (A.y & 0x7fffffff) * nsb + (Bq.y & 0x7fffffff);
          sg[u] = (((uint32_t)(A.y ^ Bq.y) >> 31) ? -1.0 : 1.0) * R.sgn;
          pd[u] = psi[x[u]];
          pj[u] = psi[J[u]];
          if (REV) {
            ld[u] = lam[x[u]];
            lj[u] = lam[J[u]];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (x[u] < 0) continue;
        if (REV) dot += R.scale * sg[u] * (lj[u] * pd[u] - ld[u] * pj[u]);
        if (apply) {
          psi[x[u]] = c * pd[u] - sg[u] * s * pj[u];
          psi[J[u]] = sg[u] * s * pd[u] + c * pj[u];
          if (REV) {
            lam[x[u]] = c * ld[u] - sg[u] * s * lj[u];
            lam[J[u]] = sg[u] * s * ld[u] + c * lj[u];
          }
        }
      }
    }
  }
  return dot;
}

// U: rotation pairs per thread in flight; CT: 16-column tiles per wave in the
// sigma GEMM (sharing each LDS read of A).  Both are picked by the autotuner.
template <int U, int CT>
__global__ __launch_bounds__(kBT) void k_batch(BatchDev B) {
  __shared__ double red[kBT / 64];
  extern __shared__ double sA[];
  double* slot = B.scratch + (size_t)blockIdx.x * B.slot_doubles;
  double* psi = slot;
  double* lam = psi + B.dmax;
  for (int li = B.wg_off[blockIdx.x]; li < B.wg_off[blockIdx.x + 1]; ++li) {
    const BatchInstance I = B.inst[B.wg_list[li]];
    const BatchSpace S = B.spaces[I.space];
    const int nsb = S.nsb;
    const int dim = S.nsa * nsb;  // < 2^30, checked at creation
    double* G = lam + B.dmax;
    const double* th = B.theta + I.theta0;
    const double* start = B.psi0[B.wg_list[li]];
    if (start) {
      for (int64_t x = threadIdx.x; x < dim; x += blockDim.x) psi[x] = start[x];
    } else {
      // psi = HF determinant (alpha string 0, beta string 0) in the engine's
      // sign convention: (-1)^{sum_{p in alpha} #{q in beta : q < p}}.
      for (int64_t x = threadIdx.x; x < dim; x += blockDim.x) psi[x] = 0.0;
      __syncthreads();
      if (threadIdx.x == 0) {
        uint32_t a = S.astr[0], b = S.bstr[0];
        int par = 0;
        for (uint32_t aa = a; aa; aa &= aa - 1) par += __popc(b & ((1u << (__ffs(aa) - 1)) - 1u));
        psi[0] = (par & 1) ? -1.0 : 1.0;
      }
    }
    __syncthreads();
    // forward sweep
    for (int r = 0; r < I.ngate; ++r) {
      const BatchRot& R = B.rots[I.gate0 + r];
      const double t0 = th[R.theta_idx];
      if (fabs(t0) >= 1e-28) {
        double s, c;
        sincos(R.scale * t0, &s, &c);
        (void)rotate_lists<U, false>(B.lists, R, nsb, c, s, true, psi, nullptr);
      }
      __syncthreads();
    }
    // stage A (K4 x kpad, transposed coefficient matrix) in LDS
    const double* Ag = B.A + I.A_off;
    for (int i = threadIdx.x; i < S.K4 * S.kpad; i += blockDim.x) sA[i] = Ag[i];
    __syncthreads();
    // G = A . D with v_mfma_f64_16x16x4f64, the B operand D[k][c] = (E+_k psi)[c]
    // (row npair: psi) generated in registers: each wave walks 16-column tiles
    // and keeps all kpad/16 row tiles in flight, so D is never stored.
    {
      typedef double double4_t __attribute__((ext_vector_type(4)));
      const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, kq = lane >> 4, li = lane & 15;
      const int nwaves = blockDim.x >> 6;
      const int MT = S.kpad / 16;
      for (int c0 = wave * 16 * CT; c0 < dim; c0 += nwaves * 16 * CT) {
        int ccol[CT], ia[CT], ib[CT];
        bool ok[CT];
        double pc[CT];
#pragma unroll
        for (int u = 0; u < CT; ++u) {
          ccol[u] = c0 + u * 16 + li;
          ok[u] = ccol[u] < dim;
          ia[u] = ok[u] ? ccol[u] / nsb : 0;
          ib[u] = ok[u] ? ccol[u] - ia[u] * nsb : 0;
          pc[u] = ok[u] ? psi[ccol[u]] : 0.0;
        }
        auto dval = [&](int u, int k) -> double {
          if (!ok[u] || k > S.npair) return 0.0;
          if (k == S.npair) return pc[u];
          const int32_t ea = S.aexc[k * S.nsa + ia[u]];
          const int32_t eb = S.bexc[k * nsb + ib[u]];
          if (c_bdiag[k]) return (double)(ea + eb) * pc[u];
          double v = 0.0;
          if (ea) v += (ea > 0 ? 1.0 : -1.0) * psi[(abs(ea) - 1) * nsb + ib[u]];
          if (eb) v += (eb > 0 ? 1.0 : -1.0) * psi[ia[u] * nsb + abs(eb) - 1];
          return v;
        };
        double4_t acc[CT][kBatchMT];
#pragma unroll
        for (int u = 0; u < CT; ++u)
#pragma unroll
          for (int t = 0; t < kBatchMT; ++t) acc[u][t] = double4_t{0.0, 0.0, 0.0, 0.0};
        double bnext[CT];
#pragma unroll
        for (int u = 0; u < CT; ++u) bnext[u] = dval(u, kq);
        for (int k4 = 0; k4 < S.K4; k4 += 4) {
          double bv[CT];
#pragma unroll
          for (int u = 0; u < CT; ++u) {
            bv[u] = bnext[u];
            if (k4 + 4 < S.K4) bnext[u] = dval(u, k4 + 4 + kq);
          }
          const double* arow = sA + (k4 + kq) * S.kpad + li;
#pragma unroll
          for (int t = 0; t < kBatchMT; ++t) {
            if (t < MT) {
              const double a = arow[t * 16];
#pragma unroll
              for (int u = 0; u < CT; ++u) acc[u][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv[u], acc[u][t], 0, 0, 0);
            }
          }
        }
#pragma unroll
        for (int u = 0; u < CT; ++u) {
          if (!ok[u]) continue;
#pragma unroll
          for (int t = 0; t < kBatchMT; ++t) {
            if (t >= MT) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int m = t * 16 + kq + 4 * r;
              if (m < S.npair) G[(size_t)m * dim + ccol[u]] = acc[u][t][r];
            }
          }
        }
      }
    }
    __syncthreads();
    // sigma gather into lam, energy
    double e = 0.0;
    for_each_det(S.nsa, nsb, [&](int ia, int ib, int x) {
      double v = I.e_core * psi[x];
      int k = 0;
      for (int p = 0; p < S.n; ++p) {
        for (int q = 0; q <= p; ++q, ++k) {
          const double* Gk = G + (size_t)k * dim;
          const int32_t ea = S.aexc[(size_t)k * S.nsa + ia];
          const int32_t eb = S.bexc[(size_t)k * nsb + ib];
          if (p == q) {
            v += (double)(ea + eb) * Gk[x];
          } else {
            if (ea) v += (ea > 0 ? 1.0 : -1.0) * Gk[(size_t)(abs(ea) - 1) * nsb + ib];
            if (eb) v += (eb > 0 ? 1.0 : -1.0) * Gk[(size_t)ia * nsb + abs(eb) - 1];
          }
        }
      }
      lam[x] = v;
      e += v * psi[x];
    });
    const double energy = block_reduce(e, red);
    if (threadIdx.x == 0) B.energy[B.wg_list[li]] = energy;
    // reverse sweep with gradient dots
    for (int p = threadIdx.x; p < I.nparam; p += blockDim.x) B.grad[I.out0 + p] = 0.0;
    __syncthreads();
    for (int r = I.ngate - 1; r >= 0; --r) {
      const BatchRot& R = B.rots[I.gate0 + r];
      const double t0 = th[R.theta_idx];
      const bool skip = fabs(t0) < 1e-28;
      double s, c;
      sincos(-R.scale * t0, &s, &c);
      const double dot = rotate_lists<U, true>(B.lists, R, nsb, c, s, !skip, psi, lam);
      const double tot = block_reduce(dot, red);
      if (threadIdx.x == 0) B.grad[I.out0 + R.grad_idx] += 2.0 * tot;
      __syncthreads();
    }
  }
}

// Large instances of one (space, gate layout) run together through the
// single-instance kernels with an instance grid dimension (tile stages, row
// sigma, gradient): one persistent workgroup per instance would leave them
// bound by one CU each.  Per-instance buffers are contiguous at fixed
// strides.
struct BatchGroup {
  const Space* space = nullptr;
  const Program* prog = nullptr;
  int n = 0, ng = 0;
  int64_t dim = 0, A_stride = 0, Ma_stride = 0, mats_stride = 0, tpart_stride = 0, spart_stride = 0;
  double hf = 1.0;  // the HF determinant's amplitude in the engine's sign convention
  int plan_gen = 0;  // the program's tile-plan generation the buffers are sized for
  int32_t *d_inst = nullptr, *d_theta0 = nullptr, *d_out0 = nullptr;  // per member
  double *theta = nullptr, *A = nullptr, *Ma = nullptr, *psi = nullptr, *lam = nullptr, *mats = nullptr,
         *tpart = nullptr, *spart = nullptr, *grad = nullptr;
};

// theta_g[m][i] = theta_all[theta0[m] + i] (member m = blockIdx.y)
__global__ void k_group_theta(const double* __restrict__ theta_all, const int32_t* __restrict__ theta0, int ng,
                              double* __restrict__ theta_g) {
  const int m = blockIdx.y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += gridDim.x * blockDim.x)
    theta_g[(size_t)m * ng + i] = theta_all[theta0[m] + i];
}

// psi_g[m] = the member's start state, or the HF determinant (amplitude hf at 0)
__global__ void k_group_init(const double* const* __restrict__ psi0, const int32_t* __restrict__ inst, int64_t dim,
                             double hf, double* __restrict__ psi) {
  const int m = blockIdx.y;
  const double* src = psi0[inst[m]];
  double* dst = psi + (size_t)m * dim;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src ? src[i] : (i == 0 ? hf : 0.0);
}

// energy[inst[m]] = sum of member m's sigma energy partials (fixed order);
// grad[out0[m] + g] = grad_g[m][g]
__global__ void k_group_out(const double* __restrict__ spart, int64_t nblk, const double* __restrict__ grad_g, int ng,
                            const int32_t* __restrict__ inst, const int32_t* __restrict__ out0,
                            double* __restrict__ energy, double* __restrict__ grad) {
  __shared__ double red[4];
  const int m = blockIdx.x;
  const double* p = spart + (size_t)m * nblk;
  double v = 0.0;
  for (int64_t i = threadIdx.x; i < nblk; i += blockDim.x) v += p[i];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) energy[inst[m]] = red[0] + red[1] + red[2] + red[3];
  for (int g = threadIdx.x; g < ng; g += blockDim.x) grad[out0[m] + g] = grad_g[(size_t)m * ng + g];
}

struct Batch {
  std::vector<BatchInstance> inst;
  int nwg = 0;
  int64_t dmax = 0, slot = 0, ngrad = 0;
  double bytes = 0.0, flops = 0.0;  // algorithmic totals of one launch
  size_t lds = 0;
  std::vector<void*> allocs;
  BatchDev dev{};
  double* d_theta = nullptr;
  const double** d_psi0 = nullptr;  // start-state pointer table [inst]
  std::vector<BatchGroup> groups;   // instances run through the tile and row-sigma kernels
  std::vector<std::pair<int, int>> group_members;  // (group, instance), in instance order
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<int32_t> grad_off;
  hipStream_t side = nullptr;  // the persistent launch beside the groups (default priority)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int rot_pairs = 2;    // rotation pairs per thread in flight (1, 2, 4)
  int gemm_tiles = 1;   // sigma GEMM column tiles per wave (1, 2)
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~Batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
    for (void* p : allocs) (void)hipFree(p);
  }
};

template <class T>
static T* bupload(Batch& B, const std::vector<T>& v) {
  T* d = nullptr;
  SQ_HIP(hipMalloc(&d, sizeof(T) * std::max<size_t>(v.size(), 1)));
  if (!v.empty()) SQ_HIP(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  B.allocs.push_back(d);
  return d;
}


// Instances this large go to groups (the persistent launch keeps the rest):
// at (10e,10o) one workgroup per instance runs a 3-layer evaluation in
// 14.5 ms against 0.25 ms through the single-instance kernels.
constexpr int64_t kGroupMinDim = 16384;

// Bytes of a program's gate layout (kinds, pairs, scales and generator
// terms): equal keys on one space give identical tile plans.
static std::string layout_key(const Program& p) {
  std::string k;
  auto put = [&](const void* x, size_t n) { k.append(reinterpret_cast<const char*>(x), n); };
  for (size_t g = 0; g < p.gates.size(); ++g) {
    const GateDev& gd = p.gates[g];
    put(&gd.kind, sizeof gd.kind);
    put(&gd.scale, sizeof gd.scale);
    const int32_t pp = g < p.pair_p.size() ? p.pair_p[g] : -1;
    put(&pp, sizeof pp);
    const int32_t nt = gd.term_end - gd.term_begin;
    put(&nt, sizeof nt);
    for (int t = gd.term_begin; t < gd.term_end; ++t) {
      const Term& T = p.terms[t];
      put(&T.coef, sizeof T.coef);
      for (const SpinOps* o : {&T.a, &T.b}) {
        put(&o->n, sizeof o->n);
        put(o->orb, (size_t)o->n);
        put(o->cre, (size_t)o->n);
      }
    }
  }
  return k;
}

template <class T>
static T* balloc(Batch& B, size_t n) {
  T* d = nullptr;
  SQ_HIP(hipMalloc(&d, sizeof(T) * std::max<size_t>(n, 1)));
  B.allocs.push_back(d);
  return d;
}

// Buffers of every group: members' theta offsets, output offsets, their
// Hamiltonians copied into contiguous A and Ma arrays, and the work vectors.
static void build_groups(Batch& B, const sq_hamiltonian* const* hams, const sq_space* const* spaces) {
  (void)spaces;
  std::vector<std::vector<int>> members(B.groups.size());
  for (const auto& gm : B.group_members) members[gm.first].push_back(gm.second);
  for (size_t gi = 0; gi < B.groups.size(); ++gi) {
    BatchGroup& G = B.groups[gi];
    const Space& s = *G.space;
    const std::vector<int>& mem = members[gi];
    G.n = (int)mem.size();
    SQ_CHECK(G.n <= 65535, SQ_ERR_ARG, "too many instances in one batched group");
    G.dim = s.dim;
    const Hamiltonian& H0 = hams[mem[0]]->h;
    const int K4 = (s.npair + 1 + 3) / 4 * 4;
    G.A_stride = (int64_t)K4 * H0.kpad;
    G.Ma_stride = (int64_t)s.nsa * s.nsa;
    G.plan_gen = G.prog->plan_gen;
    G.mats_stride = tiles_mats_doubles(s, *G.prog);
    G.tpart_stride = tiles_partial_doubles(s, *G.prog);
    G.spart_stride = row_sigma_group_partials(s);
    SQ_CHECK(G.mats_stride > 0 && G.tpart_stride > 0 && G.spart_stride > 0, SQ_ERR_ARG, "group path unavailable");
    {
      const uint32_t a = s.h_astr[0], b = s.h_bstr[0];
      int par = 0;
      for (uint32_t aa = a; aa; aa &= aa - 1) par += __builtin_popcount(b & ((1u << __builtin_ctz(aa)) - 1u));
      G.hf = (par & 1) ? -1.0 : 1.0;
    }
    std::vector<int32_t> inst(mem.begin(), mem.end()), th0, out0;
    for (int i : mem) {
      th0.push_back(B.inst[i].theta0);
      out0.push_back(B.inst[i].out0);
    }
    G.d_inst = bupload(B, inst);
    G.d_theta0 = bupload(B, th0);
    G.d_out0 = bupload(B, out0);
    const size_t n = (size_t)G.n;
    G.theta = balloc<double>(B, n * G.ng);
    G.A = balloc<double>(B, n * G.A_stride);
    G.Ma = balloc<double>(B, n * G.Ma_stride);
    G.psi = balloc<double>(B, n * G.dim);
    G.lam = balloc<double>(B, n * G.dim);
    G.mats = balloc<double>(B, n * G.mats_stride);
    G.tpart = balloc<double>(B, n * G.tpart_stride);
    G.spart = balloc<double>(B, n * G.spart_stride);
    G.grad = balloc<double>(B, n * G.ng);
    for (size_t m = 0; m < n; ++m) {
      const Hamiltonian& H = hams[mem[m]]->h;
      SQ_CHECK(H.kpad == H0.kpad && H.d_Ma, SQ_ERR_ARG, "group Hamiltonians differ in shape");
      SQ_HIP(hipMemcpy(G.A + m * G.A_stride, H.d_A, sizeof(double) * G.A_stride, hipMemcpyDeviceToDevice));
      SQ_HIP(hipMemcpy(G.Ma + m * G.Ma_stride, H.d_Ma, sizeof(double) * G.Ma_stride, hipMemcpyDeviceToDevice));
    }
  }
}

// A grouped program whose tile shape changed after the batch was created has
// a new tile plan: the group's buffers (and a captured graph) are sized for
// the old one.
// The group buffers are sized for the tile plan of each grouped program and
// the row-sigma block count of each grouped space when the batch was made.
static void check_group(const BatchGroup& G) {
  SQ_CHECK(G.prog->plan_gen == G.plan_gen, SQ_ERR_ARG,
           "a batched program's tile shape changed after the batch was created; create the batch again");
  SQ_CHECK(row_sigma_group_partials(*G.space) == G.spart_stride, SQ_ERR_ARG,
           "a batched space's sigma shape changed after the batch was created; create the batch again");
}

static void check_groups(const Batch& B) {
  for (const BatchGroup& G : B.groups) check_group(G);
}

// One group's evaluation on st: thetas and start states in, forward sweep,
// sigma and energies, reverse sweep with gradients, results out.
static void group_launch(const Batch& B, const BatchGroup& G, hipStream_t st) {
  const Space& s = *G.space;
  check_group(G);
  Multi M;
  M.n = G.n;
  M.theta = G.ng;
  M.mats = G.mats_stride;
  M.vec = G.dim;
  M.part = G.tpart_stride;
  M.grad = G.ng;
  k_group_theta<<<dim3((unsigned)((G.ng + 255) / 256), (unsigned)G.n), 256, 0, st>>>(B.dev.theta, G.d_theta0, G.ng,
                                                                                     G.theta);
  const unsigned ib = (unsigned)std::min<int64_t>((G.dim + 255) / 256, 64);
  k_group_init<<<dim3(ib, (unsigned)G.n), 256, 0, st>>>(B.dev.psi0, G.d_inst, G.dim, G.hf, G.psi);
  SQ_HIP(hipGetLastError());
  tiles_forward_group(s, *G.prog, G.theta, G.psi, G.mats, M, st);
  row_sigma_group(s, G.A, G.A_stride, G.Ma, G.Ma_stride, G.psi, G.lam, G.spart, G.n, st);
  tiles_reverse_group(s, *G.prog, G.theta, G.psi, G.lam, G.mats, G.tpart, G.grad, M, st);
  k_group_out<<<(unsigned)G.n, 256, 0, st>>>(G.spart, G.spart_stride, G.grad, G.ng, G.d_inst, G.d_out0,
                                             B.dev.energy, B.dev.grad);
  SQ_HIP(hipGetLastError());
}

}  // namespace sq

using namespace sq;

struct sq_batch {
  Batch b;
};

extern "C" {

int sq_batch_create(int n_inst, const sq_space* const* spaces, const sq_program* const* programs,
                    const sq_hamiltonian* const* hams, int n_workgroups, sq_batch** out) {
  return guarded([&] {
    SQ_CHECK(n_inst > 0 && spaces && programs && hams && out, SQ_ERR_ARG, "invalid arguments");
    {
      uint8_t diag[91] = {0};
      for (int p = 0; p < 13; ++p) diag[p * (p + 1) / 2 + p] = 1;
      SQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_bdiag), diag, sizeof(diag)));
    }
    auto* bo = new sq_batch();
    Batch& B = bo->b;
    try {
      std::map<const sq_space*, int> sp_index;
      std::vector<int2> lists;
      std::map<std::pair<const void*, std::vector<int>>, std::pair<int, int>> list_cache;
      // (string, image | parity << 31) of every string a spin part applies to
      auto spin_list = [&](const std::vector<uint32_t>& strs, const SpinOps& o, int& off, int& len) {
        spin_part_list(strs, o, lists, list_cache, off, len);
      };
      std::vector<BatchSpace> bsp;
      std::vector<BatchRot> rots;
      std::vector<double> Astore;
      std::vector<int64_t> cost(n_inst, 0);
      std::vector<char> grouped(n_inst, 0);
      std::map<std::pair<const Space*, std::string>, int> group_of;  // (space, gate layout) -> group
      const char* genv = getenv("SQ_BATCH_GROUPS");
      const bool groups_on = !(genv && genv[0] == '0');
      int theta_off = 0, grad_off = 0;
      for (int i = 0; i < n_inst; ++i) {
        const Space& s = spaces[i]->s;
        const Program& p = programs[i]->p;
        const Hamiltonian& H = hams[i]->h;
        SQ_CHECK(p.space == &s && H.space == &s, SQ_ERR_ARG, "program / Hamiltonian built for another space");
        SQ_CHECK(s.dim < (int64_t)1 << 30, SQ_ERR_ARG, "batched spaces must have fewer than 2^30 determinants");
        if (groups_on && s.dim >= kGroupMinDim && H.d_Ma && row_sigma_group_ok(s) && tiles_group_ok(s, p)) {
          // a large instance: grouped with the instances of the same space and gate layout
          for (const GateDev& gd : p.gates)
            SQ_CHECK(gd.kind == GC_ROT || gd.kind == GC_SA_SINGLE, SQ_ERR_ARG,
                     "batched launches support rotation gates (single, double, sa_single, pair doubles) only");
          const auto key = std::make_pair(&s, layout_key(p));
          auto it = group_of.find(key);
          if (it == group_of.end()) {
            it = group_of.emplace(key, (int)B.groups.size()).first;
            BatchGroup G;
            G.space = &s;
            G.prog = &p;
            G.ng = (int)p.gates.size();
            B.groups.push_back(G);
          }
          grouped[i] = 1;
          BatchInstance I{};
          I.theta0 = theta_off;
          I.nparam = (int)p.gates.size();
          theta_off += I.nparam;
          I.out0 = grad_off;
          B.grad_off.push_back(grad_off);
          grad_off += I.nparam;
          B.inst.push_back(I);
          B.group_members.push_back({it->second, i});
          continue;
        }
        if (!sp_index.count(spaces[i])) {
          sp_index[spaces[i]] = (int)bsp.size();
          BatchSpace b{};
          b.n = s.n;
          b.nsa = s.nsa;
          b.nsb = s.nsb;
          b.npair = s.npair;
          b.K4 = (s.npair + 1 + 3) / 4 * 4;
          b.kpad = (s.npair + 1 + 15) / 16 * 16;
          b.astr = s.d_astr;
          b.bstr = s.d_bstr;
          b.arank = s.d_arank;
          b.brank = s.d_brank;
          b.aexc = s.d_aexc;
          b.bexc = s.d_bexc;
          bsp.push_back(b);
        }
        const BatchSpace& bs = bsp[sp_index[spaces[i]]];
        BatchInstance I{};
        I.space = sp_index[spaces[i]];
        I.gate0 = (int)rots.size();
        for (int g = 0; g < (int)p.gates.size(); ++g) {
          const GateDev& gd = p.gates[g];
          SQ_CHECK(gd.kind == GC_ROT || gd.kind == GC_SA_SINGLE, SQ_ERR_ARG,
                   "batched launches support rotation gates (single, double, sa_single, pair doubles) only");
          const int halves = gd.kind == GC_SA_SINGLE ? 2 : 1;
          for (int h = 0; h < halves; ++h) {
            const Term& T = p.terms[gd.term_begin + 2 * h];
            BatchRot R{};
            spin_list(s.h_astr, T.a, R.aoff, R.na);
            spin_list(s.h_bstr, T.b, R.boff, R.nb);
            R.sgn = T.coef < 0 ? -1.0 : 1.0;
            R.scale = halves == 2 ? 1.0 : gd.scale;
            R.theta_idx = g;
            R.grad_idx = g;
            rots.push_back(R);
          }
        }
        I.ngate = (int)rots.size() - I.gate0;
        I.theta0 = theta_off;
        I.nparam = (int)p.gates.size();
        theta_off += I.nparam;
        I.out0 = grad_off;
        B.grad_off.push_back(grad_off);
        grad_off += I.nparam;
        // A copied from the device Hamiltonian (K4 x kpad, same packing)
        I.A_off = (int)Astore.size();
        std::vector<double> A((size_t)bs.K4 * bs.kpad);
        SQ_HIP(hipMemcpy(A.data(), H.d_A, sizeof(double) * A.size(), hipMemcpyDeviceToHost));
        Astore.insert(Astore.end(), A.begin(), A.end());
        I.e_core = H.e_core;
        B.inst.push_back(I);
        B.dmax = std::max<int64_t>(B.dmax, s.dim);
        B.slot = std::max<int64_t>(B.slot, (int64_t)(2 + bs.npair) * s.dim);
        B.lds = std::max<size_t>(B.lds, sizeof(double) * bs.K4 * bs.kpad);
        cost[i] = s.dim * (int64_t)(2 * I.ngate + 4 + 2 * bs.npair + (bs.npair * bs.K4) / 8);
      }
      SQ_CHECK(B.lds <= 65536, SQ_ERR_ARG, "batched spaces limited to 12 orbitals (LDS)");
      for (const auto& bs : bsp)
        SQ_CHECK(bs.kpad <= 16 * kBatchMT && bs.npair < 91, SQ_ERR_ARG, "batched spaces limited to 12 orbitals");
      B.ngrad = grad_off;
      // longest processing time first, each to the least-loaded workgroup
      // (the instances of the persistent launch: those not in a group)
      int n_small = 0;
      for (int i = 0; i < n_inst; ++i) n_small += grouped[i] ? 0 : 1;
      int nwg = n_workgroups > 0 ? n_workgroups : 256;
      nwg = std::min(nwg, n_small);
      std::vector<int> order;
      for (int i = 0; i < n_inst; ++i)
        if (!grouped[i]) order.push_back(i);
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
      std::vector<std::vector<int32_t>> wg_inst(nwg);
      std::vector<int64_t> load(nwg, 0);
      for (int i : order) {
        int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        wg_inst[w].push_back(i);
        load[w] += cost[i];
      }
      std::vector<int32_t> wl, wo{0};
      for (auto& l : wg_inst) {
        wl.insert(wl.end(), l.begin(), l.end());
        wo.push_back((int32_t)wl.size());
      }
      B.nwg = nwg;
      B.dev.spaces = bupload(B, bsp);
      B.dev.inst = bupload(B, B.inst);
      B.dev.rots = bupload(B, rots);
      B.dev.lists = bupload(B, lists);
      B.dev.A = bupload(B, Astore);
      B.dev.wg_list = bupload(B, wl);
      B.dev.wg_off = bupload(B, wo);
      B.dev.slot_doubles = B.slot;
      B.dev.dmax = B.dmax;
      double* scratch = nullptr;
      SQ_HIP(hipMalloc(&scratch, sizeof(double) * std::max<size_t>((size_t)B.slot * nwg, 1)));
      B.allocs.push_back(scratch);
      B.dev.scratch = scratch;
      double* en = nullptr;
      SQ_HIP(hipMalloc(&en, sizeof(double) * n_inst));
      B.allocs.push_back(en);
      B.dev.energy = en;
      double* gr = nullptr;
      SQ_HIP(hipMalloc(&gr, sizeof(double) * std::max<int64_t>(grad_off, 1)));
      B.allocs.push_back(gr);
      B.dev.grad = gr;
      SQ_HIP(hipMalloc(&B.d_theta, sizeof(double) * std::max(theta_off, 1)));
      B.allocs.push_back(B.d_theta);
      B.dev.theta = B.d_theta;
      B.d_psi0 = bupload(B, std::vector<const double*>(n_inst, nullptr));
      B.dev.psi0 = B.d_psi0;
      build_groups(B, hams, spaces);
      if (!B.groups.empty() && n_small > 0) {
        SQ_HIP(hipStreamCreateWithFlags(&B.side, hipStreamNonBlocking));
        SQ_HIP(hipEventCreateWithFlags(&B.ev_fork, hipEventDisableTiming));
        SQ_HIP(hipEventCreateWithFlags(&B.ev_join, hipEventDisableTiming));
      }
    } catch (...) {
      delete bo;
      throw;
    }
    *out = bo;
  });
}

int sq_batch_destroy(sq_batch* b) {
  return guarded([&] { delete b; });
}

int64_t sq_batch_num_params(const sq_batch* b) { return b ? b->b.ngrad : -1; }

/* theta: host array of all instances' parameters (concatenated in instance
 * order); copied into the batch's device buffer. */
int sq_batch_set_theta(sq_batch* bo, const double* theta, void* stream) {
  return guarded([&] {
    SQ_CHECK(bo && theta, SQ_ERR_ARG, "NULL argument");
    SQ_HIP(hipMemcpyAsync(bo->b.d_theta, theta, sizeof(double) * bo->b.ngrad, hipMemcpyHostToDevice,
                          as_stream(stream)));
  });
}

/* psi0: host array of n_inst device pointers, each to its instance's start
 * state (engine convention, space dim doubles) or NULL for the HF
 * determinant.  The pointers are read at every launch (graph replays
 * included): the states must stay allocated while the batch runs. */
int sq_batch_set_start_states(sq_batch* bo, const double* const* psi0, void* stream) {
  return guarded([&] {
    SQ_CHECK(bo && psi0, SQ_ERR_ARG, "NULL argument");
    Batch& B = bo->b;
    hipStream_t st = as_stream(stream);
    SQ_HIP(hipMemcpyAsync(B.d_psi0, psi0, sizeof(double*) * B.inst.size(), hipMemcpyHostToDevice, st));
    SQ_HIP(hipStreamSynchronize(st));
  });
}

using BatchKernel = void (*)(BatchDev);

static BatchKernel batch_kernel(const Batch& B) {
  static const BatchKernel table[3][2] = {{k_batch<1, 1>, k_batch<1, 2>},
                                          {k_batch<2, 1>, k_batch<2, 2>},
                                          {k_batch<4, 1>, k_batch<4, 2>}};
  const int u = B.rot_pairs == 1 ? 0 : (B.rot_pairs == 2 ? 1 : 2);
  return table[u][B.gemm_tiles - 1];
}

// The persistent launch and the groups write disjoint outputs; with both,
// the persistent launch runs on the batch's side stream between a fork and
// a join event, beside the groups' kernels on st.
static void batch_launch(Batch& B, hipStream_t st) {
  if (B.nwg > 0 && !B.groups.empty()) {
    SQ_HIP(hipEventRecord(B.ev_fork, st));
    SQ_HIP(hipStreamWaitEvent(B.side, B.ev_fork, 0));
    SQ_PROF(PC_BATCH, B.side, B.bytes, B.flops, batch_kernel(B)<<<B.nwg, kBT, B.lds, B.side>>>(B.dev));
    for (const BatchGroup& G : B.groups) group_launch(B, G, st);
    SQ_HIP(hipEventRecord(B.ev_join, B.side));
    SQ_HIP(hipStreamWaitEvent(st, B.ev_join, 0));
    return;
  }
  if (B.nwg > 0) SQ_PROF(PC_BATCH, st, B.bytes, B.flops, batch_kernel(B)<<<B.nwg, kBT, B.lds, st>>>(B.dev));
  for (const BatchGroup& G : B.groups) group_launch(B, G, st);
}

/* Kernel shape: rotation pairs per thread in flight (1, 2 or 4) and sigma
 * GEMM column tiles per wave (1 or 2); drops a captured graph. */
int sq_batch_configure(sq_batch* bo, int rot_pairs, int gemm_tiles) {
  return guarded([&] {
    SQ_CHECK(bo, SQ_ERR_ARG, "NULL argument");
    SQ_CHECK(rot_pairs == 1 || rot_pairs == 2 || rot_pairs == 4, SQ_ERR_ARG, "rot_pairs must be 1, 2 or 4");
    SQ_CHECK(gemm_tiles == 1 || gemm_tiles == 2, SQ_ERR_ARG, "gemm_tiles must be 1 or 2");
    Batch& B = bo->b;
    if (B.rot_pairs != rot_pairs || B.gemm_tiles != gemm_tiles) B.drop_graph();
    B.rot_pairs = rot_pairs;
    B.gemm_tiles = gemm_tiles;
  });
}

int sq_batch_run(sq_batch* bo, void* stream) {
  return guarded([&] {
    check_groups(bo->b);
    batch_launch(bo->b, as_stream(stream));
  });
}

/* Capture the launch in a hipGraph (once) and replay it. */
int sq_batch_run_graph(sq_batch* bo, void* stream) {
  return guarded([&] {
    Batch& B = bo->b;
    hipStream_t st = as_stream(stream);
    check_groups(B);
    if (!B.exec) {
      // the groups' kernels take their shared-memory attributes on an eager
      // launch first (not a stream operation, so kept out of the capture)
      if (!B.groups.empty()) {
        batch_launch(B, st);
        SQ_HIP(hipStreamSynchronize(st));
      }
      hipStream_t cap;
      SQ_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      const bool prof_was = prof_pause();  // event records would become graph nodes
      SQ_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
      hipError_t e = hipSuccess;
      try {
        batch_launch(B, cap);
        e = hipGetLastError();
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(cap, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cap);
        prof_resume(prof_was);
        throw;
      }
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(cap, &g);
      (void)hipStreamDestroy(cap);
      prof_resume(prof_was);
      SQ_CHECK(e == hipSuccess && e2 == hipSuccess, SQ_ERR_HIP, "graph capture failed");
      B.graph = g;
      SQ_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
    }
    SQ_HIP(hipGraphLaunch(B.exec, st));
  });
}

int sq_batch_results(const sq_batch* bo, double* energies, double* grads, void* stream) {
  return guarded([&] {
    SQ_CHECK(bo && energies && grads, SQ_ERR_ARG, "NULL argument");
    const Batch& B = bo->b;
    hipStream_t st = as_stream(stream);
    SQ_HIP(hipMemcpyAsync(energies, B.dev.energy, sizeof(double) * B.inst.size(), hipMemcpyDeviceToHost, st));
    SQ_HIP(hipMemcpyAsync(grads, B.dev.grad, sizeof(double) * B.ngrad, hipMemcpyDeviceToHost, st));
    SQ_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"
