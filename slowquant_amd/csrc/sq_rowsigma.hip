// sigma = H psi by alpha rows, with no intermediate in HBM (replaces the
// H application of propagate_state([H]) / expectation_value,
// operator_state_algebra.py:472-824, for H = hamiltonian_0i_0a,
// operators.py:476-529).
//
// Spin split.  With E+_k = E+a_k + E+b_k over packed pairs k = (p >= q),
// A(k,k') = 1/2 g(k,k') and h'_k = h_pq - 1/2 sum_r g_prrq (the alpha and
// beta parts commute):
//     H = Ha + Hb + sum_{k,k'} g(k,k') E+b_k E+a_k',
//     Ha = sum_k h'_k E+a_k + sum_{k,k'} A(k,k') E+a_k E+a_k'   (Hb alike),
// so
//     sigma = Ma C + sum_k E+b_k G_k,
//     G_k   = h'_k psi + sum_k' A(k,k') (2 E+a_k' + E+b_k') psi,
// where C is psi as an [alpha string][beta string] matrix and Ma (+ e_core
// on its diagonal) is the matrix of Ha over alpha strings, built once per
// Hamiltonian (k_alpha_ham).
//   1. k_alpha_gemm: S = Ma C, FP64 MFMA, split over the contraction index
//      into nsplit partial products (summed in a fixed order in step 2).
//   2. k_sigma_rows: one block per group of R alpha rows.  Row I_a of G is
//      A . B with B[k'][I_b] = 2 ca_k' psi[Pa_k'(I_a), I_b] + cb_k'(I_b)
//      psi[I_a, Pb_k'(I_b)] (rows npair: psi), generated in registers (the
//      alpha partner row is uniform over the row, the beta partner is read
//      from the row staged in LDS) and contracted on FP64 MFMA.  Each wave
//      computes one 16-column tile of G for all k per round; the tiles of a
//      round are staged in LDS and every target amplitude I_b of the rows
//      adds, in a fixed order, the staged G_k[I_a, J_b] of its beta
//      neighbours (J_b = partner of I_b under E+b_k).  G never leaves LDS.
// Deterministic: every sum runs in a fixed order.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include <rocblas/rocblas.h>

#include "sq_internal.h"

namespace sq {

void reduce_impl(const double* partial, int64_t n, double* out_dev, int slot, bool accumulate, double factor,
                 hipStream_t st);

namespace {

constexpr int kRowThreads = 512;
constexpr int kRowWaves = kRowThreads / 64;
constexpr int kMaxTargets = 8;  // targets (alpha row, beta string) per thread
constexpr size_t kLdsMax = 160 * 1024 - 256;  // dynamic LDS per block (the static part stays below 256 bytes)
constexpr int kMaxSplit = 4;     // alpha GEMM partial products (workspace is sized for this many)

__host__ __device__ inline int rup(int x, int m) { return (x + m - 1) / m * m; }

__global__ void k_zero(double* __restrict__ x, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Space tables.
// rta[k][s] (k < K8 + 16): {partner row * nsb, 2 * coefficient} of E+a_k on
// alpha string s (diagonal: 2 * occupation; absent / k >= npair: {own, 0}).
// rtb[k][s] (k < K8 + 32): {partner column, coefficient} of E+b_k on beta
// string s (diagonal: occupation; k == npair: {own, 1}; absent: {own, 0}).
__global__ void k_row_tables(int npair, int nstr, int nsb, int alpha, const int32_t* __restrict__ exc,
                             int2* __restrict__ tab) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (s >= nstr) return;
  const int own = alpha ? s * nsb : s;
  const int f = alpha ? 2 : 1;
  int2 e = make_int2(own, 0);
  if (k < npair) {
    int p = 0;
    while ((p + 1) * (p + 2) / 2 <= k) ++p;
    const int q = k - p * (p + 1) / 2;
    const int32_t x = exc[(size_t)k * nstr + s];
    if (p == q) e = make_int2(own, f * x);
    else if (x != 0) e = make_int2(alpha ? (abs(x) - 1) * nsb : abs(x) - 1, x > 0 ? f : -f);
  } else if (k == npair && !alpha) {
    e = make_int2(own, 1);
  }
  tab[(size_t)k * nstr + s] = e;
}

// Beta neighbour lists: for target string s, the entries (J, k, sign) with
// (E+b_k)[s][J] != 0 (J = partner, or s itself with occupation 1 on the
// diagonal), packed J | k << 14 | negative << 31 and sorted by J; tile[s][c]
// (c = 0..tpr) counts the entries with J < 16 c.
__global__ void k_nbr_count(int npair, int nsb, const int32_t* __restrict__ bexc, int32_t* __restrict__ cnt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsb) return;
  int c = 0;
  for (int k = 0; k < npair; ++k) c += bexc[(size_t)k * nsb + s] != 0;
  cnt[s] = c;
}

__global__ void k_nbr_fill(int npair, int nsb, int tpr, const int32_t* __restrict__ bexc,
                           const int32_t* __restrict__ off, uint32_t* __restrict__ nbr, uint8_t* __restrict__ tile) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsb) return;
  uint32_t* out = nbr + off[s];
  int m = 0, k = 0;
  for (int p = 0; p < 32 && k < npair; ++p) {
    for (int q = 0; q <= p; ++q, ++k) {
      const int32_t x = bexc[(size_t)k * nsb + s];
      if (x == 0) continue;
      uint32_t e;
      if (p == q) e = (uint32_t)s | ((uint32_t)k << 14);
      else e = (uint32_t)(abs(x) - 1) | ((uint32_t)k << 14) | (x < 0 ? 0x80000000u : 0u);
      // insertion by partner column (stable: equal columns keep k order)
      int j = m++;
      while (j > 0 && (out[j - 1] & 0x3FFFu) > (e & 0x3FFFu)) {
        out[j] = out[j - 1];
        --j;
      }
      out[j] = e;
    }
  }
  int e = 0;
  for (int c = 0; c <= tpr; ++c) {
    while (e < m && (int)(out[e] & 0x3FFFu) < 16 * c) ++e;
    tile[(size_t)s * (tpr + 1) + c] = (uint8_t)e;
  }
}

// ---------------------------------------------------------------------------
// Ma[J][I] = <I|Ha|J> + e_core delta_IJ (symmetric), one block per column J:
//   Ha|J> = sum_k E+a_k u_k,  u_k = h'_k |J> + sum_k' A(k,k') E+a_k' |J>,
// where E+a_k'|J> is J itself (diagonal k', coefficient = occupation) or one
// single excitation of J (distinct for distinct off-diagonal k'): u_k lives
// on at most 1 + Na (n - Na) strings ("slots"), and
//   <I|Ha|J> = sum_k c_k(I) u_k[slot(P_k(I))].
__global__ void k_alpha_ham(int nsa, int npair, int kpad, const int32_t* __restrict__ aexc,
                            const double* __restrict__ At, double e_core, double* __restrict__ Ma) {
  extern __shared__ unsigned char smem_raw[];
  const int J = blockIdx.x;
  // LDS: slot of every string (-1: none) | k' -> slot | k' coefficient | u[npair][nslot_max]
  int16_t* slotmap = reinterpret_cast<int16_t*>(smem_raw);
  int16_t* kslot = slotmap + rup(nsa, 8);
  int32_t* kcoef = reinterpret_cast<int32_t*>(kslot + rup(npair, 8));
  double* u = reinterpret_cast<double*>(kcoef + rup(npair, 2));
  __shared__ int nslot;
  for (int i = threadIdx.x; i < nsa; i += blockDim.x) slotmap[i] = -1;
  __syncthreads();
  if (threadIdx.x == 0) {
    slotmap[J] = 0;
    int ns = 1;
    int k = 0;
    for (int p = 0; k < npair; ++p)
      for (int q = 0; q <= p; ++q, ++k) {
        const int32_t x = aexc[(size_t)k * nsa + J];
        if (p == q || x == 0) {
          kslot[k] = 0;
          kcoef[k] = p == q ? x : 0;
        } else {
          const int P = abs(x) - 1;
          slotmap[P] = (int16_t)ns;
          kslot[k] = (int16_t)ns++;
          kcoef[k] = x > 0 ? 1 : -1;
        }
      }
    nslot = ns;
  }
  __syncthreads();
  const int ns = nslot;
  // u[k][slot] in a fixed k' order
  for (int i = threadIdx.x; i < npair * ns; i += blockDim.x) {
    const int k = i / ns, sl = i - k * ns;
    double v = sl == 0 ? At[(size_t)npair * kpad + k] : 0.0;
    for (int kp = 0; kp < npair; ++kp)
      if (kslot[kp] == sl && kcoef[kp] != 0) v += At[(size_t)kp * kpad + k] * (double)kcoef[kp];
    u[i] = v;
  }
  __syncthreads();
  for (int I = threadIdx.x; I < nsa; I += blockDim.x) {
    double v = I == J ? e_core : 0.0;
    int k = 0;
    for (int p = 0; k < npair; ++p)
      for (int q = 0; q <= p; ++q, ++k) {
        const int32_t x = aexc[(size_t)k * nsa + I];
        if (x == 0) continue;
        const int K = p == q ? I : abs(x) - 1;
        const int sl = slotmap[K];
        if (sl < 0) continue;
        const double c = p == q ? (double)x : (x > 0 ? 1.0 : -1.0);
        v += c * u[k * ns + sl];
      }
    Ma[(size_t)J * nsa + I] = v;
  }
}

// ---------------------------------------------------------------------------
// out[z][I][c] = sum_{J in split z} Ma[I][J] C[J][c] (Ma nsa x nsa, C nsa x nsb).
// 256 threads, 64 x 64 output tile; wave w owns rows 32 (w >> 1) and columns
// 32 (w & 1) as 2 x 2 tiles of v_mfma_f64_16x16x4f64; the contraction runs
// in 16-deep slices staged in LDS (the next slice loaded into registers
// during the MFMAs of the current one).
__global__ __launch_bounds__(256) void k_alpha_gemm(const double* __restrict__ Ma, const double* __restrict__ C,
                                                    double* __restrict__ out, int nsa, int nsb, int kchunk) {
  typedef double double4_t __attribute__((ext_vector_type(4)));
  __shared__ double sM[16][65];  // [k][i]
  __shared__ double sC[16][64];  // [k][c]
  const int i0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const int kb = blockIdx.z * kchunk, ke = min(nsa, kb + kchunk);
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, kq = lane >> 4, li = lane & 15;
  const int wi = (wave >> 1) * 32, wc = (wave & 1) * 32;
  double4_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = double4_t{0.0, 0.0, 0.0, 0.0};
  // loaders: Ma slice 64 x 16 (thread: row t/4, columns 4 (t%4) ..), C slice 16 x 64 (row t/16, columns 4 (t%16) ..)
  const int mr = t >> 2, mc = (t & 3) * 4;
  const int cr = t >> 4, cc = (t & 15) * 4;
  double rm[4], rc[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int I = i0 + mr, K = k0 + mc + j;
      rm[j] = (I < nsa && K < ke) ? Ma[(size_t)I * nsa + K] : 0.0;
      const int Kc = k0 + cr, c = c0 + cc + j;
      rc[j] = (Kc < ke && c < nsb) ? C[(size_t)Kc * nsb + c] : 0.0;
    }
  };
  if (kb < ke) load(kb);
  for (int k0 = kb; k0 < ke; k0 += 16) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      sM[mc + j][mr] = rm[j];
      sC[cr][cc + j] = rc[j];
    }
    __syncthreads();
    if (k0 + 16 < ke) load(k0 + 16);
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      const double a0 = sM[kk + kq][wi + li], a1 = sM[kk + kq][wi + 16 + li];
      const double b0 = sC[kk + kq][wc + li], b1 = sC[kk + kq][wc + 16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
  double* o = out + (size_t)blockIdx.z * nsa * nsb;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int I = i0 + wi + a * 16 + kq + 4 * r, c = c0 + wc + b * 16 + li;
        if (I < nsa && c < nsb) o[(size_t)I * nsb + c] = acc[a][b][r];
      }
}

// ---------------------------------------------------------------------------
// One 16-column tile of G for the alpha row whose psi row is ps (LDS) and
// whose alpha operand entries are ra (LDS): acc[t] holds the 16 x 16 block
// of rows m0 + 16 t (the A slice sA [K8][W] is in LDS); column ib = c*16 + li
// (columns past nsb compute garbage that is never used).  The k-steps are
// unrolled by two with ping-pong registers: step s issues the alpha
// amplitude and beta table loads of later steps before its MFMAs.
typedef double double4_t __attribute__((ext_vector_type(4)));

template <int MT, int NSTEPS = 0>
__device__ __forceinline__ void row_tile(double4_t (&acc)[MT], const double* __restrict__ sA,
                                         const double* __restrict__ ps, const int2* __restrict__ ra,
                                         const double* __restrict__ psi, const int2* __restrict__ rtb, int nsb,
                                         int ib, int kq, int li, int nsteps_rt) {
  // NSTEPS > 0: the step count is known at compile time and the loop is
  // unrolled, so the loads issued ahead stay in flight across steps (a loop
  // header makes the compiler wait for every outstanding load)
  const int nsteps = NSTEPS > 0 ? NSTEPS : nsteps_rt;
  constexpr int W = MT * 16;
  const uint32_t ibc = (uint32_t)(ib < nsb ? ib : nsb - 1);
#pragma unroll
  for (int t = 0; t < MT; ++t) acc[t] = double4_t{0.0, 0.0, 0.0, 0.0};
  const uint32_t sb = 4u * (uint32_t)nsb;
  const uint32_t tb_i = (uint32_t)kq * (uint32_t)nsb + ibc;
  // slot 0: even steps, slot 1: odd steps
  int2 tb0 = rtb[tb_i], tb1 = rtb[tb_i + sb];
  int2 a = ra[kq];
  double xa0 = psi[(uint32_t)a.x + ibc];
  int ca0 = a.y;
  double xb0 = ps[tb0.x];
  int cb0 = tb0.y;
  tb0 = rtb[tb_i + 2 * sb];
  double xa1, xb1;
  int ca1, cb1;
  const double* arow = sA + (size_t)kq * W + li;
#pragma unroll
  for (int s = 0; s < nsteps; s += 2) {
    a = ra[(s + 1) * 4 + kq];
    xa1 = psi[(uint32_t)a.x + ibc];
    ca1 = a.y;
    xb1 = ps[tb1.x];
    cb1 = tb1.y;
    tb1 = rtb[tb_i + (uint32_t)(s + 3) * sb];
    __asm__ volatile("" ::: "memory");
    {
      const double b = (double)ca0 * xa0 + (double)cb0 * xb0;
#pragma unroll
      for (int t = 0; t < MT; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(arow[t * 16], b, acc[t], 0, 0, 0);
    }
    arow += 4 * W;
    a = ra[(s + 2) * 4 + kq];
    xa0 = psi[(uint32_t)a.x + ibc];
    ca0 = a.y;
    xb0 = ps[tb0.x];
    cb0 = tb0.y;
    tb0 = rtb[tb_i + (uint32_t)(s + 4) * sb];
    __asm__ volatile("" ::: "memory");
    {
      const double b = (double)ca1 * xa1 + (double)cb1 * xb1;
#pragma unroll
      for (int t = 0; t < MT; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(arow[t * 16], b, acc[t], 0, 0, 0);
    }
    arow += 4 * W;
  }
}

// Row kernel, staged form (deterministic): LDS = A slice [K8][W] | staged G
// tiles [waves][W][16] | psi rows [R][nsb] | alpha operand rows [R][K8 + 16].
// Per round every wave computes one tile; after a barrier each target adds
// its beta neighbours' staged values in neighbour-list order.
template <int MT>
__global__ __launch_bounds__(kRowThreads) void k_sigma_rows(
    const double* __restrict__ At, int kpad, int K4, int K8, int mgroups, const double* __restrict__ psi,
    const double* __restrict__ sin, int nsplit, double* __restrict__ sigma, double* __restrict__ partial,
    const int2* __restrict__ rta, const int2* __restrict__ rtb, const uint32_t* __restrict__ nbr,
    const int32_t* __restrict__ nbr_off, const uint8_t* __restrict__ nbr_tile, int nsa, int nsb, int R) {
  constexpr int W = MT * 16;
  extern __shared__ double smem[];
  __shared__ double red[kRowWaves];
  const int tpr = (nsb + 15) >> 4;
  const int RA = K8 + 16;
  double* sA = smem;
  double* sG = sA + (size_t)K8 * W;
  double* sPsi = sG + (size_t)kRowWaves * W * 16;
  int2* sRA = reinterpret_cast<int2*>(sPsi + (size_t)R * nsb);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, kq = lane >> 4, li = lane & 15;
  const int row0 = blockIdx.x * R;
  const int nr = min(R, nsa - row0);
  const int ntg = nr * nsb;
  const size_t base = (size_t)row0 * nsb;
  for (int i = tid; i < ntg; i += kRowThreads) sPsi[i] = psi[base + i];
  for (int i = tid; i < nr * RA; i += kRowThreads) {
    const int r = i / RA, k = i - r * RA;
    sRA[i] = rta[(size_t)k * nsa + row0 + r];
  }
  double tacc[kMaxTargets];
#pragma unroll
  for (int i = 0; i < kMaxTargets; ++i) tacc[i] = 0.0;
  const int ntiles = nr * tpr;
  const int nsteps = K8 / 4;
  for (int mg = 0; mg < mgroups; ++mg) {
    const int m0 = mg * W;
    __syncthreads();  // the previous m-group's reads of sA are done
    for (int i = tid; i < K8 * W; i += kRowThreads) {
      const int k = i / W, m = m0 + (i - k * W);
      sA[i] = (k < K4 && m < kpad) ? At[(size_t)k * kpad + m] : 0.0;
    }
    __syncthreads();
    for (int f0 = 0; f0 < ntiles; f0 += kRowWaves) {
      const int f = f0 + wave;
      if (f < ntiles) {
        const int r = f / tpr, c = f - r * tpr;
        double4_t acc[MT];
        row_tile<MT>(acc, sA, sPsi + (size_t)r * nsb, sRA + (size_t)r * RA, psi, rtb, nsb, c * 16 + li, kq, li,
                     nsteps);
        double* g = sG + (size_t)wave * W * 16 + li;
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
          for (int q = 0; q < 4; ++q) g[(t * 16 + kq + 4 * q) * 16] = acc[t][q];
      }
      __syncthreads();
      // every target adds its beta neighbours' staged G_k values, by partner column
#pragma unroll
      for (int i = 0; i < kMaxTargets; ++i) {
        const int tau = tid + i * kRowThreads;
        if (tau >= ntg) break;
        const int r = tau / nsb, ib = tau - r * nsb;
        const int lo = max(f0, r * tpr), hi = min(f0 + kRowWaves, (r + 1) * tpr);
        if (lo >= hi) continue;
        const uint8_t* tl = nbr_tile + (size_t)ib * (tpr + 1);
        const int eb = nbr_off[ib];
        const int e0 = eb + tl[lo - r * tpr], e1 = eb + tl[hi - r * tpr];
        const int fb = r * tpr - f0;
        double v = tacc[i];
        for (int e = e0; e < e1; ++e) {
          const uint32_t x = nbr[e];
          const int jb = (int)(x & 0x3FFFu), k = (int)((x >> 14) & 0xFFu) - m0;
          if (k < 0 || k >= W) continue;
          const double gv = sG[((size_t)(fb + (jb >> 4)) * W + k) * 16 + (jb & 15)];
          v += (x >> 31) ? -gv : gv;
        }
        tacc[i] = v;
      }
      __syncthreads();
    }
  }
  const size_t dim = (size_t)nsa * nsb;
  double e = 0.0;
#pragma unroll
  for (int i = 0; i < kMaxTargets; ++i) {
    const int tau = tid + i * kRowThreads;
    if (tau >= ntg) break;
    double v = tacc[i];
    for (int z = 0; z < nsplit; ++z) v += sin[(size_t)z * dim + base + tau];
    sigma[base + tau] = v;
    e += v * sPsi[tau];
  }
  for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
  if (lane == 0) red[wave] = e;
  __syncthreads();
  if (tid == 0) {
    double tsum = 0.0;
    for (int w = 0; w < kRowWaves; ++w) tsum += red[w];
    partial[blockIdx.x] = tsum;
  }
}

// One 128 x 64 tile of Ma C over the contraction range [kb, ke) for the
// fused sigma launch (512 threads: wave w owns rows 32 (w >> 1) and columns
// 32 (w & 1) as 2 x 2 MFMA tiles); the tile is added to sigma by global
// atomics and its <psi|.> partial returned (thread 0's value is the block's).
constexpr int kGemmRows = 128, kGemmCols = 64;

__device__ __forceinline__ double alpha_gemm_tile(int i0, int j0, int kb, int ke, const double* __restrict__ Ma,
                                                  const double* __restrict__ C, double* __restrict__ sigma, int nsa,
                                                  int nsb, double* smem, double* red) {
  double(*sM)[kGemmRows + 1] = reinterpret_cast<double(*)[kGemmRows + 1]>(smem);          // [16][129]: [k][i]
  double(*sC)[kGemmCols + 1] = reinterpret_cast<double(*)[kGemmCols + 1]>(smem + 16 * (kGemmRows + 1));
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, kq = lane >> 4, li = lane & 15;
  const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
  double4_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = double4_t{0.0, 0.0, 0.0, 0.0};
  const int mr = t >> 2, mc = (t & 3) * 4;   // Ma slice 128 x 16
  const int cr = t >> 5, cc = (t & 31) * 2;  // C slice 16 x 64
  double rm[4], rc[2];
  auto load = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int I = i0 + mr, K = k0 + mc + j;
      rm[j] = (I < nsa && K < ke) ? Ma[(size_t)I * nsa + K] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int Kc = k0 + cr, c = j0 + cc + j;
      rc[j] = (Kc < ke && c < nsb) ? C[(size_t)Kc * nsb + c] : 0.0;
    }
  };
  if (kb < ke) load(kb);
  for (int k0 = kb; k0 < ke; k0 += 16) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) sM[mc + j][mr] = rm[j];
#pragma unroll
    for (int j = 0; j < 2; ++j) sC[cr][cc + j] = rc[j];
    __syncthreads();
    if (k0 + 16 < ke) load(k0 + 16);
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      const double a0 = sM[kk + kq][wi + li], a1 = sM[kk + kq][wi + 16 + li];
      const double b0 = sC[kk + kq][wj + li], b1 = sC[kk + kq][wj + 16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
  double e = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int I = i0 + wi + a * 16 + kq + 4 * r, c = j0 + wj + b * 16 + li;
        if (I < nsa && c < nsb) {
          const size_t x = (size_t)I * nsb + c;
          unsafeAtomicAdd(sigma + x, acc[a][b][r]);
          e += acc[a][b][r] * C[x];
        }
      }
  for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
  if (lane == 0) red[wave] = e;
  __syncthreads();
  double tsum = 0.0;
  if (t == 0)
    for (int w = 0; w < kRowWaves; ++w) tsum += red[w];
  return tsum;
}

// Row kernel, scatter form: LDS = A slice [K8][W] | sigma rows [R][nsb] |
// psi rows [R][nsb] | alpha operand rows [R][K8 + 16] (small enough for two
// blocks per CU).  Waves take tiles independently; each output G_k[I_a, J_b]
// goes to the sigma row at P_k(J_b) (E+b_k is symmetric, so the partner and
// coefficient are rtb[k][J_b]) by an LDS atomic add.  The order of those
// adds varies between runs, so results agree to rounding, not bit for bit.
// FUSED: the launch also holds the Ma C tiles as extra blocks after the row
// blocks (they fill the CUs the last row blocks leave idle); every block adds
// its part to sigma (zeroed beforehand) by global atomics.
template <int MT, bool FUSED>
__global__ __launch_bounds__(kRowThreads) __attribute__((amdgpu_waves_per_eu(MT <= 5 ? 4 : 2, 8))) void k_sigma_rows_scatter(
    const double* __restrict__ At, int kpad, int K4, int K8, int mgroups, int npair, const double* __restrict__ psi,
    const double* __restrict__ sin, int nsplit, double* __restrict__ sigma, double* __restrict__ partial,
    const int2* __restrict__ rta, const int2* __restrict__ rtb, int nsa, int nsb, int R, int xcd, int nrowblk,
    const double* __restrict__ Ma, int gkchunk, int64_t A_stride, int64_t Ma_stride, int64_t vec_stride,
    int64_t part_stride) {
  constexpr int W = MT * 16;
  extern __shared__ double smem[];
  __shared__ double red[kRowWaves];
  // instance blockIdx.y of a group (zero strides: one instance)
  At += blockIdx.y * A_stride;
  psi += blockIdx.y * vec_stride;
  sigma += blockIdx.y * vec_stride;
  partial += blockIdx.y * part_stride;
  if (FUSED) Ma += blockIdx.y * Ma_stride;
  if (FUSED && (int)blockIdx.x >= nrowblk) {
    const int g = (int)blockIdx.x - nrowblk;
    const int tn = (nsb + kGemmCols - 1) / kGemmCols, tm = (nsa + kGemmRows - 1) / kGemmRows;
    const int z = g / (tm * tn), rem = g - z * (tm * tn);
    const int kb = z * gkchunk;
    const double e = alpha_gemm_tile((rem / tn) * kGemmRows, (rem % tn) * kGemmCols, kb, min(nsa, kb + gkchunk), Ma,
                                     psi, sigma, nsa, nsb, smem, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = e;
    return;
  }
  const int tpr = (nsb + 15) >> 4;
  const int RA = K8 + 16;
  double* sA = smem;
  double* sAcc = sA + (size_t)K8 * W;
  double* sPsi = sAcc + (size_t)R * nsb;
  int2* sRA = reinterpret_cast<int2*>(sPsi + (size_t)R * nsb);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, kq = lane >> 4, li = lane & 15;
  // with xcd, XCD x (workgroups go to the eight XCDs round robin) takes a
  // contiguous range of row groups, whose alpha partner rows overlap in its L2
  int blk = blockIdx.x;
  const int nrb = FUSED ? nrowblk : (int)gridDim.x;
  if (xcd && blk < (nrb & ~7)) {  // the remainder past a multiple of 8 keeps its order
    const int q = nrb >> 3, x = blk & 7;
    blk = x * q + (blk >> 3);
  }
  const int row0 = blk * R;
  const int nr = min(R, nsa - row0);
  const int ntg = nr * nsb;
  const size_t base = (size_t)row0 * nsb;
  for (int i = tid; i < ntg; i += kRowThreads) {
    sPsi[i] = psi[base + i];
    sAcc[i] = 0.0;
  }
  for (int i = tid; i < nr * RA; i += kRowThreads) {
    const int r = i / RA, k = i - r * RA;
    sRA[i] = rta[(size_t)k * nsa + row0 + r];
  }
  const int ntiles = nr * tpr;
  const int nsteps = K8 / 4;
  for (int mg = 0; mg < mgroups; ++mg) {
    const int m0 = mg * W;
    __syncthreads();  // the previous m-group's reads of sA are done
    for (int i = tid; i < K8 * W; i += kRowThreads) {
      const int k = i / W, m = m0 + (i - k * W);
      sA[i] = (k < K4 && m < kpad) ? At[(size_t)k * kpad + m] : 0.0;
    }
    __syncthreads();
    for (int f = wave; f < ntiles; f += kRowWaves) {
      const int r = f / tpr, c = f - r * tpr;
      const int ib = c * 16 + li;
      double4_t acc[MT];
      // one m-group: the step count is 4 MT, or 4 MT - 2 where the contraction
      // ends in the first half of the last 16-row block (10 and 11 orbitals)
      if (K8 == 16 * MT)
        row_tile<MT, 4 * MT>(acc, sA, sPsi + (size_t)r * nsb, sRA + (size_t)r * RA, psi, rtb, nsb, ib, kq, li, nsteps);
      else if (K8 == 16 * MT - 8)
        row_tile<MT, 4 * MT - 2>(acc, sA, sPsi + (size_t)r * nsb, sRA + (size_t)r * RA, psi, rtb, nsb, ib, kq, li,
                                 nsteps);
      else
        row_tile<MT>(acc, sA, sPsi + (size_t)r * nsb, sRA + (size_t)r * RA, psi, rtb, nsb, ib, kq, li, nsteps);
      // partners and coefficients of all 4 MT outputs first (rows m >= npair
      // and columns past nsb read row 0 / the last column and add zero), then
      // the LDS adds, with no branch between the loads
      const uint32_t ibc = (uint32_t)(ib < nsb ? ib : nsb - 1);
      int2 tb[MT][4];
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int m = m0 + t * 16 + kq + 4 * q;
          tb[t][q] = rtb[(uint32_t)(m < npair ? m : 0) * (uint32_t)nsb + ibc];
        }
      double* acc_row = sAcc + (size_t)r * nsb;
      const bool col_ok = ib < nsb;
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int m = m0 + t * 16 + kq + 4 * q;
          const double c = (col_ok && m < npair) ? (double)tb[t][q].y : 0.0;
          atomicAdd(acc_row + tb[t][q].x, c * acc[t][q]);
        }
    }
  }
  __syncthreads();
  const size_t dim = (size_t)nsa * nsb;
  double e = 0.0;
  for (int tau = tid; tau < ntg; tau += kRowThreads) {
    double v = sAcc[tau];
    if (FUSED) {
      unsafeAtomicAdd(sigma + base + tau, v);
    } else {
      for (int z = 0; z < nsplit; ++z) v += sin[(size_t)z * dim + base + tau];
      sigma[base + tau] = v;
    }
    e += v * sPsi[tau];
  }
  for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
  if (lane == 0) red[wave] = e;
  __syncthreads();
  if (tid == 0) {
    double tsum = 0.0;
    for (int w = 0; w < kRowWaves; ++w) tsum += red[w];
    partial[blockIdx.x] = tsum;
  }
}

struct RowPlan {
  int mt, mgroups, R, ngroups, nsplit, kchunk;
  bool scatter, fused;
  int ngemm;  // fused: Ma C tile blocks after the row blocks
  size_t lds;
};

int row_k8(const Space& s) { return rup(s.npair + 1, 8); }  // contraction rows: two 4-deep k-steps per iteration

// Dynamic LDS of one block (bytes): staged form A slice | G tiles | psi rows |
// alpha operand rows; scatter form A slice | sigma rows | psi rows | alpha
// operand rows.
size_t row_lds(const Space& s, int mt, int R, bool scatter) {
  const int K8 = row_k8(s);
  const size_t W = (size_t)mt * 16;
  const size_t mid = scatter ? (size_t)R * s.nsb : (size_t)kRowWaves * W * 16;
  return 8 * ((size_t)K8 * W + mid + (size_t)R * s.nsb + (size_t)R * (K8 + 16));
}

int env_int(const char* name, int dflt) {
  const char* e = getenv(name);
  return (e && e[0]) ? atoi(e) : dflt;
}

// The space's kernel shape choices (sq_sigma_configure) override these
// defaults: rows per block, alpha GEMM split, and the deterministic (staged)
// form instead of the scatter form (default: SQ_SIGMA_DETERMINISTIC, else 0).
bool row_deterministic(const Space& s) {
  if (s.sig_det >= 0) return s.sig_det != 0;
  static int v = -1;
  if (v < 0) v = env_int("SQ_SIGMA_DETERMINISTIC", 0) != 0 ? 1 : 0;
  return v == 1;
}

// Kernel shape: the m-tile group (largest that fits, then balanced) and the
// rows per block R.  Staged form: R with the fewest block-rounds on 256 CUs
// (one block per CU).  Scatter form: the smallest R that gives every wave a
// tile (more, smaller blocks; two fit per CU at 12 orbitals).
bool row_plan(const Space& s, RowPlan* out) {
  if (s.dim == 0 || s.nsb > 16384 || s.npair + 1 > 255 || s.dim >= ((int64_t)1 << 31)) return false;
  const bool scatter = !row_deterministic(s);
  const int kpad = rup(s.npair + 1, 16);
  const int mt_total = kpad / 16;
  int mt = std::min(mt_total, 9);
  while (mt > 1 && row_lds(s, mt, 1, scatter) > kLdsMax) --mt;
  if (row_lds(s, mt, 1, scatter) > kLdsMax) return false;
  const int mgroups = (mt_total + mt - 1) / mt;
  mt = (mt_total + mgroups - 1) / mgroups;
  const int tpr = (s.nsb + 15) / 16;
  int bestR = 0;
  if (scatter) {
    bestR = std::max(1, std::min(s.nsa, (kRowWaves + tpr - 1) / tpr));
    while (bestR > 1 && row_lds(s, mt, bestR, true) > kLdsMax) --bestR;
  } else {
    int64_t best = -1;
    for (int R = 1; R <= 64 && R <= s.nsa; ++R) {
      if (row_lds(s, mt, R, false) > kLdsMax) break;
      if ((int64_t)R * s.nsb > (int64_t)kMaxTargets * kRowThreads) break;
      const int64_t groups = (s.nsa + R - 1) / R;
      const int64_t cost = ((groups + 255) / 256) * (((int64_t)R * tpr + kRowWaves - 1) / kRowWaves);
      if (best < 0 || cost < best) {
        best = cost;
        bestR = R;
      }
    }
  }
  if (bestR == 0) return false;
  if (s.sig_R > 0) {
    const int R = s.sig_R;
    if (R <= s.nsa && row_lds(s, mt, R, scatter) <= kLdsMax &&
        (scatter || (int64_t)R * s.nsb <= (int64_t)kMaxTargets * kRowThreads))
      bestR = R;
  }
  out->mt = mt;
  out->mgroups = mgroups;
  out->R = bestR;
  out->scatter = scatter;
  out->ngroups = (s.nsa + bestR - 1) / bestR;
  int nsplit = s.sig_nsplit > 0 ? s.sig_nsplit : (s.nsa >= 2048 ? 1 : 2);
  nsplit = std::max(1, std::min({nsplit, kMaxSplit, (s.nsa + 15) / 16}));
  out->kchunk = rup((s.nsa + nsplit - 1) / nsplit, 16);
  out->nsplit = (s.nsa + out->kchunk - 1) / out->kchunk;
  out->lds = row_lds(s, mt, bestR, scatter);
  // fused form: the Ma C tiles (128 x 64, contraction split nsplit ways) as
  // extra blocks of the scatter launch, where two row blocks share a CU (the
  // tile blocks then fill the slots the last row blocks leave: (12e,12o)
  // 258 + 54 -> 296 us); with one row block per CU (14 orbitals) the tile
  // blocks would run at that block's LDS footprint, so the GEMM stays separate
  static const int fuse_env = env_int("SQ_SIGMA_FUSED_GEMM", 1);
  out->fused = scatter && fuse_env != 0 && 2 * out->lds <= kLdsMax;
  out->ngemm = 0;
  if (out->fused) {
    if (s.sig_nsplit <= 0) {
      out->kchunk = rup((s.nsa + 3) / 4, 16);
      out->nsplit = (s.nsa + out->kchunk - 1) / out->kchunk;
    }
    out->ngemm = ((s.nsa + kGemmRows - 1) / kGemmRows) * ((s.nsb + kGemmCols - 1) / kGemmCols) * out->nsplit;
    const size_t glds = 8 * (size_t)16 * (kGemmRows + 1 + kGemmCols + 1);
    out->lds = std::max(out->lds, glds);
  }
  return true;
}

bool rows_enabled() {
  static int v = -1;
  if (v < 0) v = env_int("SQ_SIGMA_ROWS", 1) != 0 ? 1 : 0;
  return v == 1;
}

}  // namespace

bool row_sigma_applies(const Space& s) {
  RowPlan pl;
  return rows_enabled() && s.d_rtb != nullptr && row_plan(s, &pl);
}

// doubles of workspace: alpha GEMM partials [nsplit][dim] | block energy
// partials [ngroups], sized for the largest shape (nsplit = kMaxSplit,
// ngroups = nsa) | scalars [64] at the end.
int64_t row_sigma_work_doubles(const Space& s) {
  return (int64_t)kMaxSplit * s.dim + s.nsa + 64;
}

void build_row_tables(Space& s) {
  RowPlan pl;
  if (!row_plan(s, &pl)) return;
  const int K8 = row_k8(s);
  const int tpr = (s.nsb + 15) / 16;
  SQ_HIP(hipMalloc(&s.d_rta, sizeof(int2) * (size_t)(K8 + 16) * s.nsa));
  SQ_HIP(hipMalloc(&s.d_rtb, sizeof(int2) * (size_t)(K8 + 32) * s.nsb));
  k_row_tables<<<dim3((s.nsa + 255) / 256, K8 + 16), 256>>>(s.npair, s.nsa, s.nsb, 1, s.d_aexc, s.d_rta);
  SQ_HIP(hipGetLastError());
  k_row_tables<<<dim3((s.nsb + 255) / 256, K8 + 32), 256>>>(s.npair, s.nsb, s.nsb, 0, s.d_bexc, s.d_rtb);
  SQ_HIP(hipGetLastError());
  int32_t* cnt = nullptr;
  SQ_HIP(hipMalloc(&cnt, sizeof(int32_t) * s.nsb));
  k_nbr_count<<<(s.nsb + 255) / 256, 256>>>(s.npair, s.nsb, s.d_bexc, cnt);
  SQ_HIP(hipGetLastError());
  std::vector<int32_t> h(s.nsb), off(s.nsb + 1, 0);
  SQ_HIP(hipMemcpy(h.data(), cnt, sizeof(int32_t) * s.nsb, hipMemcpyDeviceToHost));
  for (int i = 0; i < s.nsb; ++i) off[i + 1] = off[i] + h[i];
  (void)hipFree(cnt);
  SQ_HIP(hipMalloc(&s.d_nbr_off, sizeof(int32_t) * (s.nsb + 1)));
  SQ_HIP(hipMemcpy(s.d_nbr_off, off.data(), sizeof(int32_t) * (s.nsb + 1), hipMemcpyHostToDevice));
  SQ_HIP(hipMalloc(&s.d_nbr, sizeof(uint32_t) * std::max(off[s.nsb], 1)));
  SQ_HIP(hipMalloc(&s.d_nbr_tile, (size_t)s.nsb * (tpr + 1)));
  k_nbr_fill<<<(s.nsb + 255) / 256, 256>>>(s.npair, s.nsb, tpr, s.d_bexc, s.d_nbr_off, s.d_nbr, s.d_nbr_tile);
  SQ_HIP(hipGetLastError());
  SQ_HIP(hipDeviceSynchronize());
}

void free_row_tables(Space& s) {
  for (void* p : {(void*)s.d_rta, (void*)s.d_rtb, (void*)s.d_nbr, (void*)s.d_nbr_off, (void*)s.d_nbr_tile})
    if (p) (void)hipFree(p);
  s.d_rta = s.d_rtb = nullptr;
  s.d_nbr = nullptr;
  s.d_nbr_off = nullptr;
  s.d_nbr_tile = nullptr;
}

// Ma (+ e_core on the diagonal) for a Hamiltonian whose coefficient matrix
// is already on the device.
void build_alpha_ham(const Space& s, Hamiltonian& H) {
  if (s.d_rtb == nullptr || s.nsa == 0) return;
  SQ_HIP(hipMalloc(&H.d_Ma, sizeof(double) * (size_t)s.nsa * s.nsa));
  const int nslot_max = 1 + s.na * (s.n - s.na);
  const size_t lds = 2 * (size_t)rup(s.nsa, 8) + 2 * (size_t)rup(s.npair, 8) + 4 * (size_t)rup(s.npair, 2) +
                     8 * (size_t)s.npair * nslot_max;
  SQ_CHECK(lds <= kLdsMax, SQ_ERR_ARG, "alpha Hamiltonian block does not fit in LDS");
  allow_lds(k_alpha_ham, lds);
  k_alpha_ham<<<s.nsa, 256, lds>>>(s.nsa, s.npair, H.kpad, s.d_aexc, H.d_A, H.e_core, H.d_Ma);
  SQ_HIP(hipGetLastError());
  SQ_HIP(hipDeviceSynchronize());
}

namespace {

// rocBLAS handle of the current device, created on first use and kept until
// the process ends (creation allocates the library's workspace).
rocblas_handle blas_handle() {
  static std::mutex mu;
  static std::map<int, rocblas_handle> handles;
  int dev = 0;
  SQ_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = handles.find(dev);
  if (it != handles.end()) return it->second;
  rocblas_handle h = nullptr;
  SQ_CHECK(rocblas_create_handle(&h) == rocblas_status_success, SQ_ERR_HIP, "rocblas_create_handle failed");
  handles[dev] = h;
  return h;
}

}  // namespace

// sigma = H psi and the block partials of <psi|sigma> (*partial_out, the
// count is returned).  Where the alpha-string product S = Ma C runs as its
// own launch (spaces whose row blocks fill a CU's LDS), it is a plain dense
// GEMM and goes to rocBLAS unless allow_library is false (the no-allocation
// entry point) or SQ_ALPHA_GEMM_LIB=0; the library reaches 66.5 TFLOP/s on
// the (14e,14o) shape (3432^3) against 48 for k_alpha_gemm.
int64_t row_sigma(const Space& s, const Hamiltonian& H, const double* psi, double* sigma, double* work,
                  double** partial_out, hipStream_t st, bool allow_library) {
  RowPlan pl;
  SQ_CHECK(row_plan(s, &pl) && H.d_Ma, SQ_ERR_ARG, "row sigma does not apply to this space");
  SQ_CHECK(pl.nsplit <= kMaxSplit && pl.ngroups <= s.nsa, SQ_ERR_ARG, "row sigma shape exceeds its workspace");
  SQ_CHECK(!pl.fused || pl.ngroups + pl.ngemm <= row_sigma_work_doubles(s) - 64, SQ_ERR_ARG,
           "row sigma shape exceeds its workspace");
  double* sinp = work;
  double* partial = work + (size_t)pl.nsplit * s.dim;
  *partial_out = partial;
  int nsplit = pl.nsplit;  // partial products of S the row kernels add
  if (!pl.fused) {
    static const int lib_env = env_int("SQ_ALPHA_GEMM_LIB", 1);
    const double flops = 2.0 * (double)s.nsa * s.nsa * s.nsb;
    if (allow_library && lib_env != 0) {
      // row-major S = Ma C is the column-major product C' Ma' of the same
      // buffers (C' the nsb x nsa view of psi, Ma' that of Ma)
      rocblas_handle h = blas_handle();
      SQ_CHECK(rocblas_set_stream(h, st) == rocblas_status_success, SQ_ERR_HIP, "rocblas_set_stream failed");
      const double one = 1.0, zero = 0.0;
      rocblas_status rs = rocblas_status_success;
      const double bytes = 8.0 * ((double)s.nsa * s.nsa + 2.0 * (double)s.dim);
      SQ_PROF(PC_ALPHA_GEMM, st, bytes, flops,
              rs = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, s.nsb, s.nsa, s.nsa, &one, psi,
                                 s.nsb, H.d_Ma, s.nsa, &zero, sinp, s.nsb));
      SQ_CHECK(rs == rocblas_status_success, SQ_ERR_HIP, "rocblas_dgemm failed");
      nsplit = 1;
    } else {
      dim3 grid((s.nsb + 63) / 64, (s.nsa + 63) / 64, pl.nsplit);
      const double bytes = 8.0 * ((double)s.nsa * s.nsa + (double)s.dim * (1 + pl.nsplit));
      SQ_PROF(PC_ALPHA_GEMM, st, bytes, flops,
              k_alpha_gemm<<<grid, 256, 0, st>>>(H.d_Ma, psi, sinp, s.nsa, s.nsb, pl.kchunk));
    }
  }
  const int K8 = row_k8(s);
  const int K4 = rup(s.npair + 1, 4);
  const double flops = 2.0 * (double)s.npair * (s.npair + 1) * (double)s.dim;
  // algorithmic traffic: psi read (rows and alpha partners through the cache), the alpha partials read, sigma written
  const double bytes = 8.0 * (double)s.dim * (2 + nsplit);
  if (pl.scatter && pl.fused) {
    static const int xcd = env_int("SQ_ROWS_XCD", 1) != 0 ? 1 : 0;
    k_zero<<<(unsigned)std::min<int64_t>((s.dim + 255) / 256, 8192), 256, 0, st>>>(sigma, s.dim);
    SQ_HIP(hipGetLastError());
    double* fpart = work;
    *partial_out = fpart;
    const double fl = flops + 2.0 * (double)s.nsa * s.nsa * s.nsb;
    const int nblk = pl.ngroups + pl.ngemm;
    switch (pl.mt) {
#define SQ_RSF(MT)                                                                                              \
  case MT:                                                                                                      \
    allow_lds(k_sigma_rows_scatter<MT, true>, pl.lds);                                                          \
    SQ_PROF(PC_SIGMA_ROWS, st, bytes, fl,                                                                       \
            k_sigma_rows_scatter<MT, true><<<nblk, kRowThreads, pl.lds, st>>>(                                  \
                H.d_A, H.kpad, K4, K8, pl.mgroups, s.npair, psi, nullptr, 0, sigma, fpart, s.d_rta, s.d_rtb,    \
                s.nsa, s.nsb, pl.R, xcd, pl.ngroups, H.d_Ma, pl.kchunk, 0, 0, 0, 0));                           \
    break;
      SQ_RSF(1) SQ_RSF(2) SQ_RSF(3) SQ_RSF(4) SQ_RSF(5) SQ_RSF(6) SQ_RSF(7) SQ_RSF(8) SQ_RSF(9)
#undef SQ_RSF
      default:
        throw Error{SQ_ERR_ARG, "unsupported orbital count for the row sigma"};
    }
    return nblk;
  }
  if (pl.scatter) {
    static const int xcd = env_int("SQ_ROWS_XCD", 1) != 0 ? 1 : 0;
    switch (pl.mt) {
#define SQ_RSS(MT)                                                                                              \
  case MT:                                                                                                      \
    allow_lds(k_sigma_rows_scatter<MT, false>, pl.lds);                                                         \
    SQ_PROF(PC_SIGMA_ROWS, st, bytes, flops,                                                                    \
            k_sigma_rows_scatter<MT, false><<<pl.ngroups, kRowThreads, pl.lds, st>>>(                           \
                H.d_A, H.kpad, K4, K8, pl.mgroups, s.npair, psi, sinp, nsplit, sigma, partial, s.d_rta,         \
                s.d_rtb, s.nsa, s.nsb, pl.R, xcd, pl.ngroups, nullptr, 0, 0, 0, 0, 0));                         \
    break;
      SQ_RSS(1) SQ_RSS(2) SQ_RSS(3) SQ_RSS(4) SQ_RSS(5) SQ_RSS(6) SQ_RSS(7) SQ_RSS(8) SQ_RSS(9)
#undef SQ_RSS
      default:
        throw Error{SQ_ERR_ARG, "unsupported orbital count for the row sigma"};
    }
    return pl.ngroups;
  }
  switch (pl.mt) {
#define SQ_RS(MT)                                                                                               \
  case MT:                                                                                                      \
    allow_lds(k_sigma_rows<MT>, pl.lds);                                                                        \
    SQ_PROF(PC_SIGMA_ROWS, st, bytes, flops,                                                                    \
            k_sigma_rows<MT><<<pl.ngroups, kRowThreads, pl.lds, st>>>(                                          \
                H.d_A, H.kpad, K4, K8, pl.mgroups, psi, sinp, nsplit, sigma, partial, s.d_rta, s.d_rtb,         \
                s.d_nbr, s.d_nbr_off, s.d_nbr_tile, s.nsa, s.nsb, pl.R));                                       \
    break;
    SQ_RS(1) SQ_RS(2) SQ_RS(3) SQ_RS(4) SQ_RS(5) SQ_RS(6) SQ_RS(7) SQ_RS(8) SQ_RS(9)
#undef SQ_RS
    default:
      throw Error{SQ_ERR_ARG, "unsupported orbital count for the row sigma"};
  }
  return pl.ngroups;
}

// Doubles of workspace the row sigma uses at its chosen shape (alpha GEMM
// partials and block partials), or -1 when it does not apply.
int64_t row_sigma_used_doubles(const Space& s) {
  RowPlan pl;
  if (!row_plan(s, &pl)) return -1;
  if (pl.fused) return (int64_t)pl.ngroups + pl.ngemm;
  return (int64_t)pl.nsplit * s.dim + pl.ngroups;
}

bool row_sigma_group_ok(const Space& s) {
  RowPlan pl;
  return rows_enabled() && s.d_rtb != nullptr && row_plan(s, &pl) && pl.scatter && pl.fused;
}

int64_t row_sigma_group_partials(const Space& s) {
  RowPlan pl;
  if (!row_plan(s, &pl) || !pl.fused) return 0;
  return (int64_t)pl.ngroups + pl.ngemm;
}

void row_sigma_group(const Space& s, const double* A, int64_t A_stride, const double* Ma, int64_t Ma_stride,
                     const double* psi, double* sigma, double* partial, int n, hipStream_t st) {
  RowPlan pl;
  SQ_CHECK(row_sigma_group_ok(s) && row_plan(s, &pl) && n > 0 && n <= 65535, SQ_ERR_ARG,
           "the grouped row sigma does not apply to this space");
  static const int xcd = env_int("SQ_ROWS_XCD", 1) != 0 ? 1 : 0;
  const int K8 = row_k8(s);
  const int K4 = rup(s.npair + 1, 4);
  const int nblk = pl.ngroups + pl.ngemm;
  const int64_t total = (int64_t)n * s.dim;
  k_zero<<<(unsigned)std::min<int64_t>((total + 255) / 256, 16384), 256, 0, st>>>(sigma, total);
  SQ_HIP(hipGetLastError());
  const double fl = (2.0 * (double)s.npair * (s.npair + 1) * (double)s.dim + 2.0 * (double)s.nsa * s.nsa * s.nsb) * n;
  const double bytes = 8.0 * (double)s.dim * 3 * n;
  const int kpad = rup(s.npair + 1, 16);
  switch (pl.mt) {
#define SQ_RSG(MT)                                                                                              \
  case MT:                                                                                                      \
    allow_lds(k_sigma_rows_scatter<MT, true>, pl.lds);                                                          \
    SQ_PROF(PC_SIGMA_ROWS, st, bytes, fl,                                                                       \
            (k_sigma_rows_scatter<MT, true><<<dim3(nblk, n), kRowThreads, pl.lds, st>>>(                        \
                A, kpad, K4, K8, pl.mgroups, s.npair, psi, nullptr, 0, sigma, partial, s.d_rta, s.d_rtb, s.nsa, \
                s.nsb, pl.R, xcd, pl.ngroups, Ma, pl.kchunk, A_stride, Ma_stride, s.dim, nblk)));               \
    break;
    SQ_RSG(1) SQ_RSG(2) SQ_RSG(3) SQ_RSG(4) SQ_RSG(5) SQ_RSG(6) SQ_RSG(7) SQ_RSG(8) SQ_RSG(9)
#undef SQ_RSG
    default:
      throw Error{SQ_ERR_ARG, "unsupported orbital count for the row sigma"};
  }
}

}  // namespace sq

extern "C" int sq_sigma_configure(sq_space* sp, int rows_per_block, int nsplit, int deterministic) {
  return sq::guarded([&] {
    SQ_CHECK(sp, SQ_ERR_ARG, "NULL space");
    SQ_CHECK(rows_per_block >= 0 && rows_per_block <= 64 && nsplit >= 0 && nsplit <= 4, SQ_ERR_ARG,
             "rows_per_block must be in [0, 64] and nsplit in [0, 4] (0: automatic)");
    sq::Space& s = sp->s;
    sq::RowPlan pl;
    const bool before = sq::row_plan(s, &pl);
    const int old[3] = {s.sig_R, s.sig_nsplit, s.sig_det};
    s.sig_R = rows_per_block;
    s.sig_nsplit = nsplit;
    s.sig_det = deterministic < 0 ? -1 : (deterministic != 0);
    if (before && !sq::row_plan(s, &pl)) {  // the workspace sizes assume the row path
      s.sig_R = old[0];
      s.sig_nsplit = old[1];
      s.sig_det = old[2];
      throw sq::Error{SQ_ERR_ARG, "this sigma form does not fit the space"};
    }
  });
}
