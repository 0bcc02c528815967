// sigma = H psi, energies and RDMs (replaces propagate_state([H]) /
// expectation_value, operator_state_algebra.py:472-824, and the rdm1/rdm2
// expectation-value loops of ups_wavefunction.py:409-476).
//
// Packed pair index k = p(p+1)/2 + q (p >= q), E+_k = E_pq + E_qp (p > q),
// E+_pp = E_pp.  With h'_pq = h_pq - 1/2 sum_r g_prrq and the 8-fold symmetry
// of g:
//     sigma = e_core psi + sum_k E+_k G'_k,
//     G'_k  = sum_{k'} A_{k k'} D_{k'},   D_{k'} = E+_{k'} psi (k' < npair),
//             D_npair = psi,  A = [ 1/2 g(k,k') | h'_k ].
// Kernel 1 (bandwidth): D build by single-excitation gathers.
// Kernel 2 (FP64 MFMA, v_mfma_f64_16x16x4f64): G' = A . D.
// Kernel 3 (bandwidth): sigma gather + <psi|sigma> partials.

#include <cmath>
#include <cstdlib>
#include <cstring>

#include "sq_internal.h"

namespace sq {

void copy_vec(const double* in, double* out, int64_t n, hipStream_t st);
void reduce_impl(const double* partial, int64_t n, double* out_dev, int slot, bool accumulate, double factor,
                 hipStream_t st);

constexpr int kB = 256;

static inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

struct SigmaDims {
  int npair, K4, kpad;
  int64_t dim;
};

static SigmaDims sigma_dims(const Space& s) {
  SigmaDims d;
  d.npair = s.npair;
  d.K4 = round_up(s.npair + 1, 4);
  d.kpad = round_up(s.npair + 1, 16);
  d.dim = s.dim;
  return d;
}

// Work layout (doubles): D [K4*dim] | G [npair*dim] | partials [nsa*ceil(nsb/256)] | scalars [8];
// the row path (sq_rowsigma.hip) needs only its alpha GEMM partials and block partials.
int64_t sigma_work_bytes(const Space& s) {
  if (row_sigma_applies(s)) return 8 * row_sigma_work_doubles(s);
  SigmaDims d = sigma_dims(s);
  int64_t nb = (int64_t)s.nsa * ((s.nsb + kB - 1) / kB);
  return 8 * ((int64_t)d.K4 * d.dim + (int64_t)d.npair * d.dim + nb + 64);
}

// ---------------------------------------------------------------------------
// Kernel 1: D[k][I] = (E+_k psi)[I]; row npair = psi; rows npair+1..K4-1 = 0.
// Excitation tables are k-major: exc[k*nstr + s] = sign*(target+1), or the
// occupation of orbital p for diagonal k.
__global__ void k_dbuild(int n, int nsa, int nsb, int npair, int K4, const int32_t* __restrict__ aexc,
                         const int32_t* __restrict__ bexc, const double* __restrict__ psi, double* __restrict__ D) {
  int ia = blockIdx.y;
  int ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= nsb) return;
  const size_t dim = (size_t)nsa * nsb;
  const size_t I = (size_t)ia * nsb + ib;
  const double* row = psi + (size_t)ia * nsb;
  const double pI = psi[I];
  int k = 0;
  for (int p = 0; p < n; ++p) {
    for (int q = 0; q <= p; ++q, ++k) {
      int32_t ea = aexc[(size_t)k * nsa + ia];
      int32_t eb = bexc[(size_t)k * nsb + ib];
      double v;
      if (p == q) {
        v = (double)(ea + eb) * pI;
      } else {
        v = 0.0;
        if (ea) v += (ea > 0 ? 1.0 : -1.0) * psi[(size_t)(abs(ea) - 1) * nsb + ib];
        if (eb) v += (eb > 0 ? 1.0 : -1.0) * row[abs(eb) - 1];
      }
      D[(size_t)k * dim + I] = v;
    }
  }
  D[(size_t)npair * dim + I] = pI;
  for (int r = npair + 1; r < K4; ++r) D[(size_t)r * dim + I] = 0.0;
}

// Kernel 3: sigma[I] = e_core psi[I] + sum_k (E+_k G_k)[I]; partial <psi|sigma>.
__global__ void k_sigma_gather(int n, int nsa, int nsb, const int32_t* __restrict__ aexc,
                               const int32_t* __restrict__ bexc, const double* __restrict__ G,
                               const double* __restrict__ psi, double e_core, double* __restrict__ sigma,
                               double* __restrict__ partial) {
  __shared__ double red[16];  // up to 1024 threads
  int ia = blockIdx.y;
  int ib = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t dim = (size_t)nsa * nsb;
  double e = 0.0;
  if (ib < nsb) {
    const size_t I = (size_t)ia * nsb + ib;
    double v = e_core * psi[I];
    int k = 0;
    for (int p = 0; p < n; ++p) {
      for (int q = 0; q <= p; ++q, ++k) {
        const double* Gk = G + (size_t)k * dim;
        int32_t ea = aexc[(size_t)k * nsa + ia];
        int32_t eb = bexc[(size_t)k * nsb + ib];
        if (p == q) {
          v += (double)(ea + eb) * Gk[I];
        } else {
          if (ea) v += (ea > 0 ? 1.0 : -1.0) * Gk[(size_t)(abs(ea) - 1) * nsb + ib];
          if (eb) v += (eb > 0 ? 1.0 : -1.0) * Gk[(size_t)ia * nsb + abs(eb) - 1];
        }
      }
    }
    sigma[I] = v;
    e = v * psi[I];
  }
  for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
  }
}

// Kernel 2: G[m][c] = sum_k At[k][m] D[k][c] for m < npair, k < K4, with
// v_mfma_f64_16x16x4f64.  Lane l supplies A[l%16][l/16] and B[l/16][l%16];
// accumulator register r of lane l holds C[(l/16) + 4r][l%16] (measured on
// gfx950 against a host GEMM).  One wave owns 16 columns and MT
// row tiles starting at row m0 = blockIdx.y * MT * 16; the matching column
// slice of At (K4 x 16*MT) is staged in LDS and shared by the block.
template <int MT>
__global__ __launch_bounds__(512) void k_gemm_mfma(const double* __restrict__ At, const double* __restrict__ D,
                                                   double* __restrict__ G, int npair, int K4, int kpad,
                                                   int64_t dim) {
  extern __shared__ double sA[];
  constexpr int W = MT * 16;
  const int m0 = blockIdx.y * W;
  for (int i = threadIdx.x; i < K4 * W; i += blockDim.x) {
    int k = i / W, m = m0 + i % W;
    sA[i] = m < kpad ? At[(size_t)k * kpad + m] : 0.0;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t c = ((int64_t)blockIdx.x * (blockDim.x >> 6) + wave) * 16 + (lane & 15);
  const int kq = lane >> 4;
  const bool valid = c < dim;
  typedef double double4_t __attribute__((ext_vector_type(4)));
  double4_t acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) acc[t] = double4_t{0.0, 0.0, 0.0, 0.0};
  const double* Dp = D + (size_t)kq * dim + (valid ? c : 0);
  double bnext = valid ? Dp[0] : 0.0;
  for (int k4 = 0; k4 < K4; k4 += 4) {
    double b = bnext;
    if (k4 + 4 < K4) bnext = valid ? Dp[(size_t)(k4 + 4) * dim] : 0.0;
    const double* arow = sA + (size_t)(k4 + kq) * W + (lane & 15);
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      double a = arow[t * 16];
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
    }
  }
  if (!valid) return;
#pragma unroll
  for (int t = 0; t < MT; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int m = m0 + t * 16 + kq + 4 * r;
      if (m < npair) G[(size_t)m * dim + c] = acc[t][r];
    }
  }
}

// Kernel 2' (fused): the same contraction with the B operand D[k][c]
// generated in registers from psi by the single-excitation gathers of
// k_dbuild, so D never touches memory.  Lane l computes D[k4 + l/16][c] for
// its column c; row npair is psi itself, rows above it are zero.
__constant__ uint8_t c_kp[160], c_kq[160];

// Fused G = A D with D generated in registers.  Row k of D reads one
// amplitude per spin part: D[k][ia, ib] = ca * psi[ra + ib] + cb * psi[ia*nsb + cb_off]
// with {ra, ca} = gta[k][ia] (alpha partner row offset, sign / occupation)
// and {cb_off, cb} = gtb[k][ib]; absent parts carry coefficient 0 and point
// at the amplitude itself, so the loop has no branches.  The k-steps are
// unrolled by two with ping-pong registers: step s issues the amplitude
// loads of step s+1 and the table loads of step s+3 before its MFMAs, and
// nothing loaded moves between registers.  Steps run to K8 = K4 rounded up
// to 8 (At slice rows K4..K8 are zero in LDS).
template <int MT>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(MT <= 5 ? 6 : 4, 8)))
void k_gemm_fused(const double* __restrict__ At, const double* __restrict__ psi,
                                                    const int2* __restrict__ gta, const int2* __restrict__ gtb,
                                                    double* __restrict__ G, int npair, int K4, int kpad, int nsa,
                                                    int nsb) {
  extern __shared__ double sA[];
  constexpr int W = MT * 16;
  const int K8 = (K4 + 7) & ~7;
  const int m0 = blockIdx.y * W;
  const int64_t dim = (int64_t)nsa * nsb;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t c = ((int64_t)blockIdx.x * (blockDim.x >> 6) + wave) * 16 + (lane & 15);
  const int kq = lane >> 4;
  const bool valid = c < dim;
  const int64_t cc = valid ? c : dim - 1;  // columns past the end compute garbage that is never stored
  const int ia = (int)(cc / nsb);
  const int ib = (int)(cc - (int64_t)ia * nsb);
  // 32-bit element indices over uniform bases (dim < 2^29 keeps byte offsets in 32 bits)
  const uint32_t ta_i = (uint32_t)(kq * nsa + ia), tb_i = (uint32_t)(kq * nsb + ib);
  const uint32_t sa = 4u * (uint32_t)nsa, sb = 4u * (uint32_t)nsb;
  const uint32_t row = (uint32_t)ia * (uint32_t)nsb, col = (uint32_t)ib;
  typedef double double4_t __attribute__((ext_vector_type(4)));
  double4_t acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) acc[t] = double4_t{0.0, 0.0, 0.0, 0.0};
  // slot 0: even steps, slot 1: odd steps
  int2 ta0 = gta[ta_i], tb0 = gtb[tb_i];
  int2 ta1 = gta[ta_i + sa], tb1 = gtb[tb_i + sb];
  double xa0 = psi[(uint32_t)ta0.x + col], xb0 = psi[row + (uint32_t)tb0.x];
  int ca0 = ta0.y, cb0 = tb0.y;
  double xa1, xb1;
  int ca1, cb1;
  ta0 = gta[ta_i + 2 * sa];
  tb0 = gtb[tb_i + 2 * sb];
  // the At slice goes to LDS while the operand prologue loads are in flight
  for (int i = threadIdx.x; i < K8 * W; i += blockDim.x) {
    int k = i / W, m = m0 + i % W;
    sA[i] = (m < kpad && k < K4) ? At[(size_t)k * kpad + m] : 0.0;
  }
  __syncthreads();
  const double* arow = sA + (size_t)kq * W + (lane & 15);
  const int nsteps = K8 / 4;
  for (int s = 0; s < nsteps; s += 2) {
    // step s
    xa1 = psi[(uint32_t)ta1.x + col];
    xb1 = psi[row + (uint32_t)tb1.x];
    ca1 = ta1.y;
    cb1 = tb1.y;
    ta1 = gta[ta_i + (uint32_t)(s + 3) * sa];
    tb1 = gtb[tb_i + (uint32_t)(s + 3) * sb];
    __asm__ volatile("" ::: "memory");  // keep the loads ahead of this step's MFMAs
    {
      const double b = (double)ca0 * xa0 + (double)cb0 * xb0;
#pragma unroll
      for (int t = 0; t < MT; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(arow[t * 16], b, acc[t], 0, 0, 0);
    }
    arow += 4 * W;
    // step s + 1
    xa0 = psi[(uint32_t)ta0.x + col];
    xb0 = psi[row + (uint32_t)tb0.x];
    ca0 = ta0.y;
    cb0 = tb0.y;
    ta0 = gta[ta_i + (uint32_t)(s + 4) * sa];
    tb0 = gtb[tb_i + (uint32_t)(s + 4) * sb];
    __asm__ volatile("" ::: "memory");
    {
      const double b = (double)ca1 * xa1 + (double)cb1 * xb1;
#pragma unroll
      for (int t = 0; t < MT; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(arow[t * 16], b, acc[t], 0, 0, 0);
    }
    arow += 4 * W;
  }
  if (!valid) return;
#pragma unroll
  for (int t = 0; t < MT; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int m = m0 + t * 16 + kq + 4 * r;
      if (m < npair) G[(size_t)m * dim + c] = acc[t][r];
    }
  }
}

// GEMM operand tables (rows k < K4 + 16; the rows past K4 feed the steps
// up to K8 and the prefetch past the end): alpha strings {partner row * nsb, coefficient}, beta strings
// {partner column, coefficient}.  Diagonal pairs (p, p) read the amplitude
// itself with the occupation of p; row npair is psi (alpha part, 1); absent
// parts and padding rows read the amplitude itself with coefficient 0.
__global__ void k_gemm_tables(int npair, int nstr, int nsb, int alpha, const int32_t* __restrict__ exc,
                              int2* __restrict__ tab) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (s >= nstr) return;
  const int own = alpha ? s * nsb : s;
  int2 e = make_int2(own, 0);
  if (k < npair) {
    int p = 0;
    while ((p + 1) * (p + 2) / 2 <= k) ++p;
    const int q = k - p * (p + 1) / 2;
    const int32_t x = exc[(size_t)k * nstr + s];
    if (p == q) e = make_int2(own, x);
    else if (x != 0) e = make_int2(alpha ? (abs(x) - 1) * nsb : abs(x) - 1, x > 0 ? 1 : -1);
  } else if (k == npair && alpha) {
    e = make_int2(own, 1);
  }
  tab[(size_t)k * nstr + s] = e;
}

void build_gemm_tables(Space& s) {
  const SigmaDims d = sigma_dims(s);
  const int rows = d.K4 + 16;
  SQ_HIP(hipMalloc(&s.d_gta, sizeof(int2) * (size_t)rows * std::max(s.nsa, 1)));
  SQ_HIP(hipMalloc(&s.d_gtb, sizeof(int2) * (size_t)rows * std::max(s.nsb, 1)));
  if (s.nsa > 0 && s.nsb > 0) {
    k_gemm_tables<<<dim3((s.nsa + 255) / 256, rows), 256>>>(s.npair, s.nsa, s.nsb, 1, s.d_aexc, s.d_gta);
    SQ_HIP(hipGetLastError());
    k_gemm_tables<<<dim3((s.nsb + 255) / 256, rows), 256>>>(s.npair, s.nsb, s.nsb, 0, s.d_bexc, s.d_gtb);
    SQ_HIP(hipGetLastError());
  }
  SQ_HIP(hipDeviceSynchronize());
}

static void upload_pair_tables(int n) {
  static int done_for = -1;
  if (done_for == n) return;
  uint8_t kp[160] = {0}, kq[160] = {0};
  for (int p = 0, k = 0; p < n; ++p)
    for (int q = 0; q <= p; ++q, ++k) kp[k] = (uint8_t)p, kq[k] = (uint8_t)q;
  SQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_kp), kp, sizeof(kp)));
  SQ_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_kq), kq, sizeof(kq)));
  done_for = n;
}

static void launch_gemm_fused(const Space& s, const SigmaDims& d, const double* At, const double* psi, double* G,
                              hipStream_t st) {
  upload_pair_tables(s.n);
  const int threads = 512;
  const int64_t cols_per_block = (threads / 64) * 16;
  const int mt_total = d.kpad / 16;
  const int K8 = (d.K4 + 7) & ~7;
  int mt = mt_total;
  while (mt > 1 && (size_t)K8 * mt * 16 * sizeof(double) > 65536) --mt;
  const int groups = (mt_total + mt - 1) / mt;
  dim3 grid((unsigned)((d.dim + cols_per_block - 1) / cols_per_block), groups);
  size_t lds = sizeof(double) * (size_t)K8 * mt * 16;
  // algorithmic traffic: psi read once per m-group (+ gathers hitting cache), G written
  const double gbytes = 8.0 * (double)d.dim * (groups + d.npair);
  const double gflops = 2.0 * (double)d.npair * (d.npair + 1) * (double)d.dim;
  switch (mt) {
#define SQ_GF(MT)                                                                                          \
  case MT:                                                                                                 \
    SQ_PROF(PC_GEMM, st, gbytes, gflops,                                                                   \
            k_gemm_fused<MT><<<grid, threads, lds, st>>>(At, psi, s.d_gta, s.d_gtb, G, d.npair, d.K4,      \
                                                         d.kpad, s.nsa, s.nsb));                           \
    break;
    SQ_GF(1) SQ_GF(2) SQ_GF(3) SQ_GF(4) SQ_GF(5) SQ_GF(6) SQ_GF(7) SQ_GF(8) SQ_GF(9)
#undef SQ_GF
    default:
      throw Error{SQ_ERR_ARG, "unsupported orbital count for the sigma GEMM"};
  }
}

static void launch_gemm(const SigmaDims& d, const double* At, const double* D, double* G, hipStream_t st) {
  const int threads = 512;
  const int64_t cols_per_block = (threads / 64) * 16;
  const int mt_total = d.kpad / 16;
  // largest tile group whose At slice fits in 64 KiB of LDS
  int mt = mt_total;
  while (mt > 1 && (size_t)d.K4 * mt * 16 * sizeof(double) > 65536) --mt;
  const int groups = (mt_total + mt - 1) / mt;
  dim3 grid((unsigned)((d.dim + cols_per_block - 1) / cols_per_block), groups);
  size_t lds = sizeof(double) * (size_t)d.K4 * mt * 16;
  const double gbytes = 8.0 * (double)d.dim * (d.K4 + d.npair);
  const double gflops = 2.0 * (double)d.npair * (d.npair + 1) * (double)d.dim;
  switch (mt) {
#define SQ_G(MT)                                                                              \
  case MT:                                                                                    \
    SQ_PROF(PC_GEMM, st, gbytes, gflops,                                                      \
            k_gemm_mfma<MT><<<grid, threads, lds, st>>>(At, D, G, d.npair, d.K4, d.kpad, d.dim)); \
    break;
    SQ_G(1) SQ_G(2) SQ_G(3) SQ_G(4) SQ_G(5) SQ_G(6) SQ_G(7) SQ_G(8) SQ_G(9)
#undef SQ_G
    default:
      throw Error{SQ_ERR_ARG, "unsupported orbital count for the sigma GEMM"};
  }
}

// The sigma gather walks whole beta rows per block (up to 1024 threads), so a
// block reads each row of every G_k once.
static int gather_threads(const Space& s) { return (int)std::min<int64_t>(1024, (s.nsb + 63) / 64 * 64); }

static bool fused_sigma() {
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("SQ_SIGMA_FUSED");
    v = (e && e[0] == '0') ? 0 : 1;
  }
  return v == 1;
}

void sigma_impl(const Space& s, const Hamiltonian& H, const double* psi, double* sigma, double* energy_host,
                double* work, hipStream_t st) {
  SigmaDims d = sigma_dims(s);
  if (s.dim == 0) return;
  if (H.d_Ma && row_sigma_applies(s)) {
    double* partial = nullptr;
    const int64_t nb = row_sigma(s, H, psi, sigma, work, &partial, st);
    if (energy_host) {
      double* scal = work + (row_sigma_work_doubles(s) - 64);
      reduce_impl(partial, nb, scal, 0, false, 1.0, st);
      SQ_HIP(hipMemcpyAsync(energy_host, scal, sizeof(double), hipMemcpyDeviceToHost, st));
      SQ_HIP(hipStreamSynchronize(st));
    }
    return;
  }
  const int32_t* ae = s.d_aexc;
  const int32_t* be = s.d_bexc;
  double* D = work;
  double* G = D + (size_t)d.K4 * d.dim;
  double* partial = G + (size_t)d.npair * d.dim;
  int64_t nb = (int64_t)s.nsa * ((s.nsb + kB - 1) / kB);
  double* scal = partial + nb;
  dim3 grid((s.nsb + kB - 1) / kB, s.nsa);
  const double d8 = 8.0 * (double)d.dim;
  const int gth = gather_threads(s);
  const dim3 ggrid((s.nsb + gth - 1) / gth, s.nsa);
  const int64_t gnb = (int64_t)ggrid.x * ggrid.y;
  if (fused_sigma()) {
    launch_gemm_fused(s, d, H.d_A, psi, G, st);
  } else {
    SQ_PROF(PC_DBUILD, st, d8 * (d.K4 + 1), 0.0,
            k_dbuild<<<grid, kB, 0, st>>>(s.n, s.nsa, s.nsb, d.npair, d.K4, ae, be, psi, D));
    launch_gemm(d, H.d_A, D, G, st);
  }
  SQ_PROF(PC_SIGMA_GATHER, st, d8 * (d.npair + 2), 0.0,
          k_sigma_gather<<<ggrid, gth, 0, st>>>(s.n, s.nsa, s.nsb, ae, be, G, psi, H.e_core, sigma, partial));
  if (energy_host) {
    reduce_impl(partial, gnb, scal, 0, false, 1.0, st);
    SQ_HIP(hipMemcpyAsync(energy_host, scal, sizeof(double), hipMemcpyDeviceToHost, st));
    SQ_HIP(hipStreamSynchronize(st));
  }
}

// Device-resident energy: same as sigma_impl but leaves <psi|sigma> in e_dev[0].
void sigma_energy_dev(const Space& s, const Hamiltonian& H, const double* psi, double* sigma, double* e_dev,
                      double* work, hipStream_t st) {
  SigmaDims d = sigma_dims(s);
  if (H.d_Ma && row_sigma_applies(s)) {
    double* partial = nullptr;
    const int64_t nb = row_sigma(s, H, psi, sigma, work, &partial, st);
    reduce_impl(partial, nb, e_dev, 0, false, 1.0, st);
    return;
  }
  const int32_t* ae = s.d_aexc;
  const int32_t* be = s.d_bexc;
  double* D = work;
  double* G = D + (size_t)d.K4 * d.dim;
  double* partial = G + (size_t)d.npair * d.dim;
  int64_t nb = (int64_t)s.nsa * ((s.nsb + kB - 1) / kB);
  dim3 grid((s.nsb + kB - 1) / kB, s.nsa);
  const double d8 = 8.0 * (double)d.dim;
  const int gth = gather_threads(s);
  const dim3 ggrid((s.nsb + gth - 1) / gth, s.nsa);
  const int64_t gnb = (int64_t)ggrid.x * ggrid.y;
  if (fused_sigma()) {
    launch_gemm_fused(s, d, H.d_A, psi, G, st);
  } else {
    SQ_PROF(PC_DBUILD, st, d8 * (d.K4 + 1), 0.0,
            k_dbuild<<<grid, kB, 0, st>>>(s.n, s.nsa, s.nsb, d.npair, d.K4, ae, be, psi, D));
    launch_gemm(d, H.d_A, D, G, st);
  }
  SQ_PROF(PC_SIGMA_GATHER, st, d8 * (d.npair + 2), 0.0,
          k_sigma_gather<<<ggrid, gth, 0, st>>>(s.n, s.nsa, s.nsb, ae, be, G, psi, H.e_core, sigma, partial));
  reduce_impl(partial, gnb, e_dev, 0, false, 1.0, st);
}

// ---------------------------------------------------------------------------
// Dot products (two-stage, deterministic).
__global__ void k_dot_partial(const double* __restrict__ x, const double* __restrict__ y, int64_t n,
                              double* __restrict__ partial) {
  __shared__ double red[kB / 64];
  double v = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v += x[i] * y[i];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    partial[blockIdx.x] = t;
  }
}

void dot_dev(const double* x, const double* y, int64_t n, double* out_dev, int slot, double factor, double* partial,
             hipStream_t st) {
  int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + kB - 1) / kB));
  k_dot_partial<<<nb, kB, 0, st>>>(x, y, n, partial);
  SQ_HIP(hipGetLastError());
  reduce_impl(partial, nb, out_dev, slot, false, factor, st);
}

// Partial M = D D^T over one chunk of determinants (split over the
// determinant axis), D generated while it is staged, so D never touches
// memory.  Rows ("slots") of D, in three blocks of whole 16-row tiles:
//   L  slots [0, 16 MTL):          E_ab psi for the pairs a < b (lexicographic)
//   U  slots [16 MTL, 32 MTL):     E_ba psi for the same pairs, in the same order
//   Dg slots [32 MTL, 32 MTL + 16): E_kk psi (k < n), then psi, then zero rows
// (padding rows zero), with (E_pq psi)[I] = sign * psi[J] summed over the
// alpha and beta images J of I (full excitation tables, entry x = q*n+p as
// k_exc writes them).  Besides M = M^T, M is invariant under transposing
// both pairs up to 1-RDM terms (real psi, [E_pq, E_rs] = d_qr E_ps -
// d_ps E_rq):
//   M[(a,b)][(c,d)] = M[(b,a)][(d,c)] - d_bd g_ca + d_ac g_bd,
// so the U x U and U x Dg blocks and the lower tile triangle of L x U are
// skipped: only L x L (upper tiles), L x U (tiles i <= j), L x Dg and Dg x Dg
// are accumulated, NT = MTL (MTL + 1) + MTL + 1 tiles instead of the
// RT (RT + 1) / 2 of the plain upper triangle (RT = 2 MTL + 1); the host
// rebuilds the rest.  WAVES waves per block; wave w owns tiles w, w + WAVES,
// ....  Determinants are staged KS at a time in LDS, transposed (sD[k][row],
// row stride LD + 1).  Each thread owns NLOAD fixed (row, determinant
// offset) elements of a slice; their table entries are loaded two slices
// ahead and their amplitudes one slice ahead, so both overlap the MFMAs of
// the current slice.  dim < 2^31.
template <int MTL, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void k_gram_mfma(int n, int nsa, int nsb, const int32_t* __restrict__ afull,
                                                          const int32_t* __restrict__ bfull,
                                                          const double* __restrict__ psi, int64_t chunk,
                                                          int nchunks, int xcd, double* __restrict__ partial) {
  typedef double double4_t __attribute__((ext_vector_type(4)));
  constexpr int RT = 2 * MTL + 1;
  constexpr int LD = RT * 16, LDP = LD + 1;
  constexpr int KS = 16;
  constexpr int NTH = WAVES * 64;
  constexpr int NLOAD = (LD * KS + NTH - 1) / NTH;
  constexpr int NA = MTL * (MTL + 1) / 2;
  constexpr int NT = 2 * NA + MTL + 1;
  constexpr int PER = (NT + WAVES - 1) / WAVES;
  __shared__ double sD[KS * LDP];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int kq = lane >> 4, li = lane & 15;
  int tile_i[PER], tile_j[PER];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    int tile = wave + WAVES * t, ti = 0, tj = 0;
    if (tile < 2 * NA) {  // L x L, then L x U: upper triangles of MTL x MTL tiles
      const int blk = tile >= NA ? 1 : 0;
      int u = tile - blk * NA;
      while (u >= MTL - ti && ti < MTL) {
        u -= MTL - ti;
        ++ti;
      }
      tj = ti + u + blk * MTL;
    } else if (tile < 2 * NA + MTL) {  // L x Dg
      ti = tile - 2 * NA;
      tj = 2 * MTL;
    } else {  // Dg x Dg
      ti = tj = 2 * MTL;
    }
    tile_i[t] = ti;
    tile_j[t] = tj;
  }
  double4_t acc[PER];
#pragma unroll
  for (int t = 0; t < PER; ++t) acc[t] = double4_t{0.0, 0.0, 0.0, 0.0};
  const uint32_t dim = (uint32_t)nsa * (uint32_t)nsb;
  // chunk of this block: with xcd = 1, XCD b % 8 (workgroups go to the XCDs
  // round robin) takes a contiguous range of chunks, so the blocks running
  // together on one XCD cover neighbouring alpha rows
  int cidx = blockIdx.x;
  if (xcd) cidx = (int)((blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3));
  if (cidx >= nchunks) return;
  const uint32_t c0 = (uint32_t)((int64_t)cidx * chunk);
  const uint32_t c1 = (uint32_t)min((int64_t)c0 + chunk, (int64_t)dim);
  const double inv_nsb = 1.0 / (double)nsb;
  // owned elements: kind 0 zero row, 1 diagonal E_pp, 2 E_pq (p != q), 3 psi row
  const int npair = n * (n - 1) / 2;
  int kind[NLOAD], ex[NLOAD];
#pragma unroll
  for (int u = 0; u < NLOAD; ++u) {
    const int i = threadIdx.x + u * NTH;
    const int r = i / KS;
    kind[u] = 0;
    ex[u] = 0;
    if (i >= LD * KS) continue;
    if (r < 32 * MTL) {
      const bool up = r >= 16 * MTL;
      int pi = up ? r - 16 * MTL : r;
      if (pi < npair) {
        int a = 0;
        while (pi >= n - 1 - a) {  // pair pi = (a, b), a < b, lexicographic
          pi -= n - 1 - a;
          ++a;
        }
        const int b = a + 1 + pi;
        const int p = up ? b : a, q = up ? a : b;
        kind[u] = 2;
        ex[u] = q * n + p;
      }
    } else {
      const int k = r - 32 * MTL;
      if (k < n) {
        kind[u] = 1;
        ex[u] = k * n + k;
      } else if (k == n) {
        kind[u] = 3;
      }
    }
  }
  int ta[NLOAD], tb[NLOAD], tia[NLOAD], tib[NLOAD];
  auto tables = [&](uint32_t k0) {
#pragma unroll
    for (int u = 0; u < NLOAD; ++u) {
      const uint32_t I = k0 + (uint32_t)((threadIdx.x + u * NTH) % KS);
      tia[u] = -1;
      if (kind[u] != 0 && I < c1) {
        int ia = (int)((double)I * inv_nsb);
        int ib = (int)I - ia * nsb;
        if (ib < 0) {
          --ia;
          ib += nsb;
        } else if (ib >= nsb) {
          ++ia;
          ib -= nsb;
        }
        tia[u] = ia;
        tib[u] = ib;
        if (kind[u] <= 2) {
          ta[u] = afull[(size_t)ex[u] * nsa + ia];
          tb[u] = bfull[(size_t)ex[u] * nsb + ib];
        }
      }
    }
  };
  double xa[NLOAD], xb[NLOAD];
  int ca[NLOAD], cb[NLOAD];  // coefficients: signs or an occupation sum
  auto amps = [&](uint32_t k0) {
#pragma unroll
    for (int u = 0; u < NLOAD; ++u) {
      const uint32_t I = k0 + (uint32_t)((threadIdx.x + u * NTH) % KS);
      ca[u] = 0;
      cb[u] = 0;
      xa[u] = 0.0;
      xb[u] = 0.0;
      if (tia[u] < 0) continue;
      const int ia = tia[u], ib = tib[u];
      uint32_t oa = I, ob = I;
      if (kind[u] == 3) {
        ca[u] = 1;
      } else if (kind[u] == 1) {
        ca[u] = ta[u] + tb[u];
      } else {
        if (ta[u]) {
          oa = (uint32_t)(abs(ta[u]) - 1) * (uint32_t)nsb + (uint32_t)ib;
          ca[u] = ta[u] > 0 ? 1 : -1;
        }
        if (tb[u]) {
          ob = (uint32_t)ia * (uint32_t)nsb + (uint32_t)(abs(tb[u]) - 1);
          cb[u] = tb[u] > 0 ? 1 : -1;
        }
      }
      xa[u] = psi[oa];
      xb[u] = psi[ob];
    }
  };
  tables(c0);
  amps(c0);
  if (c0 + KS < c1) tables(c0 + KS);
  for (uint32_t k0 = c0; k0 < c1; k0 += KS) {
#pragma unroll
    for (int u = 0; u < NLOAD; ++u) {
      const int i = threadIdx.x + u * NTH;
      if (i < LD * KS) sD[(i % KS) * LDP + i / KS] = (double)ca[u] * xa[u] + (double)cb[u] * xb[u];
    }
    __syncthreads();
    if (k0 + KS < c1) amps(k0 + KS);
    if (k0 + 2 * KS < c1) tables(k0 + 2 * KS);
#pragma unroll
    for (int kk = 0; kk < KS; kk += 4) {
      const double* row = sD + (kk + kq) * LDP + li;
#pragma unroll
      for (int t = 0; t < PER; ++t) {
        if (wave + WAVES * t < NT) {
          // A[i][k] = D[ti*16+i][k0+kk+k], B[k][j] = D[tj*16+j][k0+kk+k]
          const double a = row[tile_i[t] * 16];
          const double b = row[tile_j[t] * 16];
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
        }
      }
    }
    __syncthreads();  // the next slice overwrites sD
  }
  const int ld = RT * 16;
  double* out = partial + (size_t)cidx * ld * ld;
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    if (wave + WAVES * t < NT) {
      const int ti = tile_i[t], tj = tile_j[t];
#pragma unroll
      for (int r = 0; r < 4; ++r) out[(size_t)(ti * 16 + kq + 4 * r) * ld + tj * 16 + li] = acc[t][r];
    }
  }
}

__global__ void k_sum_partials(const double* __restrict__ partial, int nchunks, int64_t len,
                               double* __restrict__ out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= len) return;
  double v = 0.0;
  for (int c = 0; c < nchunks; ++c) v += partial[(size_t)c * len + i];
  out[i] = v;
}

// RDMs.  With D laid out in the L | U | Dg slot blocks of k_gram_mfma, the
// kernel accumulates the needed tiles of M = D D^T over determinant chunks
// (one partial per chunk, summed in a fixed order); the host reads, for pair
// indices x = (p, q) (E_pq psi) and the psi row,
//   G(x, y) = <E_x psi | E_y psi> from a computed tile, directly or through
//             M = M^T, or else through the pair transposition
//             G((a,b),(c,d)) = G((b,a),(d,c)) - d_bd g_ca + d_ac g_bd,
//   rdm1_pq   = g_pq = <psi|E_pq psi>,
//   rdm2_pqrs = G((q,p),(r,s)) - d_qr rdm1_ps  (<psi|E_pq E_rs|psi> - ...).
static int rdm_mtl(int n) { return std::max(1, (n * (n - 1) / 2 + 15) / 16); }

int64_t rdm_work_bytes(const Space& s, int* rt_out, int* nchunks_out, int64_t* chunk_out) {
  const int rt = 2 * rdm_mtl(s.n) + 1;
  int64_t chunk = 4096;
  int64_t nchunks = (s.dim + chunk - 1) / chunk;
  while (nchunks > 1024) {
    chunk *= 2;
    nchunks = (s.dim + chunk - 1) / chunk;
  }
  if (nchunks < 1) nchunks = 1;
  if (rt_out) *rt_out = rt;
  if (nchunks_out) *nchunks_out = (int)nchunks;
  if (chunk_out) *chunk_out = chunk;
  const int64_t ld = (int64_t)rt * 16;
  return 8 * (nchunks * ld * ld + ld * ld + 64);
}

void rdm_impl(const Space& s, const double* psi, double* rdm1, double* rdm2, double* work, hipStream_t st) {
  int rt, nchunks;
  int64_t chunk;
  rdm_work_bytes(s, &rt, &nchunks, &chunk);
  SQ_CHECK(s.dim < ((int64_t)1 << 31), SQ_ERR_ARG, "RDMs need fewer than 2^31 determinants");
  SQ_CHECK(s.n <= 15, SQ_ERR_ARG, "RDMs support at most 15 orbitals");
  const int n = s.n;
  const int mtl = rdm_mtl(n);
  const int64_t ld = (int64_t)rt * 16;
  double* partial = work;
  double* M = partial + (size_t)nchunks * ld * ld;
  const int nt = mtl * (mtl + 1) + mtl + 1;
  // algorithmic traffic: psi and the two excitation-table rows per E_pq; flops: the accumulated tiles
  const double gram_bytes = 8.0 * (double)s.dim + 4.0 * (double)(n * n) * (s.nsa + s.nsb);
  const double gram_flops = 2.0 * 256.0 * nt * (double)s.dim;
  const char* xenv = getenv("SQ_GRAM_XCD");
  const int xcd = (xenv && xenv[0] == '0') ? 0 : 1;
  const unsigned gblocks = xcd ? (unsigned)((nchunks + 7) / 8 * 8) : (unsigned)nchunks;
  switch (mtl) {
#define SQ_R(MTL)                                                                                          \
  case MTL:                                                                                                \
    if (MTL >= 3)                                                                                          \
      SQ_PROF(PC_RDM_GRAM, st, gram_bytes, gram_flops,                                                     \
              k_gram_mfma<MTL, 16><<<gblocks, 1024, 0, st>>>(n, s.nsa, s.nsb, s.d_aexc_full, s.d_bexc_full, \
                                                             psi, chunk, nchunks, xcd, partial));          \
    else                                                                                                   \
      SQ_PROF(PC_RDM_GRAM, st, gram_bytes, gram_flops,                                                     \
              k_gram_mfma<MTL, 4><<<gblocks, 256, 0, st>>>(n, s.nsa, s.nsb, s.d_aexc_full, s.d_bexc_full,   \
                                                           psi, chunk, nchunks, xcd, partial));            \
    break;
    SQ_R(1) SQ_R(2) SQ_R(3) SQ_R(4) SQ_R(5) SQ_R(6) SQ_R(7)
#undef SQ_R
    default:
      throw Error{SQ_ERR_ARG, "unsupported orbital count for the RDM kernel"};
  }
  SQ_HIP(hipGetLastError());
  const int64_t len = ld * ld;
  k_sum_partials<<<(unsigned)((len + kB - 1) / kB), kB, 0, st>>>(partial, nchunks, len, M);
  SQ_HIP(hipGetLastError());
  std::vector<double> hM(len);
  SQ_HIP(hipMemcpyAsync(hM.data(), M, sizeof(double) * len, hipMemcpyDeviceToHost, st));
  SQ_HIP(hipStreamSynchronize(st));
  // slot of pair (p, q) and of the psi row
  std::vector<int> pidx((size_t)n * n, -1);
  for (int a = 0, i = 0; a < n; ++a)
    for (int b = a + 1; b < n; ++b, ++i) pidx[(size_t)a * n + b] = i;
  const int psi_slot = 32 * mtl + n;
  auto slot = [&](int p, int q) {
    if (p == q) return 32 * mtl + p;
    return p < q ? pidx[(size_t)p * n + q] : 16 * mtl + pidx[(size_t)q * n + p];
  };
  // accumulated tile pairs (t1 <= t2)
  auto computed = [&](int t1, int t2) {
    if (t2 < mtl) return true;                                    // L x L
    if (t1 < mtl && t2 < 2 * mtl) return t1 <= t2 - mtl;         // L x U, upper tiles
    if (t1 < mtl && t2 == 2 * mtl) return true;                   // L x Dg
    return t1 == 2 * mtl && t2 == 2 * mtl;                        // Dg x Dg
  };
  // M at slots (s1, s2) when its tile pair (in either order) was accumulated
  auto get = [&](int s1, int s2, double& v) {
    if (s1 / 16 > s2 / 16) std::swap(s1, s2);
    if (!computed(s1 / 16, s2 / 16)) return false;
    v = hM[(size_t)s1 * ld + s2];
    return true;
  };
  // 1-RDM g_pq = <psi|E_pq psi>: psi x L and psi x Dg are accumulated; g is symmetric
  std::vector<double> g1((size_t)n * n);
  for (int p = 0; p < n; ++p)
    for (int q = p; q < n; ++q) {
      double v = 0.0;
      SQ_CHECK(get(psi_slot, slot(p, q), v), SQ_ERR_ARG, "RDM tile layout is incomplete");
      g1[(size_t)p * n + q] = g1[(size_t)q * n + p] = v;
    }
  for (size_t i = 0; i < g1.size(); ++i) rdm1[i] = g1[i];
  // G((a,b),(c,d)) = <E_ab psi|E_cd psi>
  auto G = [&](int a, int b, int c, int d) {
    double v = 0.0;
    if (get(slot(a, b), slot(c, d), v)) return v;
    const bool ok = get(slot(b, a), slot(d, c), v);
    SQ_CHECK(ok, SQ_ERR_ARG, "RDM tile layout is incomplete");
    if (b == d) v -= g1[(size_t)c * n + a];
    if (a == c) v += g1[(size_t)b * n + d];
    return v;
  };
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q)
      for (int r = 0; r < n; ++r)
        for (int t = 0; t < n; ++t) {
          double v = G(q, p, r, t);
          if (q == r) v -= g1[(size_t)p * n + t];
          rdm2[((size_t)(p * n + q) * n + r) * n + t] = v;
        }
}

}  // namespace sq

using namespace sq;

extern "C" {

int sq_hamiltonian_create(const sq_space* sp, const double* h, const double* g, double e_core, sq_hamiltonian** out) {
  return guarded([&] {
    SQ_CHECK(sp && h && g && out, SQ_ERR_ARG, "NULL argument");
    const Space& s = sp->s;
    SigmaDims d = sigma_dims(s);
    const int n = s.n;
    auto H = [&](int p, int q) { double v = h[p * n + q]; return std::fabs(v) > 1e-14 ? v : 0.0; };
    auto Gf = [&](int p, int q, int r, int t) {
      double v = g[((size_t)(p * n + q) * n + r) * n + t];
      return std::fabs(v) > 1e-14 ? v : 0.0;
    };
    // symmetry check (the packed contraction relies on it)
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q) {
        SQ_CHECK(std::fabs(H(p, q) - H(q, p)) <= 1e-12 * (1.0 + std::fabs(H(p, q))), SQ_ERR_ARG,
                 "h must be symmetric");
        for (int r = 0; r < n; ++r)
          for (int t = 0; t < n; ++t) {
            double v = Gf(p, q, r, t);
            double tol = 1e-12 * (1.0 + std::fabs(v));
            SQ_CHECK(std::fabs(v - Gf(q, p, r, t)) <= tol && std::fabs(v - Gf(p, q, t, r)) <= tol &&
                         std::fabs(v - Gf(r, t, p, q)) <= tol,
                     SQ_ERR_ARG, "g must have the 8-fold real symmetry");
          }
      }
    std::vector<double> At((size_t)d.K4 * d.kpad, 0.0);
    std::vector<int> kp(d.npair), kq(d.npair);
    for (int p = 0, k = 0; p < n; ++p)
      for (int q = 0; q <= p; ++q, ++k) kp[k] = p, kq[k] = q;
    for (int m = 0; m < d.npair; ++m) {
      const int p = kp[m], q = kq[m];
      for (int k = 0; k < d.npair; ++k) At[(size_t)k * d.kpad + m] = 0.5 * Gf(p, q, kp[k], kq[k]);
      double hp = H(p, q);
      for (int r = 0; r < n; ++r) hp -= 0.5 * Gf(p, r, r, q);
      At[(size_t)d.npair * d.kpad + m] = hp;
    }
    auto* ho = new sq_hamiltonian();
    ho->h.space = &s;
    ho->h.npair = d.npair;
    ho->h.kpad = d.kpad;
    ho->h.e_core = e_core;
    hipError_t e = hipMalloc(&ho->h.d_A, sizeof(double) * At.size());
    if (e == hipSuccess) e = hipMemcpy(ho->h.d_A, At.data(), sizeof(double) * At.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      if (ho->h.d_A) (void)hipFree(ho->h.d_A);
      delete ho;
      throw Error{SQ_ERR_HIP, hipGetErrorString(e)};
    }
    try {
      if (row_sigma_applies(s)) build_alpha_ham(s, ho->h);
    } catch (...) {
      if (ho->h.d_Ma) (void)hipFree(ho->h.d_Ma);
      (void)hipFree(ho->h.d_A);
      delete ho;
      throw;
    }
    *out = ho;
  });
}

int sq_hamiltonian_destroy(sq_hamiltonian* H) {
  return guarded([&] {
    if (!H) return;
    if (H->h.d_A) (void)hipFree(H->h.d_A);
    if (H->h.d_Ma) (void)hipFree(H->h.d_Ma);
    delete H;
  });
}

int sq_sigma(const sq_space* sp, const sq_hamiltonian* H, const double* psi, double* sigma, double* energy,
             double* work, void* stream) {
  return guarded([&] {
    SQ_CHECK(sp && H && psi && sigma && work, SQ_ERR_ARG, "NULL argument");
    SQ_CHECK(psi != sigma, SQ_ERR_ARG, "psi and sigma must differ");
    SQ_CHECK(H->h.space == &sp->s, SQ_ERR_ARG, "Hamiltonian built for another space");
    sigma_impl(sp->s, H->h, psi, sigma, energy, work, as_stream(stream));
  });
}

int sq_rdm12(const sq_space* sp, const double* psi, double* rdm1, double* rdm2, double* work, void* stream) {
  return guarded([&] {
    SQ_CHECK(sp && psi && rdm1 && rdm2 && work, SQ_ERR_ARG, "NULL argument");
    SQ_CHECK(sp->s.n * sp->s.n + 1 <= 13 * 16, SQ_ERR_ARG, "RDM kernel supports at most 14 orbitals");
    rdm_impl(sp->s, psi, rdm1, rdm2, work, as_stream(stream));
  });
}

}  // extern "C"

extern "C" int sq_debug_gemm(const double* At, const double* D, double* G, int npair, int K4, int kpad, int64_t dim,
                             void* stream) {
  return guarded([&] {
    SQ_CHECK(At && D && G && npair > 0 && K4 % 4 == 0 && kpad % 16 == 0 && npair <= kpad && K4 <= kpad, SQ_ERR_ARG,
             "invalid GEMM shape");
    SigmaDims d;
    d.npair = npair;
    d.K4 = K4;
    d.kpad = kpad;
    d.dim = dim;
    launch_gemm(d, At, D, G, as_stream(stream));
  });
}
