// Orbital rotations, AO -> MO integral transforms and the orbital gradient on
// the device (replaces c_mo = C expm(-kappa), ups_wavefunction.py:367-385;
// one_/two_electron_integral_transform, integralfunctions.py:186-211;
// get_orbital_gradient, density_matrix.py:181-230).
//
// Every contraction is one strided, batched FP64-MFMA GEMM (k_dgemm):
//   C[b](i,j) = alpha sum_k A[b](i,k) B[b](k,j) + beta C[b](i,j),
// with element addresses base + b*bs + i*rs + j*cs for each operand, so
// transposes and index permutations of the quarter transforms are strides.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sq_internal.h"

namespace sq {

namespace {

struct Mat {
  const double* p;
  int64_t rs, cs, bs;
};

// 256 threads, 64 x 64 output tile, 16-deep contraction slices staged in LDS;
// wave w owns rows 32 (w >> 1) and columns 32 (w & 1) as 2 x 2 tiles of
// v_mfma_f64_16x16x4f64.  Slice loads follow the operand's unit stride
// (row- or column-major) so that neighbouring lanes read neighbouring
// addresses.
__global__ __launch_bounds__(256) void k_dgemm(int m, int n, int k, double alpha, Mat A, Mat B, double beta,
                                               double* __restrict__ C, int64_t c_rs, int64_t c_cs, int64_t c_bs) {
  typedef double double4_t __attribute__((ext_vector_type(4)));
  __shared__ double sA[16][65];  // [k][i]
  __shared__ double sB[16][65];  // [k][j]
  const int tn = (n + 63) / 64;  // output tiles flattened over blockIdx.x, column tiles fastest
  const int i0 = (int)(blockIdx.x / tn) * 64, j0 = (int)(blockIdx.x % tn) * 64;
  const int64_t bz = blockIdx.z;
  const double* Ap = A.p + bz * A.bs;
  const double* Bp = B.p + bz * B.bs;
  double* Cp = C + bz * c_bs;
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, kq = lane >> 4, li = lane & 15;
  const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
  const bool a_row = A.cs == 1;  // A row-major: k contiguous
  const bool b_row = B.cs == 1;  // B row-major: j contiguous
  double4_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = double4_t{0.0, 0.0, 0.0, 0.0};
  double ra[4], rb[4];
  int ai[4], ak[4], bk[4], bj[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (a_row) {
      ai[u] = t >> 2;
      ak[u] = (t & 3) * 4 + u;
    } else {
      ai[u] = t & 63;
      ak[u] = (t >> 6) * 4 + u;
    }
    if (b_row) {
      bk[u] = t >> 4;
      bj[u] = (t & 15) * 4 + u;
    } else {
      bk[u] = t & 15;
      bj[u] = (t >> 4) * 4 + u;
    }
  }
  auto load = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + ai[u], kk = k0 + ak[u];
      ra[u] = (i < m && kk < k) ? Ap[(int64_t)i * A.rs + (int64_t)kk * A.cs] : 0.0;
      const int kb = k0 + bk[u], j = j0 + bj[u];
      rb[u] = (kb < k && j < n) ? Bp[(int64_t)kb * B.rs + (int64_t)j * B.cs] : 0.0;
    }
  };
  if (k > 0) load(0);
  for (int k0 = 0; k0 < k; k0 += 16) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      sA[ak[u]][ai[u]] = ra[u];
      sB[bk[u]][bj[u]] = rb[u];
    }
    __syncthreads();
    if (k0 + 16 < k) load(k0 + 16);
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      const double a0 = sA[kk + kq][wi + li], a1 = sA[kk + kq][wi + 16 + li];
      const double b0 = sB[kk + kq][wj + li], b1 = sB[kk + kq][wj + 16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wi + a * 16 + kq + 4 * r, j = j0 + wj + b * 16 + li;
        if (i < m && j < n) {
          double* c = Cp + (int64_t)i * c_rs + (int64_t)j * c_cs;
          const double v = alpha * acc[a][b][r];
          *c = beta == 0.0 ? v : v + beta * *c;
        }
      }
}

void gemm(int m, int n, int k, double alpha, Mat A, Mat B, double beta, double* C, int64_t c_rs, int64_t c_cs,
          int64_t c_bs, int batch, hipStream_t st) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  SQ_CHECK(batch <= 65535, SQ_ERR_ARG, "GEMM batch too large");
  const int64_t tiles = (int64_t)((n + 63) / 64) * ((m + 63) / 64);
  SQ_CHECK(tiles < ((int64_t)1 << 31), SQ_ERR_ARG, "GEMM too large");
  dim3 grid((unsigned)tiles, 1, batch);
  SQ_PROF(PC_ORBITAL, st, 8.0 * batch * ((double)m * k + (double)k * n + (double)m * n),
          2.0 * batch * (double)m * n * k,
          k_dgemm<<<grid, 256, 0, st>>>(m, n, k, alpha, A, B, beta, C, c_rs, c_cs, c_bs));
}

Mat rowmaj(const double* p, int64_t ld, int64_t bs = 0) { return Mat{p, ld, 1, bs}; }
Mat trans(const double* p, int64_t ld, int64_t bs = 0) { return Mat{p, 1, ld, bs}; }

// out = a * x + b * I (n x n)
__global__ void k_axpby_eye(int n, double a, const double* __restrict__ x, double b, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * n) return;
  out[i] = a * x[i] + ((i / n == i % n) ? b : 0.0);
}

// exp(X) and c_new = c_old exp(X) in one block for n <= kSmallExpm (X, E and
// T in LDS, 3 n^2 doubles): the Horner recurrence E = I + X/14,
// E = I + X E / j (j = 13 .. 1) and s squarings of the GEMM path, with the
// block's threads taking the entries of each product (a few microseconds for
// the orbital counts of small molecules, where the GEMM path is ~30 launches).
constexpr int kSmallExpm = 48;

__global__ __launch_bounds__(256) void k_expm_small(int n, int n_ao, int s, const double* __restrict__ X,
                                                    const double* __restrict__ c_old, double* __restrict__ c_new) {
  extern __shared__ double sm[];
  const int nn = n * n;
  double* sX = sm;
  double* sE = sm + nn;
  double* sT = sm + 2 * nn;
  for (int i = threadIdx.x; i < nn; i += blockDim.x) {
    const double x = X[i];
    sX[i] = x;
    sE[i] = (i / n == i % n ? 1.0 : 0.0) + x / 14.0;
  }
  __syncthreads();
  for (int j = 13; j >= 1; --j) {
    for (int i = threadIdx.x; i < nn; i += blockDim.x) {
      const int r = i / n, c = i - r * n;
      double acc = 0.0;
      for (int l = 0; l < n; ++l) acc += sX[r * n + l] * sE[l * n + c];
      sT[i] = (r == c ? 1.0 : 0.0) + acc / j;
    }
    __syncthreads();  // every read of sE is done before it becomes the next output
    double* t = sE;
    sE = sT;
    sT = t;
  }
  for (int q = 0; q < s; ++q) {
    for (int i = threadIdx.x; i < nn; i += blockDim.x) {
      const int r = i / n, c = i - r * n;
      double acc = 0.0;
      for (int l = 0; l < n; ++l) acc += sE[r * n + l] * sE[l * n + c];
      sT[i] = acc;
    }
    __syncthreads();
    double* t = sE;
    sE = sT;
    sT = t;
  }
  for (int i = threadIdx.x; i < n_ao * n; i += blockDim.x) {
    const int a = i / n, c = i - a * n;
    double acc = 0.0;
    for (int l = 0; l < n; ++l) acc += c_old[(size_t)a * n + l] * sE[l * n + c];
    c_new[i] = acc;
  }
}

// h_mo = C^T h C and the four quarter transforms of g in one block for
// n_ao, n_mo <= kSmallAo (C is n_ao x n_mo row-major): the intermediates
// T1 (m a^3), T2 (m^2 a^2), T1 (m^3 a) live in two LDS buffers of kSmallAo^4
// doubles, with a barrier between passes; the same contractions as the GEMM
// path of sq_ao2mo, for the basis sizes of small molecules.
constexpr int kSmallAo = 8;

__global__ __launch_bounds__(256) void k_ao2mo_small(int a, int m, const double* __restrict__ c,
                                                     const double* __restrict__ h_ao, const double* __restrict__ g_ao,
                                                     double* __restrict__ h_mo, double* __restrict__ g_mo) {
  extern __shared__ double sm[];
  constexpr int kBuf = kSmallAo * kSmallAo * kSmallAo * kSmallAo;
  double* A = sm;
  double* B = sm + kBuf;
  const int a2 = a * a, a3 = a2 * a, m2 = m * m, m3 = m2 * m;
  if (h_ao) {
    // Q[i][b] = sum_x C[x][i] h[x][b];  h_mo[i][j] = sum_b Q[i][b] C[b][j]
    for (int x = threadIdx.x; x < m * a; x += blockDim.x) {
      const int i = x / a, b = x % a;
      double v = 0.0;
      for (int y = 0; y < a; ++y) v += c[y * m + i] * h_ao[y * a + b];
      A[x] = v;
    }
    __syncthreads();
    for (int x = threadIdx.x; x < m2; x += blockDim.x) {
      const int i = x / m, j = x % m;
      double v = 0.0;
      for (int b = 0; b < a; ++b) v += A[i * a + b] * c[b * m + j];
      h_mo[x] = v;
    }
    __syncthreads();
  }
  if (g_ao) {
    // A[i][bcd] = sum_x C[x][i] g[x][bcd]
    for (int x = threadIdx.x; x < m * a3; x += blockDim.x) {
      const int i = x / a3, r = x % a3;
      double v = 0.0;
      for (int y = 0; y < a; ++y) v += c[y * m + i] * g_ao[(int64_t)y * a3 + r];
      A[x] = v;
    }
    __syncthreads();
    // B[i][j][cd] = sum_b C[b][j] A[i][b][cd]
    for (int x = threadIdx.x; x < m2 * a2; x += blockDim.x) {
      const int i = x / (m * a2), j = (x / a2) % m, r = x % a2;
      double v = 0.0;
      for (int b = 0; b < a; ++b) v += c[b * m + j] * A[i * a3 + b * a2 + r];
      B[x] = v;
    }
    __syncthreads();
    // A[i][j][k][d] = sum_x C[x][k] B[i][j][x][d]
    for (int x = threadIdx.x; x < m3 * a; x += blockDim.x) {
      const int ij = x / (m * a), k = (x / a) % m, d = x % a;
      double v = 0.0;
      for (int y = 0; y < a; ++y) v += c[y * m + k] * B[ij * a2 + y * a + d];
      A[x] = v;
    }
    __syncthreads();
    // g_mo[ijk][l] = sum_d A[ijk][d] C[d][l]
    for (int x = threadIdx.x; x < m3 * m; x += blockDim.x) {
      const int ijk = x / m, l = x % m;
      double v = 0.0;
      for (int d = 0; d < a; ++d) v += A[ijk * a + d] * c[d * m + l];
      g_mo[x] = v;
    }
  }
}

// Full-space Gamma elements over the o = ni + na inactive + active orbitals
// (density_matrix.py:5-136), from the active rdm1 (na x na) and rdm2 (na^4).
__device__ __forceinline__ double full_rdm1_elem(int p, int q, int ni, int na, const double* __restrict__ rdm1) {
  if (p >= ni && q >= ni) return rdm1[(p - ni) * na + (q - ni)];
  if (p < ni && q < ni) return p == q ? 2.0 : 0.0;
  return 0.0;
}

__device__ __forceinline__ double full_rdm2_elem(int p, int q, int r, int s, int ni, int na,
                                                 const double* __restrict__ rdm1, const double* __restrict__ rdm2) {
  const bool ap = p >= ni, aq = q >= ni, ar = r >= ni, as = s >= ni;
  auto d1 = [&](int a, int b) { return rdm1[(a - ni) * na + (b - ni)]; };
  double v = 0.0;
  if (ap && aq && ar && as) {
    v = rdm2[(((int64_t)(p - ni) * na + (q - ni)) * na + (r - ni)) * na + (s - ni)];
  } else if (!ap && aq && ar && !as) {
    v = p == s ? -d1(q, r) : 0.0;
  } else if (ap && !aq && !ar && as) {
    v = q == r ? -d1(p, s) : 0.0;
  } else if (ap && aq && !ar && !as) {
    v = r == s ? 2.0 * d1(p, q) : 0.0;
  } else if (!ap && !aq && ar && as) {
    v = p == q ? 2.0 * d1(r, s) : 0.0;
  } else if (!(ap || aq || ar || as)) {
    if (p == q && r == s) v += 4.0;
    if (q == r && p == s) v -= 2.0;
  }
  return v;
}

// G1[p][q] (N x o, rows >= o zero) and the compact G2c[p][q][r][s] (o^4).
__global__ void k_full_rdm1(int N, int ni, int na, const double* __restrict__ rdm1, double* __restrict__ G1) {
  const int o = ni + na;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N * o) return;
  const int p = x / o, q = x % o;
  G1[x] = p < o ? full_rdm1_elem(p, q, ni, na, rdm1) : 0.0;
}

__global__ void k_full_rdm2(int ni, int na, const double* __restrict__ rdm1, const double* __restrict__ rdm2,
                            double* __restrict__ G2) {
  const int o = ni + na;
  const int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t o4 = (int64_t)o * o * o * o;
  if (x >= o4) return;
  const int s = (int)(x % o), r = (int)((x / o) % o), q = (int)((x / ((int64_t)o * o)) % o),
            p = (int)(x / ((int64_t)o * o * o));
  G2[x] = full_rdm2_elem(p, q, r, s, ni, na, rdm1, rdm2);
}

// The whole orbital gradient in one block for o <= kSmallGradO: Gamma1 and
// Gamma2 over the o orbitals in LDS; one wave per kappa element (m, n), its
// lanes splitting the o^3 index triples of the four sums Z(n,m), W(m,n),
// Z(m,n), W(n,m) of k_kappa_grad (and the o terms of Y and V), summed by a
// fixed-order shuffle reduction; g is read from global memory.
constexpr int kSmallGradO = 8;

__global__ __launch_bounds__(256) void k_kappa_grad_small(int N, int ni, int na, const double* __restrict__ h,
                                                          const double* __restrict__ g,
                                                          const double* __restrict__ rdm1,
                                                          const double* __restrict__ rdm2, int nk,
                                                          const int32_t* __restrict__ idx, double* __restrict__ grad) {
  __shared__ double sG1[kSmallGradO * kSmallGradO];
  __shared__ double sG2[kSmallGradO * kSmallGradO * kSmallGradO * kSmallGradO];
  const int o = ni + na, o2 = o * o, o3 = o2 * o, o4 = o3 * o;
  for (int x = threadIdx.x; x < o2; x += blockDim.x) sG1[x] = full_rdm1_elem(x / o, x % o, ni, na, rdm1);
  for (int x = threadIdx.x; x < o4; x += blockDim.x)
    sG2[x] = full_rdm2_elem(x / o3, (x / o2) % o, (x / o) % o, x % o, ni, na, rdm1, rdm2);
  __syncthreads();
  const int64_t N2 = (int64_t)N * N, N3 = N2 * N;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
  for (int i = wave; i < nk; i += nw) {
    const int m = idx[2 * i], n = idx[2 * i + 1];
    double t = 0.0;
    // grad = 2 Y[n][m] - 2 V[m][n] + Z(n,m) - W(m,n) - Z(m,n) + W(n,m), with
    // Z(a,b) = sum_pqr g[a][p][q][r] G2[b][p][q][r], W(a,b) = sum_pqr g[p][a][q][r] G2[p][b][q][r]
    // (terms with an index b >= o vanish with Gamma)
    for (int x = lane; x < o3; x += 64) {
      const int p = x / o2, q = (x / o) % o, r = x % o;
      const int64_t qr = (int64_t)q * N + r;
      const int pqr = p * o2 + q * o + r;  // = x
      if (m < o)
        t += g[n * N3 + p * N2 + qr] * sG2[m * o3 + pqr] + g[p * N3 + n * N2 + qr] * sG2[p * o3 + m * o2 + q * o + r];
      if (n < o)
        t -= g[m * N3 + p * N2 + qr] * sG2[n * o3 + pqr] + g[p * N3 + m * N2 + qr] * sG2[p * o3 + n * o2 + q * o + r];
    }
    if (lane < o) {
      if (m < o) t += 2.0 * h[(int64_t)n * N + lane] * sG1[m * o + lane];
      if (n < o) t -= 2.0 * h[(int64_t)lane * N + m] * sG1[lane * o + n];
    }
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    if (lane == 0) grad[i] = t;
  }
}

// gc[a][pqr] = g[a][p][q][r] and gt[a][pqr] = g[p][a][q][r] over p, q, r < o.
__global__ void k_g_blocks(int N, int o, const double* __restrict__ g, double* __restrict__ gc,
                           double* __restrict__ gt) {
  const int64_t o3 = (int64_t)o * o * o;
  const int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (x >= (int64_t)N * o3) return;
  const int a = (int)(x / o3);
  const int64_t rem = x % o3;
  const int p = (int)(rem / ((int64_t)o * o)), q = (int)((rem / o) % o), r = (int)(rem % o);
  const int64_t N2 = (int64_t)N * N;
  gc[x] = g[((int64_t)a * N + p) * N2 + (int64_t)q * N + r];
  gt[x] = g[((int64_t)p * N + a) * N2 + (int64_t)q * N + r];
}

// G2t[b][pqr] = G2c[p][b][q][r]
__global__ void k_perm01(int o, const double* __restrict__ G2, double* __restrict__ G2t) {
  const int64_t o3 = (int64_t)o * o * o;
  const int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (x >= (int64_t)o * o3) return;
  const int b = (int)(x / o3);
  const int64_t rem = x % o3;
  const int p = (int)(rem / ((int64_t)o * o));
  const int64_t qr = rem % ((int64_t)o * o);
  G2t[x] = G2[((int64_t)p * o + b) * o * o + qr];
}

// grad[i] = 2 Y[n][m] - 2 V[m][n] + Z[n][m] - W[m][n] - Z[m][n] + W[n][m]
// with Y = h G1^T (N x N) and V = h^T G1, Z = gc G2c^T, W = gt G2t^T stored
// N x o (their columns >= o vanish with Gamma outside the o orbitals).
__global__ void k_kappa_grad(int N, int o, int nk, const int32_t* __restrict__ idx, const double* __restrict__ Y,
                             const double* __restrict__ V, const double* __restrict__ Z,
                             const double* __restrict__ W, double* __restrict__ grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk) return;
  const int m = idx[2 * i], n = idx[2 * i + 1];
  auto z = [&](const double* M, int a, int b) { return b < o ? M[(int64_t)a * o + b] : 0.0; };
  grad[i] = 2.0 * Y[(int64_t)n * N + m] - 2.0 * z(V, m, n) + z(Z, n, m) - z(W, m, n) - z(Z, m, n) + z(W, n, m);
}

int64_t ao2mo_work_doubles(int nao, int nmo) {
  const int64_t a = nao, m = nmo;
  const int64_t t = std::max({m * a * a * a, m * m * a * a, m * m * m * a});
  return 2 * t + m * a;
}

int64_t grad_work_doubles(int N, int o, int nk) {
  const int64_t o3 = (int64_t)o * o * o;
  return (int64_t)N * o + o * o3 + o * o3 + 2 * (int64_t)N * o3 + 2 * (int64_t)N * N + 2 * (int64_t)N * o + nk +
         2 * (int64_t)nk + 64;
}

}  // namespace

}  // namespace sq

using namespace sq;

extern "C" {

int64_t sq_orbital_work_bytes(int n_ao, int n_mo, int n_inactive, int n_active, int n_kappa) {
  if (n_ao <= 0 || n_mo <= 0 || n_inactive < 0 || n_active < 0 || n_inactive + n_active > n_mo || n_kappa < 0)
    return -1;
  const int64_t rot = 3 * (int64_t)n_mo * n_mo;
  return 8 * std::max({rot, ao2mo_work_doubles(n_ao, n_mo), grad_work_doubles(n_mo, n_inactive + n_active, n_kappa)});
}

int sq_orbital_rotation(int n, int n_ao, const double* c_old, const double* kappa, double* c_new, double* work,
                        void* stream) {
  return guarded([&] {
    SQ_CHECK(n > 0 && n_ao > 0 && c_old && kappa && c_new && work, SQ_ERR_ARG, "invalid arguments");
    SQ_CHECK(c_old != c_new, SQ_ERR_ARG, "c_old and c_new must differ");
    hipStream_t st = as_stream(stream);
    // X = -kappa / 2^s with ||X||_1 <= 1/4; exp(X) by a degree-14 Taylor
    // polynomial in Horner form (truncation below 1e-21 relative), then s squarings
    double norm = 0.0;
    for (int j = 0; j < n; ++j) {
      double c = 0.0;
      for (int i = 0; i < n; ++i) c += std::fabs(kappa[(size_t)i * n + j]);
      norm = std::max(norm, c);
    }
    int s = 0;
    while (norm / std::ldexp(1.0, s) > 0.25 && s < 60) ++s;
    const size_t nn = (size_t)n * n;
    std::vector<double> X(nn);
    for (size_t i = 0; i < nn; ++i) X[i] = -std::ldexp(kappa[i], -s);
    double* dX = work;
    double* E = work + nn;
    double* T = work + 2 * nn;
    SQ_HIP(hipMemcpyAsync(dX, X.data(), sizeof(double) * nn, hipMemcpyHostToDevice, st));
    if (n <= kSmallExpm) {
      k_expm_small<<<1, 256, sizeof(double) * 3 * nn, st>>>(n, n_ao, s, dX, c_old, c_new);
      SQ_HIP(hipGetLastError());
      SQ_HIP(hipStreamSynchronize(st));  // X lives in host memory owned by this call
      return;
    }
    const unsigned nb = (unsigned)((nn + 255) / 256);
    // E = I + X/14; E = I + X E / j for j = 13 .. 1
    k_axpby_eye<<<nb, 256, 0, st>>>(n, 1.0 / 14.0, dX, 1.0, E);
    SQ_HIP(hipGetLastError());
    for (int j = 13; j >= 1; --j) {
      k_axpby_eye<<<nb, 256, 0, st>>>(n, 0.0, E, 1.0, T);  // T = I
      SQ_HIP(hipGetLastError());
      gemm(n, n, n, 1.0 / j, rowmaj(dX, n), rowmaj(E, n), 1.0, T, n, 1, 0, 1, st);  // T = I + X E / j
      std::swap(E, T);
    }
    for (int q = 0; q < s; ++q) {
      gemm(n, n, n, 1.0, rowmaj(E, n), rowmaj(E, n), 0.0, T, n, 1, 0, 1, st);
      std::swap(E, T);
    }
    gemm(n_ao, n, n, 1.0, rowmaj(c_old, n), rowmaj(E, n), 0.0, c_new, n, 1, 0, 1, st);
    SQ_HIP(hipStreamSynchronize(st));  // X lives in host memory owned by this call
  });
}

int sq_ao2mo(int n_ao, int n_mo, const double* c, const double* h_ao, const double* g_ao, double* h_mo,
             double* g_mo, void* work_v, void* stream) {
  double* work = static_cast<double*>(work_v);
  return guarded([&] {
    SQ_CHECK(n_ao > 0 && n_mo > 0 && n_mo <= 255 && c && work, SQ_ERR_ARG, "invalid arguments");
    SQ_CHECK((h_ao != nullptr) == (h_mo != nullptr) && (g_ao != nullptr) == (g_mo != nullptr), SQ_ERR_ARG,
             "each transform needs both its input and its output");
    hipStream_t st = as_stream(stream);
    if (n_ao <= kSmallAo && n_mo <= kSmallAo) {
      const size_t lds = sizeof(double) * 2 * kSmallAo * kSmallAo * kSmallAo * kSmallAo;
      k_ao2mo_small<<<1, 256, lds, st>>>(n_ao, n_mo, c, h_ao, g_ao, h_mo, g_mo);
      SQ_HIP(hipGetLastError());
      return;
    }
    const int64_t a = n_ao, m = n_mo;
    const int64_t tmax = std::max({m * a * a * a, m * m * a * a, m * m * m * a});
    double* T1 = work;
    double* T2 = work + tmax;
    double* Q = work + 2 * tmax;  // m x a
    if (h_ao) {
      // Q[i][b] = sum_a C[a][i] h[a][b]; h_mo[i][j] = sum_b Q[i][b] C[b][j]
      gemm((int)m, (int)a, (int)a, 1.0, trans(c, m), rowmaj(h_ao, a), 0.0, Q, a, 1, 0, 1, st);
      gemm((int)m, (int)m, (int)a, 1.0, rowmaj(Q, a), rowmaj(c, m), 0.0, h_mo, m, 1, 0, 1, st);
    }
    if (g_ao) {
      SQ_CHECK(a * a * a <= ((int64_t)1 << 40), SQ_ERR_ARG, "AO basis too large");
      // T1[i][bcd] = sum_a C[a][i] g[a][bcd]
      gemm((int)m, (int)(a * a * a), (int)a, 1.0, trans(c, m), rowmaj(g_ao, a * a * a), 0.0, T1, a * a * a, 1, 0, 1,
           st);
      // T2[i][j][cd] = sum_b C[b][j] T1[i][b][cd]   (batch i)
      gemm((int)m, (int)(a * a), (int)a, 1.0, trans(c, m), Mat{T1, a * a, 1, a * a * a}, 0.0, T2, a * a, 1, m * a * a,
           (int)m, st);
      // T1[ij][k][d] = sum_c C[c][k] T2[ij][c][d]   (batch ij)
      gemm((int)m, (int)a, (int)a, 1.0, trans(c, m), Mat{T2, a, 1, a * a}, 0.0, T1, a, 1, m * a, (int)(m * m), st);
      // g_mo[ijk][l] = sum_d T1[ijk][d] C[d][l]
      gemm((int)(m * m * m), (int)m, (int)a, 1.0, rowmaj(T1, a), rowmaj(c, m), 0.0, g_mo, m, 1, 0, 1, st);
    }
  });
}

int sq_orbital_gradient(int n_orb, int n_inactive, int n_active, const double* h_mo, const double* g_mo,
                        const double* rdm1, const double* rdm2, int n_kappa, const int32_t* kappa_idx, double* grad,
                        void* work_v, void* stream) {
  double* work = static_cast<double*>(work_v);
  return guarded([&] {
    const int N = n_orb, ni = n_inactive, na = n_active, o = ni + na;
    SQ_CHECK(N > 0 && ni >= 0 && na > 0 && o <= N && n_kappa >= 0, SQ_ERR_ARG, "invalid orbital counts");
    SQ_CHECK(h_mo && g_mo && rdm1 && rdm2 && work && (n_kappa == 0 || (kappa_idx && grad)), SQ_ERR_ARG,
             "NULL argument");
    for (int i = 0; i < 2 * n_kappa; ++i)
      SQ_CHECK(kappa_idx[i] >= 0 && kappa_idx[i] < N, SQ_ERR_ARG, "kappa index out of range");
    if (n_kappa == 0) return;
    hipStream_t st = as_stream(stream);
    const int64_t o3 = (int64_t)o * o * o;
    double* G1 = work;                 // N x o
    double* G2 = G1 + (int64_t)N * o;  // o^4 (compact)
    double* G2t = G2 + o * o3;         // o x o^3
    double* gc = G2t + o * o3;         // N x o^3
    double* gt = gc + (int64_t)N * o3;
    double* Y = gt + (int64_t)N * o3;  // N x N
    double* V = Y + (int64_t)N * N;    // N x o
    double* Z = V + (int64_t)N * o;    // N x o
    double* W = Z + (int64_t)N * o;
    double* dgrad = W + (int64_t)N * o;
    int32_t* didx = reinterpret_cast<int32_t*>(dgrad + n_kappa);
    SQ_HIP(hipMemcpyAsync(didx, kappa_idx, sizeof(int32_t) * 2 * n_kappa, hipMemcpyHostToDevice, st));
    if (o <= kSmallGradO) {
      k_kappa_grad_small<<<1, 256, 0, st>>>(N, ni, na, h_mo, g_mo, rdm1, rdm2, n_kappa, didx, dgrad);
      SQ_HIP(hipGetLastError());
      SQ_HIP(hipMemcpyAsync(grad, dgrad, sizeof(double) * n_kappa, hipMemcpyDeviceToHost, st));
      SQ_HIP(hipStreamSynchronize(st));
      return;
    }
    k_full_rdm1<<<(unsigned)((N * o + 255) / 256), 256, 0, st>>>(N, ni, na, rdm1, G1);
    SQ_HIP(hipGetLastError());
    k_full_rdm2<<<(unsigned)((o * o3 + 255) / 256), 256, 0, st>>>(ni, na, rdm1, rdm2, G2);
    SQ_HIP(hipGetLastError());
    k_perm01<<<(unsigned)((o * o3 + 255) / 256), 256, 0, st>>>(o, G2, G2t);
    SQ_HIP(hipGetLastError());
    k_g_blocks<<<(unsigned)(((int64_t)N * o3 + 255) / 256), 256, 0, st>>>(N, o, g_mo, gc, gt);
    SQ_HIP(hipGetLastError());
    // Y[n][m] = sum_{p<o} h[n][p] G1[m][p]
    gemm(N, N, o, 1.0, rowmaj(h_mo, N), trans(G1, o), 0.0, Y, N, 1, 0, 1, st);
    // V[m][n] = sum_{p<o} h[p][m] G1[p][n] for n < o (Gamma1 vanishes for n >= o)
    gemm(N, o, o, 1.0, trans(h_mo, N), rowmaj(G1, o), 0.0, V, o, 1, 0, 1, st);
    // Z[a][b] = sum_{pqr} g[a][pqr] G2[b][pqr];  W[a][b] = sum_{pqr} g[p][a][qr] G2[p][b][qr]
    gemm(N, o, (int)o3, 1.0, rowmaj(gc, o3), trans(G2, o3), 0.0, Z, o, 1, 0, 1, st);
    gemm(N, o, (int)o3, 1.0, rowmaj(gt, o3), trans(G2t, o3), 0.0, W, o, 1, 0, 1, st);
    k_kappa_grad<<<(unsigned)((n_kappa + 255) / 256), 256, 0, st>>>(N, o, n_kappa, didx, Y, V, Z, W, dgrad);
    SQ_HIP(hipGetLastError());
    SQ_HIP(hipMemcpyAsync(grad, dgrad, sizeof(double) * n_kappa, hipMemcpyDeviceToHost, st));
    SQ_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"
