// CI-space addressing on the device: string tables, closed-form ranks and
// E_pq single-excitation tables (replaces ci_spaces.get_indexing,
// ci_spaces.py:56-116, and the per-call parity_check / det2idx lookups of
// operator_state_algebra.py:517-522,118-135).
//
// Strings are bit masks with bit p = spatial orbital p.  The lexicographic
// order of itertools.combinations is reproduced by the combinatorial number
// system (rank = number of combinations that precede it), so no hash map is
// needed: rank tables are direct-indexed by the 2^n mask.

#include <cstdio>
#include <cstring>
#include <mutex>

#include "sq_internal.h"

namespace sq {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

__device__ __forceinline__ int64_t binom_dev(int n, int k) {
  if (k < 0 || k > n) return 0;
  int64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

static int64_t binom_host(int n, int k) {
  if (k < 0 || k > n) return 0;
  int64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// Unrank r in lexicographic combination order of k elements out of n.
__global__ void k_unrank(int n, int k, int64_t count, uint32_t* out) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= count) return;
  int64_t rem = r;
  uint32_t mask = 0;
  int j = 0;
  for (int i = 0; i < k; ++i) {
    while (true) {
      int64_t c = binom_dev(n - 1 - j, k - 1 - i);
      if (rem < c) break;
      rem -= c;
      ++j;
    }
    mask |= 1u << j;
    ++j;
  }
  out[r] = mask;
}

// Rank table over all 2^n masks: lexicographic rank if popcount == k, else -1.
__global__ void k_rank(int n, int k, int32_t* rank) {
  uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= (1u << n)) return;
  if (__popc(m) != k) {
    rank[m] = -1;
    return;
  }
  int64_t r = 0;
  int prev = -1, i = 0;
  for (int c = 0; c < n; ++c) {
    if (!((m >> c) & 1u)) continue;
    for (int j = prev + 1; j < c; ++j) r += binom_dev(n - 1 - j, k - 1 - i);
    prev = c;
    ++i;
  }
  rank[m] = (int32_t)r;
}

__device__ __forceinline__ uint32_t between_mask(int p, int q) {
  int lo = p < q ? p : q, hi = p < q ? q : p;
  return ((1u << hi) - 1u) & ~((1u << (lo + 1)) - 1u);
}

// Packed E+_k table (k = p(p+1)/2 + q, p >= q) and full E_pq table, both
// k-major (entry [k * nstr + s]) so that threads over consecutive strings
// read consecutive words.
__global__ void k_exc(int n, int nstr, const uint32_t* str, const int32_t* rank, int32_t* packed,
                      int32_t* full) {
  int s = blockIdx.x;
  if (s >= nstr) return;
  uint32_t m = str[s];
  int npair = n * (n + 1) / 2;
  for (int x = threadIdx.x; x < n * n; x += blockDim.x) {
    int p = x / n, q = x % n;
    int32_t v = 0;
    if (p == q) {
      v = (m >> p) & 1u;
    } else if (((m >> q) & 1u) && !((m >> p) & 1u)) {
      uint32_t t = (m & ~(1u << q)) | (1u << p);
      int sign = (__popc(m & between_mask(p, q)) & 1) ? -1 : 1;
      v = sign * (rank[t] + 1);
    }
    full[(size_t)x * nstr + s] = v;
    if (p >= q) {
      int k = p * (p + 1) / 2 + q;
      int32_t w;
      if (p == q) {
        w = v;
      } else if (((m >> q) & 1u) && !((m >> p) & 1u)) {
        w = v;  // E_pq acts
      } else if (((m >> p) & 1u) && !((m >> q) & 1u)) {
        uint32_t t = (m & ~(1u << p)) | (1u << q);
        int sign = (__popc(m & between_mask(p, q)) & 1) ? -1 : 1;
        w = sign * (rank[t] + 1);  // E_qp acts
      } else {
        w = 0;
      }
      packed[(size_t)k * nstr + s] = w;
    }
  }
}

// Blocked <-> interleaved sign: s(I) = (-1)^{sum_{p in alpha} #{q in beta : q < p}}.
__global__ void k_phase(int nsa, int nsb, const uint32_t* astr, const uint32_t* bstr, const double* in,
                        double* out) {
  int ia = blockIdx.y;
  int ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= nsb) return;
  uint32_t a = astr[ia], b = bstr[ib];
  int par = 0;
  uint32_t aa = a;
  while (aa) {
    int p = __ffs(aa) - 1;
    aa &= aa - 1;
    par += __popc(b & ((1u << p) - 1u));
  }
  size_t I = (size_t)ia * nsb + ib;
  double v = in[I];
  out[I] = (par & 1) ? -v : v;
}

// Reference determinant integers: spin orbital k = 2p (+1 for beta) at bit 2n-1-k.
__global__ void k_idx2det(int n, int nsa, int nsb, const uint32_t* astr, const uint32_t* bstr, int64_t* out) {
  int ia = blockIdx.y;
  int ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= nsb) return;
  uint32_t a = astr[ia], b = bstr[ib];
  int64_t det = 0;
  for (int p = 0; p < n; ++p) {
    if ((a >> p) & 1u) det |= (int64_t)1 << (2 * n - 1 - 2 * p);
    if ((b >> p) & 1u) det |= (int64_t)1 << (2 * n - 1 - (2 * p + 1));
  }
  out[(size_t)ia * nsb + ib] = det;
}

static void build_spin(int n, int k, int nstr, uint32_t** d_str, int32_t** d_rank, int32_t** d_packed,
                       int32_t** d_full, std::vector<uint32_t>& h_str) {
  SQ_HIP(hipMalloc(d_str, sizeof(uint32_t) * (size_t)std::max(nstr, 1)));
  SQ_HIP(hipMalloc(d_rank, sizeof(int32_t) * ((size_t)1 << n)));
  int npair = n * (n + 1) / 2;
  SQ_HIP(hipMalloc(d_packed, sizeof(int32_t) * (size_t)std::max(nstr, 1) * npair));
  SQ_HIP(hipMalloc(d_full, sizeof(int32_t) * (size_t)std::max(nstr, 1) * n * n));
  if (nstr > 0) {
    k_unrank<<<(nstr + 255) / 256, 256>>>(n, k, nstr, *d_str);
    SQ_HIP(hipGetLastError());
  }
  k_rank<<<((1u << n) + 255) / 256, 256>>>(n, k, *d_rank);
  SQ_HIP(hipGetLastError());
  if (nstr > 0) {
    k_exc<<<nstr, 128>>>(n, nstr, *d_str, *d_rank, *d_packed, *d_full);
    SQ_HIP(hipGetLastError());
  }
  SQ_HIP(hipDeviceSynchronize());
  h_str.resize(nstr);
  if (nstr > 0) SQ_HIP(hipMemcpy(h_str.data(), *d_str, sizeof(uint32_t) * nstr, hipMemcpyDeviceToHost));
}

}  // namespace sq

using namespace sq;

extern "C" {

const char* sq_last_error(void) { return g_last_error.c_str(); }

const char* sq_version(void) { return "sqhip 1.0 gfx950"; }

int sq_space_create(int n_orb, int n_alpha, int n_beta, sq_space** out) {
  return guarded([&] {
    SQ_CHECK(out != nullptr, SQ_ERR_ARG, "out is NULL");
    SQ_CHECK(n_orb >= 1 && n_orb <= kMaxOrb, SQ_ERR_ARG, "n_orb must be in [1, 16]");
    SQ_CHECK(n_alpha >= 0 && n_alpha <= n_orb && n_beta >= 0 && n_beta <= n_orb, SQ_ERR_ARG,
             "electron counts out of range");
    auto* sp = new sq_space();
    Space& s = sp->s;
    s.n = n_orb;
    s.na = n_alpha;
    s.nb = n_beta;
    s.nsa = (int)binom_host(n_orb, n_alpha);
    s.nsb = (int)binom_host(n_orb, n_beta);
    s.dim = (int64_t)s.nsa * s.nsb;
    s.npair = n_orb * (n_orb + 1) / 2;
    SQ_HIP(hipGetDevice(&s.device));
    try {
      build_spin(n_orb, n_alpha, s.nsa, &s.d_astr, &s.d_arank, &s.d_aexc, &s.d_aexc_full, s.h_astr);
      build_spin(n_orb, n_beta, s.nsb, &s.d_bstr, &s.d_brank, &s.d_bexc, &s.d_bexc_full, s.h_bstr);
      build_gemm_tables(s);
      build_row_tables(s);
    } catch (...) {
      sq_space_destroy(sp);
      throw;
    }
    *out = sp;
  });
}

int sq_space_destroy(sq_space* sp) {
  return guarded([&] {
    if (!sp) return;
    Space& s = sp->s;
    free_row_tables(s);
    for (void* p : {(void*)s.d_astr, (void*)s.d_bstr, (void*)s.d_arank, (void*)s.d_brank, (void*)s.d_aexc,
                    (void*)s.d_bexc, (void*)s.d_aexc_full, (void*)s.d_bexc_full, (void*)s.d_gta, (void*)s.d_gtb})
      if (p) (void)hipFree(p);
    delete sp;
  });
}

int64_t sq_space_dim(const sq_space* sp) { return sp ? sp->s.dim : -1; }

int sq_space_nstrings(const sq_space* sp, int32_t* na, int32_t* nb) {
  return guarded([&] {
    SQ_CHECK(sp && na && nb, SQ_ERR_ARG, "NULL argument");
    *na = sp->s.nsa;
    *nb = sp->s.nsb;
  });
}

int sq_space_copy_strings(const sq_space* sp, uint32_t* ha, uint32_t* hb) {
  return guarded([&] {
    SQ_CHECK(sp && ha && hb, SQ_ERR_ARG, "NULL argument");
    std::memcpy(ha, sp->s.h_astr.data(), sizeof(uint32_t) * sp->s.nsa);
    std::memcpy(hb, sp->s.h_bstr.data(), sizeof(uint32_t) * sp->s.nsb);
  });
}

int sq_space_idx2det(const sq_space* sp, int64_t* h, void* stream) {
  return guarded([&] {
    SQ_CHECK(sp && h, SQ_ERR_ARG, "NULL argument");
    const Space& s = sp->s;
    SQ_CHECK(2 * s.n <= 62, SQ_ERR_ARG, "determinant integers need 2n <= 62");
    hipStream_t st = as_stream(stream);
    int64_t* d = nullptr;
    SQ_HIP(hipMalloc(&d, sizeof(int64_t) * std::max<int64_t>(s.dim, 1)));
    dim3 grid((s.nsb + 255) / 256, s.nsa);
    if (s.dim > 0) k_idx2det<<<grid, 256, 0, st>>>(s.n, s.nsa, s.nsb, s.d_astr, s.d_bstr, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(int64_t) * s.dim, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    SQ_CHECK(e == hipSuccess, SQ_ERR_HIP, hipGetErrorString(e));
  });
}

static int phase_convert(const sq_space* sp, const double* in, double* out, void* stream) {
  return guarded([&] {
    SQ_CHECK(sp && in && out, SQ_ERR_ARG, "NULL argument");
    const Space& s = sp->s;
    if (s.dim == 0) return;
    dim3 grid((s.nsb + 255) / 256, s.nsa);
    k_phase<<<grid, 256, 0, as_stream(stream)>>>(s.nsa, s.nsb, s.d_astr, s.d_bstr, in, out);
    SQ_HIP(hipGetLastError());
  });
}

int sq_state_from_reference(const sq_space* sp, const double* ref, double* internal, void* stream) {
  return phase_convert(sp, ref, internal, stream);
}

int sq_state_to_reference(const sq_space* sp, const double* internal, double* ref, void* stream) {
  return phase_convert(sp, internal, ref, stream);
}

}  // extern "C"
