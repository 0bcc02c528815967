// sigma = H psi with no allocation: a variant of sq_sigma (sq_sigma.hip)
// whose scratch memory is a file-scope device array instead of a caller's
// workspace, for callers that cannot allocate.
//
// The row sigma's alpha GEMM partials and block energy partials live in
// `scratch`, kScratchDoubles doubles (64 MiB) that every GPU the library runs
// on holds until the process ends.  Host code never passes the array's name
// to a kernel: at the first call on a device, hipGetSymbolAddress gives that
// device's address, kept per device and used for every kernel argument.
// Every kernel writes the scratch it reads in the same call (nothing needs
// clearing); calls that overlap on one GPU share the array and race.

#include <mutex>

#include "sq_internal.h"

namespace sq {

void reduce_impl(const double* partial, int64_t n, double* out_dev, int slot, bool accumulate, double factor,
                 hipStream_t st);

constexpr int64_t kScratchDoubles = (int64_t)8 << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

double* scratch_for_current_device() {
  int dev = -1;
  SQ_HIP(hipGetDevice(&dev));
  SQ_CHECK(dev >= 0 && dev < kMaxDevices, SQ_ERR_ARG, "device index out of range for the sigma scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SQ_CHECK(p != nullptr, SQ_ERR_HIP, "no device address for the sigma scratch");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

}  // namespace sq

using namespace sq;

extern "C" int sq_sigma_noalloc(const sq_space* sp, const sq_hamiltonian* H, const double* psi, double* sigma,
                                double* energy, void* stream) {
  return guarded([&] {
    SQ_CHECK(sp && H && psi && sigma, SQ_ERR_ARG, "NULL argument");
    SQ_CHECK(psi != sigma, SQ_ERR_ARG, "psi and sigma must differ");
    SQ_CHECK(H->h.space == &sp->s, SQ_ERR_ARG, "Hamiltonian built for another space");
    const Space& s = sp->s;
    SQ_CHECK(H->h.d_Ma && row_sigma_applies(s), SQ_ERR_ARG,
             "sq_sigma_noalloc needs the row sigma (at most 4096 beta strings and 2^31 determinants)");
    const int64_t need = row_sigma_used_doubles(s);
    SQ_CHECK(need >= 0 && need <= kScratchDoubles, SQ_ERR_ARG,
             "space too large for the 64 MiB sigma scratch of sq_sigma_noalloc");
    double* w = scratch_for_current_device();
    hipStream_t st = as_stream(stream);
    double* partial = nullptr;
    const int64_t nb = row_sigma(s, H->h, psi, sigma, w, &partial, st, false);
    if (energy) reduce_impl(partial, nb, energy, 0, false, 1.0, st);
  });
}
