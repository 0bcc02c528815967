// Optional per-launch kernel timing with HIP events on the launch stream.
// When enabled (sq_profile_enable), every instrumented launch records an
// event pair; sq_profile_collect synchronises and aggregates elapsed time,
// launch counts and the launches' algorithmic bytes / flops per kernel class.

#include <array>
#include <mutex>

#include "sq_internal.h"

namespace sq {

namespace {
struct Rec {
  int cls;
  hipEvent_t a, b;
  double bytes, flops;
};
struct Prof {
  bool on = false;
  uint32_t mask = 0xffffffffu;  // kernel classes to time (sq_profile_select)
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  std::mutex mu;
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    SQ_HIP(hipEventCreate(&e));
    return e;
  }
};
Prof& prof() {
  static Prof p;
  return p;
}
thread_local hipEvent_t t_start = nullptr;
}  // namespace

const char* kClassNames[kProfClasses] = {"rotate", "rotate_adjoint", "t_apply", "d_build", "sigma_gemm_mfma",
                                         "sigma_gather", "reduce", "copy", "tile_forward", "tile_reverse",
                                         "perm", "rdm_gram_mfma", "batch", "other",
                                         "sigma_rows_mfma", "alpha_gemm_mfma", "orbital_gemm_mfma"};

bool prof_on() { return prof().on; }

// Profiling off for a stretch of launches (e.g. while a graph is captured:
// its event records would become graph nodes); returns the previous state.
bool prof_pause() {
  Prof& P = prof();
  const bool was = P.on;
  P.on = false;
  return was;
}

void prof_resume(bool was) { prof().on = was; }

void prof_begin(int cls, hipStream_t st) {
  Prof& P = prof();
  if (!P.on || !((P.mask >> cls) & 1u)) return;
  std::lock_guard<std::mutex> g(P.mu);
  t_start = P.get();
  SQ_HIP(hipEventRecord(t_start, st));
}

void prof_end(int cls, hipStream_t st, double bytes, double flops) {
  Prof& P = prof();
  if (!P.on || !((P.mask >> cls) & 1u)) return;
  std::lock_guard<std::mutex> g(P.mu);
  hipEvent_t b = P.get();
  SQ_HIP(hipEventRecord(b, st));
  P.recs.push_back({cls, t_start, b, bytes, flops});
}

}  // namespace sq

using namespace sq;

extern "C" {

int sq_profile_enable(int on) {
  return guarded([&] {
    Prof& P = prof();
    std::lock_guard<std::mutex> g(P.mu);
    P.on = on != 0;
    for (auto& r : P.recs) {
      P.pool.push_back(r.a);
      P.pool.push_back(r.b);
    }
    P.recs.clear();
  });
}

int sq_profile_collect(int n_classes, double* total_ms, int64_t* count, double* bytes, double* flops) {
  return guarded([&] {
    SQ_CHECK(n_classes >= kProfClasses && total_ms && count && bytes && flops, SQ_ERR_ARG, "arrays too small");
    Prof& P = prof();
    std::lock_guard<std::mutex> g(P.mu);
    for (int c = 0; c < n_classes; ++c) total_ms[c] = 0.0, count[c] = 0, bytes[c] = 0.0, flops[c] = 0.0;
    for (auto& r : P.recs) {
      SQ_HIP(hipEventSynchronize(r.b));
      float ms = 0.f;
      SQ_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      total_ms[r.cls] += ms;
      count[r.cls] += 1;
      bytes[r.cls] += r.bytes;
      flops[r.cls] += r.flops;
      P.pool.push_back(r.a);
      P.pool.push_back(r.b);
    }
    P.recs.clear();
  });
}

const char* sq_profile_class_name(int cls) { return (cls >= 0 && cls < kProfClasses) ? kClassNames[cls] : ""; }

int sq_profile_select(uint32_t class_mask) {
  return guarded([&] {
    Prof& P = prof();
    std::lock_guard<std::mutex> g(P.mu);
    P.mask = class_mask;
  });
}

int sq_profile_nclasses(void) { return kProfClasses; }

}  // extern "C"
