// Deferred stage timing on one HIP stream: events are recorded between
// launches without waiting, and resolved once when the stream has drained,
// so timing adds no host round trip to a level.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace am {

enum TimerSlot { T_CHAIN64 = 0, T_RAY64, T_SCAN64, T_REGION64, T_GEMM, T_WALK, T_INSERT, T_VERT64, T_NEWTON, T_SLOTS };

struct StageTimer {
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  struct Span {
    size_t a, b;
    int slot;
  };
  std::vector<Span> spans;
  double acc[T_SLOTS] = {0};
  bool enabled = true;   // off while launches are captured into a graph (no event nodes)

  static void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e) + " in " + what);
  }
  // Record an event on `st`; returns its index for span().
  size_t mark(hipStream_t st) {
    if (!enabled) return 0;
    if (used == pool.size()) {
      hipEvent_t e;
      check(hipEventCreate(&e), "hipEventCreate");
      pool.push_back(e);
    }
    check(hipEventRecord(pool[used], st), "hipEventRecord");
    return used++;
  }
  void span(size_t a, size_t b, int slot) {
    if (enabled) spans.push_back({a, b, slot});
  }
  // The stream must have drained (or the events completed).
  void resolve() {
    for (const Span& s : spans) {
      float ms = 0;
      check(hipEventElapsedTime(&ms, pool[s.a], pool[s.b]), "hipEventElapsedTime");
      acc[s.slot] += ms;
    }
    spans.clear();
    used = 0;
  }
  void reset() {
    spans.clear();
    used = 0;
    for (double& a : acc) a = 0;
  }
  void release() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
    reset();
  }
};

}  // namespace am
