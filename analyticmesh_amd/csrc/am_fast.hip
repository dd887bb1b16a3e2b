// Reduced-precision plane path: placeholder until the MFMA kernels land.
#include <stdexcept>

#include "am_fast.h"

namespace am {

void fast_prepare(FastNet& f, FastBuffers&, const NetDev&, int precision, const void* const*, const void* const*,
                  int, hipStream_t) {
  f.precision = precision;
  throw std::runtime_error("reduced-precision planes are not available in this build");
}

int fast_level(FastNet&, FastBuffers&, const NetDev&, const FastLevel&, hipStream_t, hipEvent_t, hipEvent_t,
               FastTimes& t) {
  t.gemm_ms = t.walk_ms = 0;
  t.launches = 0;
  throw std::runtime_error("reduced-precision planes are not available in this build");
}

void redo_fp64(FastBuffers&, const NetDev&, const FastLevel&, int, hipStream_t, hipEvent_t, hipEvent_t, float* ms,
               size_t, int, int) {
  *ms = 0;
}

void fast_release(FastBuffers& b) {
  for (int i = 0; i < 16; ++i)
    if (b.mem[i]) (void)hipFree(b.mem[i]);
}

}  // namespace am
