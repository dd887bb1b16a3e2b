// Reduced-precision plane path (FP32 MFMA / split-BF16 MFMA) with FP64
// re-decision of regions whose predicates fall inside the error bound.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "am_kernels.h"

namespace am {

struct FastNet {
  int precision = 0;
  bool ready = false;
};

struct FastBuffers {
  void* mem[16] = {nullptr};
  size_t bytes[16] = {0};
};

struct FastLevel {
  int first, count;
  const uint32_t* reg_mask;
  const int* reg_start;
  const double* reg_point;
  int* reg_info;
  int* reg_tab;
  double* reg_verts;
  int* reg_vpair;
  uint32_t* cand_mask;
  int* cand_start;
  double* cand_point;
};

struct FastTimes {
  float gemm_ms, walk_ms;
  int launches;
};

void fast_prepare(FastNet& f, FastBuffers& b, const NetDev& net, int precision, const void* const* weights,
                  const void* const* biases, int dtype, hipStream_t st);
int fast_level(FastNet& f, FastBuffers& b, const NetDev& net, const FastLevel& lvl, hipStream_t st,
               hipEvent_t e0, hipEvent_t e1, FastTimes& t);
void redo_fp64(FastBuffers& b, const NetDev& net, const FastLevel& lvl, int nredo, hipStream_t st, hipEvent_t e0,
               hipEvent_t e1, float* ms, size_t lds, int Cs, int Csort);
void fast_release(FastBuffers& b);

}  // namespace am
