// Adjoint kernels (device side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "am_limits.h"

namespace am {

struct AdjointIO {
  const uint32_t* reg_mask;   // [R][Wd]
  const int* info;            // [R][AM_INFO_STRIDE]
  const int* vpair;           // [R][vm][2]
  const int* vfirst;          // [R*vm] 1 where the item is its vertex's first occurrence
  const int* vid;             // [R*vm] mesh vertex index of the item
  const double* gradv;        // [V][3] dL/dv
  int r0, count, vm;
  double* Gst[AM_MAX_LINEAR]; // li = 0..D: [count][4][n_{li+1}]
  double* Pst[AM_MAX_LINEAR]; // li = 1..D: [count][4][n_li]
};

__global__ void k_adjoint_fp64(NetDev net, AdjointIO io);
__global__ void k_dw_gemm(const double* G, const double* P, int K, int M, int N, double* dW, int kchunk);
__global__ void k_col_sum(const double* G, int count, int M, int qsel, double* out, int stride_out, int bchunk);

}  // namespace am
