// Weight layouts built on the device from the caller's arrays (float or
// double, row-major), so a run whose weights already sit in device memory
// needs no host round trip to set up.
#include <algorithm>

#include "am_kernels.h"

namespace am {

__device__ __forceinline__ double src_at(const void* p, int f64, size_t i) {
  return f64 ? static_cast<const double*>(p)[i] : (double)static_cast<const float*>(p)[i];
}

// dst[j * ni + i] = src[i * cols + j] for i < rows, j < cols; +0 elsewhere
// (i < ni, j < nj).  A plain copy is the case cols = nj = 1.
__global__ void k_transpose_pad_f64(const void* src, int f64, int rows, int cols, double* dst, int ni, int nj) {
  const size_t n = (size_t)ni * nj;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(idx / ni), i = (int)(idx - (size_t)j * ni);
    dst[idx] = (i < rows && j < cols) ? src_at(src, f64, (size_t)i * cols + j) : 0.0;
  }
}

__global__ void k_transpose_pad_f32(const void* src, int f64, int rows, int cols, float* dst, int ni, int nj) {
  const size_t n = (size_t)ni * nj;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(idx / ni), i = (int)(idx - (size_t)j * ni);
    dst[idx] = (i < rows && j < cols) ? (float)src_at(src, f64, (size_t)i * cols + j) : 0.f;
  }
}

void launch_transpose_pad(const void* src, int f64, int rows, int cols, double* dst, int ni, int nj,
                          hipStream_t st) {
  const size_t n = (size_t)ni * nj;
  if (n == 0) return;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_transpose_pad_f64, dim3(blocks), dim3(256), 0, st, src, f64, rows, cols, dst, ni, nj);
}

void launch_transpose_pad(const void* src, int f64, int rows, int cols, float* dst, int ni, int nj,
                          hipStream_t st) {
  const size_t n = (size_t)ni * nj;
  if (n == 0) return;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_transpose_pad_f32, dim3(blocks), dim3(256), 0, st, src, f64, rows, cols, dst, ni, nj);
}

}  // namespace am
