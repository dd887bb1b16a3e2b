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
  // vertex uses ("items") of the combined mesh, numbered region by region:
  // item of (region r, vertex k) = (int)item_off[r] + k
  const unsigned long long* item_off;  // [R] low 32 bits: first item of the region
  const unsigned* vfirst;     // [items] 1 where the item is its vertex's first occurrence
  const int* vid;             // [items] mesh vertex index of the item
  const double* gradv;        // [V][3] dL/dv
  int r0, count, vm;
  double* Gst[AM_MAX_LINEAR]; // li = 0..D: [count][4][n_{li+1}]
  double* Pst[AM_MAX_LINEAR]; // li = 1..D: [count][4][n_li]
  // tiled path: each region's direct terms, one entry per distinct hidden
  // neuron its vertices make tight: count [count], neuron [count][dcap],
  // values [count][dcap][4] (x, y, z, c)
  int* dn;
  int* dg;
  double* dv;
  int dcap;
};

__global__ void k_adjoint_fp64(NetDev net, AdjointIO io);
// planes / eq: the tiled chain's store of regions r0 .. r0 + count - 1
__global__ void k_adjoint_direct(NetDev net, AdjointIO io, const double* planes, const double* eq, int Cs);
template <bool TA, int MODE>
__global__ void k_adj_gemm(const double* A, int lda, const double* B, int ldb, double* C, int ldc, int M, int N, int K,
                           int kchunk, const uint32_t* mask, int Wd, int n0);
template <bool TA, int MODE>
__global__ void k_adj_mfma(const double* A, int lda, const double* B, int ldb, double* C, int ldc, int M, int N, int K,
                           int kchunk, const uint32_t* mask, int Wd, int n0);
__global__ void k_adj_scatter(const int* dn, const int* dg, const double* dv, int count, int cap, int o, int n,
                              double* G);
// Column sums of G over regions, one partial per block of 256 regions
// (k_col_sum: component qsel, part [blocks][M]; k_col_sum4: all four, part
// [blocks][4][M]), and the ordered sum of partial slices into an output.
__global__ void k_col_sum(const double* G, int count, int M, int qsel, double* part);
__global__ void k_col_sum4(const double* G, int count, int M, double* part);
__global__ void k_slice_sum(const double* part, int nslice, size_t sstride, int n, double* out, int ostride);

}  // namespace am
