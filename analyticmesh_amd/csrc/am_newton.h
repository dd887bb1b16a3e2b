// FP64 Newton step for vertices solved from reduced-precision planes (am_newton.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "am_kernels.h"

namespace am {

#define AM_NT 32                   // vertices per Newton workgroup (one column each)
#define AM_NEWTON_THREADS 512      // 8 waves, 32 rows each: rows 0 .. 255 of a layer

struct NewtonIO {
  const uint32_t* reg_mask;        // [R][Wd]
  const int* vpair;                // [R][vm][2]
  double* verts;                   // [R][vm][3], refined in place
  const uint16_t* jinv;            // [R][vm][9] bf16 (the walk's inverse systems)
  const int* items;                // [*count] vertex uses to refine (first use of each vertex)
  const int* count;                // device count of items
  const int* item_region;          // vertex use -> region (the combine's item map)
  const unsigned long long* roff;  // region -> faces before << 32 | uses before
};

// LDS of k_newton_fp64 for a network with Wd mask words.
inline size_t newton_lds_bytes(int Wd) {
  return (size_t)AM_CHAIN_MAXW * AM_NT * sizeof(double) + (size_t)AM_NT * (Wd | 1) * 4;
}

// Grid: one workgroup per AM_NT items of the host's upper bound; blocks past
// the device count exit.  cio supplies the padded FP64 weights (chain layout).
__global__ void k_newton_fp64(NetDev net, ChainIO cio, NewtonIO io);
__global__ void k_newton_scatter(NewtonIO io, const uint32_t* table, const int* slot, const unsigned* vflag,
                                 const int* info, int N, int vm);

}  // namespace am
