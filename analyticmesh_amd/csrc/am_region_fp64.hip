// FP64 region kernel: one wave64 per visited region.
//
// For one activation region it builds every constraint plane through the
// masked affine chain, orders the constraints by their signed distance to the
// region's seed point, finds the start plane (seed regions only), walks the
// polygon of f = iso inside the region edge by edge, orients it, writes the
// region's record and emits one neighbour candidate per crossed hidden-neuron
// edge.  The semantics are the reference's (backend/inc/process.h:18-264,
// backend/inc/kernel.h:799-1506) with deterministic choices where the
// reference races: the lowest feasible pair index for the start search and
// the lowest feasible candidate position in the walk.
//
// LDS holds the planes as four SoA arrays (x, y, z, c) so that a lane-strided
// sweep over constraints is bank-conflict free and a uniform read of one
// plane is a broadcast.
#include "am_device.h"
#include "am_kernels.h"
#include "am_walk.h"
#include "am_chain.h"

namespace am {

__global__ __launch_bounds__(64) void k_region_fp64(NetDev net, RegionIO io) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = net.C, L = net.L;
  const int Cs = io.Cs;        // stride of one SoA plane array (>= C)
  const int N2 = io.Csort;     // power of two >= max(C, 64)
  double* Px = reinterpret_cast<double*>(smem_raw);
  double* Py = Px + Cs;
  double* Pz = Py + Cs;
  double* Pc = Pz + Cs;
  double* keys = Pc + Cs;
  int* order = reinterpret_cast<int*>(keys + N2);
  uint32_t* used = reinterpret_cast<uint32_t*>(order + N2);
  uint32_t* smask = used + ((C + 31) >> 5);

  __shared__ double s_eq[4];

  const int lane = threadIdx.x;
  if ((int)blockIdx.x >= io.count) return;
  const int rid = io.ids ? io.ids[blockIdx.x] : io.first + blockIdx.x;

  for (int w = lane; w < net.Wd; w += 64) smask[w] = io.reg_mask[(size_t)rid * net.Wd + w];
  for (int w = lane; w < ((C + 31) >> 5); w += 64) used[w] = 0u;
  double p[3];
  p[0] = io.reg_point[(size_t)rid * 3 + 0];
  p[1] = io.reg_point[(size_t)rid * 3 + 1];
  p[2] = io.reg_point[(size_t)rid * 3 + 2];

  chain_fp64(net, smask, Px, Py, Pz, Pc, s_eq);
  if (lane == 0) s_eq[3] = s_eq[3] - net.iso;
  // sign flip (backend/inc/process.h:181-186) and extra planes (:189-192)
  for (int g = lane; g < L; g += 64) {
    if (mbit(smask, g)) { Px[g] = -Px[g]; Py[g] = -Py[g]; Pz[g] = -Pz[g]; Pc[g] = -Pc[g]; }
  }
  for (int e = lane; e < net.E; e += 64) {
    Px[L + e] = net.ew[e * 3 + 0];
    Py[L + e] = net.ew[e * 3 + 1];
    Pz[L + e] = net.ew[e * 3 + 2];
    Pc[L + e] = net.eb[e];
  }
  __syncthreads();
  const double eq[4] = {s_eq[0], s_eq[1], s_eq[2], s_eq[3]};

  const WalkParams wp{0.f, 0.f, 0.f};
  walk_region<double, false>(net, io, wp, rid, rid - io.first, Px, Py, Pz, Pc, keys, order, used, smask, eq, p);
}

}  // namespace am
