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

namespace am {

__global__ __launch_bounds__(64) void k_region_fp64(NetDev net, RegionIO io) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = net.C, L = net.L, vm = net.vm;
  const int Cs = io.Cs;        // stride of one SoA plane array (>= C)
  const int N2 = io.Csort;     // power of two >= C
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
  if (blockIdx.x >= io.count) return;
  const int rid = io.ids ? io.ids[blockIdx.x] : io.first + blockIdx.x;

  for (int w = lane; w < net.Wd; w += 64) smask[w] = io.reg_mask[(size_t)rid * net.Wd + w];
  for (int w = lane; w < ((C + 31) >> 5); w += 64) used[w] = 0u;
  double p[3];
  p[0] = io.reg_point[(size_t)rid * 3 + 0];
  p[1] = io.reg_point[(size_t)rid * 3 + 1];
  p[2] = io.reg_point[(size_t)rid * 3 + 2];

  // ---- planes: hidden node 1 is W0 itself
  for (int j = lane; j < net.nodes[1]; j += 64) {
    Px[j] = net.W0[j * 3 + 0];
    Py[j] = net.W0[j * 3 + 1];
    Pz[j] = net.W0[j * 3 + 2];
    Pc[j] = net.b[0][j];
  }
  __syncthreads();
  // ---- masked affine chain, one output neuron per lane and sub-step
  for (int li = 1; li <= net.D; ++li) {
    const int nin = net.nodes[li], nout = net.nodes[li + 1], in0 = net.off[li - 1];
    const double* WT = net.WT[li];
    const double* bb = net.b[li];
    const bool last = (li == net.D);
    const int dst0 = last ? 0 : net.off[li];
    for (int jb = 0; jb < nout; jb += 256) {
      double ax[4], ay[4], az[4], ac[4];
      int jj[4];
      bool ok[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        jj[t] = jb + lane + 64 * t;
        ok[t] = jj[t] < nout;
        ax[t] = 0.0; ay[t] = 0.0; az[t] = 0.0;
        ac[t] = ok[t] ? bb[jj[t]] : 0.0;
      }
      for (int k = 0; k < nin; ++k) {
        if (!mbit(smask, in0 + k)) continue;
        const double px = Px[in0 + k], py = Py[in0 + k], pz = Pz[in0 + k], pc = Pc[in0 + k];
        const double* wrow = WT + (size_t)k * nout;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (ok[t]) {
            const double w = wrow[jj[t]];
            ax[t] = fmaT(w, px, ax[t]);
            ay[t] = fmaT(w, py, ay[t]);
            az[t] = fmaT(w, pz, az[t]);
            ac[t] = fmaT(w, pc, ac[t]);
          }
        }
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (!ok[t]) continue;
        if (last) {
          s_eq[0] = ax[t]; s_eq[1] = ay[t]; s_eq[2] = az[t]; s_eq[3] = ac[t];
        } else {
          const int g = dst0 + jj[t];
          Px[g] = ax[t]; Py[g] = ay[t]; Pz[g] = az[t]; Pc[g] = ac[t];
        }
      }
    }
    __syncthreads();
    if (!last) {
      // additive skips into this node (backend/inc/process.h:73-146)
      for (int s = 0; s < net.nskip[li]; ++s) {
        const int from = net.skip_from[li][s], tm = net.skip_tm[li][s];
        const bool ident = net.tm_h[tm] == 0 && net.tm_w[tm] == 0;
        if (from == 0) {
          if (ident) {
            if (lane < 3 && lane < nout) {
              const int g = dst0 + lane;
              if (lane == 0) Px[g] = Px[g] + 1.0;
              if (lane == 1) Py[g] = Py[g] + 1.0;
              if (lane == 2) Pz[g] = Pz[g] + 1.0;
            }
          } else {
            const double* T = net.tm[tm];
            for (int j = lane; j < net.tm_h[tm]; j += 64) {
              const int g = dst0 + j;
              Px[g] = Px[g] + T[j * 3 + 0];
              Py[g] = Py[g] + T[j * 3 + 1];
              Pz[g] = Pz[g] + T[j * 3 + 2];
            }
          }
        } else {
          const int os = net.off[from - 1], dec = net.nodes[from];
          if (ident) {
            for (int j = lane; j < dec; j += 64) {
              if (!mbit(smask, os + j)) continue;
              const int g = dst0 + j;
              Px[g] = Px[g] + Px[os + j];
              Py[g] = Py[g] + Py[os + j];
              Pz[g] = Pz[g] + Pz[os + j];
              Pc[g] = Pc[g] + Pc[os + j];
            }
          } else {
            const double* TT = net.tmT[tm];
            const int h = net.tm_h[tm], w = net.tm_w[tm];
            for (int j = lane; j < h; j += 64) {
              const int g = dst0 + j;
              double a0 = Px[g], a1 = Py[g], a2 = Pz[g], a3 = Pc[g];
              for (int k = 0; k < w; ++k) {
                if (!mbit(smask, os + k)) continue;
                const double t = TT[(size_t)k * h + j];
                a0 = fmaT(t, Px[os + k], a0);
                a1 = fmaT(t, Py[os + k], a1);
                a2 = fmaT(t, Pz[os + k], a2);
                a3 = fmaT(t, Pc[os + k], a3);
              }
              Px[g] = a0; Py[g] = a1; Pz[g] = a2; Pc[g] = a3;
            }
          }
        }
        __syncthreads();
      }
    }
  }
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
