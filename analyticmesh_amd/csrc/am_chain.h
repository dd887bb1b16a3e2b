// FP64 masked affine chain of one region, shared by the forward region
// kernel and the adjoint kernel (backend/inc/process.h:18-146).  One wave64;
// the unsigned pre-activation planes of every hidden neuron land in LDS as
// SoA arrays and the output plane in s_eq.  Canonical order: for each output
// neuron the accumulators start at (0, 0, 0, b_j + 0) and take fma(W[j][k], P_k)
// over active k ascending; skips are added afterwards in connection order.
#pragma once
#include "am_device.h"

namespace am {

__device__ __forceinline__ void chain_fp64(const NetDev& net, const uint32_t* smask, double* Px, double* Py,
                                           double* Pz, double* Pc, double* s_eq) {
  const int lane = threadIdx.x;
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
        ac[t] = ok[t] ? bb[jj[t]] + 0.0 : 0.0;   // -0 bias -> +0: the tiled chain zero-fills inactive inputs
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
}

}  // namespace am
