// FP64 planes of many regions at once, bit-identical to the one-region chain.
//
// The masked affine chain (backend/inc/process.h:18-193) is a GEMM per layer:
// Y = W_li * X with X = diag(m) P_{li-1}, one column per plane component (x, y,
// z, c) of each region.  A workgroup takes AM_CT regions (AM_CN columns); each
// thread owns 4 output rows x 8 columns and runs the whole k loop in ascending
// order with one fma per step, so every output is the same sequence of
// roundings as the canonical chain: accumulators start at (0, 0, 0, b_j + 0)
// and take fma(W[j][k], m_k ? P_k : +0).  An inactive k adds fma(w, +0, acc) =
// acc exactly (acc is never -0: every accumulator starts at +0 or a nonzero
// bias, and a sum of opposite-signed zeros rounds to +0), so zero-filling the
// inactive inputs reproduces skipping them.  W streams from L2 (coalesced
// rows of W^T, shared by the four waves); X and the outputs live in one LDS
// buffer: after the k loop every wave parks its accumulators there, then the
// tile is written to HBM signed (a.x + c <= 0 inside) and masked in place as
// the next layer's input.
//
// k_walk_fp64 is the polygon walk of the region kernel on those planes.
#include "am_device.h"
#include "am_kernels.h"
#include "am_walk.h"

namespace am {

__global__ __launch_bounds__(256) void k_chain_tile_fp64(NetDev net, ChainIO io) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int Kp = AM_CHAIN_MAXW;
  double* Xs = reinterpret_cast<double*>(smem_raw);                          // [Kp][AM_CXS]
  uint32_t* mk = reinterpret_cast<uint32_t*>(Xs + (size_t)Kp * AM_CXS);       // [AM_CT][Wd]
  const int r0 = blockIdx.x * AM_CT;
  const int nreg = min(AM_CT, io.count - r0);
  for (int idx = tid; idx < AM_CT * net.Wd; idx += 256) {
    const int t = idx / net.Wd, w = idx - t * net.Wd;
    uint32_t v = 0u;
    if (t < nreg) {
      const int rid = io.ids ? io.ids[r0 + t] : io.first + r0 + t;
      v = io.reg_mask[(size_t)rid * net.Wd + w];
    }
    mk[idx] = v;
  }
  __syncthreads();
  const size_t Cs = (size_t)io.Cs;
  // hidden node 1: the rows of W0 (unsigned into X, signed into HBM)
  const int n1 = net.nodes[1];
  for (int idx = tid; idx < n1 * AM_CN; idx += 256) {
    const int k = idx / AM_CN, col = idx - k * AM_CN, t = col >> 2, q = col & 3;
    const double val = q < 3 ? net.W0[k * 3 + q] : net.b[0][k];
    Xs[k * AM_CXS + col] = mbit(mk + t * net.Wd, k) ? val : 0.0;
  }
  for (int idx = tid; idx < AM_CN * n1; idx += 256) {
    const int col = idx / n1, k = idx - col * n1, t = col >> 2, q = col & 3;
    if (t < nreg) {
      const double val = q < 3 ? net.W0[k * 3 + q] : net.b[0][k];
      io.planes[((size_t)(r0 + t) * 4 + q) * Cs + k] = mbit(mk + t * net.Wd, k) ? -val : val;
    }
  }
  for (int idx = tid; idx < AM_CN * net.E; idx += 256) {
    const int col = idx / net.E, e = idx - col * net.E, t = col >> 2, q = col & 3;
    if (t < nreg) io.planes[((size_t)(r0 + t) * 4 + q) * Cs + net.L + e] = q < 3 ? net.ew[e * 3 + q] : net.eb[e];
  }
  __syncthreads();
  const int c0 = 8 * wave;            // this wave's columns: regions 2*wave, 2*wave + 1
  for (int li = 1; li <= net.D; ++li) {
    const int nin = net.nodes[li], nout = net.nodes[li + 1];
    const double* WT = net.WT[li];
    const double* bias = net.b[li];
    double acc[4][8];
    bool ok[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = lane + 64 * i;
      ok[i] = row < nout;
      const double bz = ok[i] ? bias[row] + 0.0 : 0.0;
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[i][j] = (j & 3) == 3 ? bz : 0.0;
    }
    const double* wk = WT + lane;
#pragma unroll 2
    for (int k = 0; k < nin; ++k) {
      const double* xr = Xs + k * AM_CXS + c0;
      double x[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = xr[j];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (ok[i]) {
          const double w = wk[(size_t)k * nout + 64 * i];
#pragma unroll
          for (int j = 0; j < 8; ++j) acc[i][j] = fmaT(w, x[j], acc[i][j]);
        }
      }
    }
    __syncthreads();                  // every wave is done reading this layer's input
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (!ok[i]) continue;
      double* yr = Xs + (lane + 64 * i) * AM_CXS + c0;
#pragma unroll
      for (int j = 0; j < 8; ++j) yr[j] = acc[i][j];
    }
    __syncthreads();
    if (li < net.D) {
      const int out0 = net.off[li];
      for (int idx = tid; idx < AM_CN * nout; idx += 256) {
        const int col = idx / nout, j = idx - col * nout, t = col >> 2, q = col & 3;
        if (t < nreg) {
          const double y = Xs[j * AM_CXS + col];
          io.planes[((size_t)(r0 + t) * 4 + q) * Cs + out0 + j] = mbit(mk + t * net.Wd, out0 + j) ? -y : y;
        }
      }
      __syncthreads();
      for (int idx = tid; idx < nout * AM_CN; idx += 256) {
        const int j = idx / AM_CN, col = idx - j * AM_CN, t = col >> 2;
        if (!mbit(mk + t * net.Wd, out0 + j)) Xs[j * AM_CXS + col] = 0.0;
      }
    } else if (tid < AM_CN) {
      const int t = tid >> 2, q = tid & 3;
      if (t < nreg) {
        const double v = Xs[tid];
        io.eq[(size_t)(r0 + t) * 4 + q] = q == 3 ? v - net.iso : v;
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(64) void k_walk_fp64(NetDev net, RegionIO io, int rid0, const double* planes,
                                                  const double* eqb, int Cs_hbm) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = net.C;
  const int Cs = io.Cs;
  const int N2 = io.Csort;
  double* Px = reinterpret_cast<double*>(smem_raw);
  double* Py = Px + Cs;
  double* Pz = Py + Cs;
  double* Pc = Pz + Cs;
  double* keys = Pc + Cs;
  int* order = reinterpret_cast<int*>(keys + N2);
  uint32_t* used = reinterpret_cast<uint32_t*>(order + N2);
  uint32_t* smask = used + ((C + 31) >> 5);

  const int lane = threadIdx.x;
  if (blockIdx.x >= io.count) return;
  const int rid = io.ids ? io.ids[blockIdx.x] : rid0 + blockIdx.x;
  for (int w = lane; w < net.Wd; w += 64) smask[w] = io.reg_mask[(size_t)rid * net.Wd + w];
  for (int w = lane; w < ((C + 31) >> 5); w += 64) used[w] = 0u;
  const double* src = planes + (size_t)blockIdx.x * 4 * Cs_hbm;
  for (int g = lane; g < C; g += 64) {
    Px[g] = src[g];
    Py[g] = src[Cs_hbm + g];
    Pz[g] = src[2 * (size_t)Cs_hbm + g];
    Pc[g] = src[3 * (size_t)Cs_hbm + g];
  }
  double p[3];
  p[0] = io.reg_point[(size_t)rid * 3 + 0];
  p[1] = io.reg_point[(size_t)rid * 3 + 1];
  p[2] = io.reg_point[(size_t)rid * 3 + 2];
  const double eq[4] = {eqb[(size_t)blockIdx.x * 4 + 0], eqb[(size_t)blockIdx.x * 4 + 1],
                        eqb[(size_t)blockIdx.x * 4 + 2], eqb[(size_t)blockIdx.x * 4 + 3]};
  __syncthreads();
  const WalkParams wp{0.f, 0.f, 0.f};
  walk_region<double, false>(net, io, wp, rid, rid - io.first, Px, Py, Pz, Pc, keys, order, used, smask, eq, p);
}

}  // namespace am
