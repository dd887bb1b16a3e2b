// Adjoint of the mesh vertices with respect to the network parameters.
//
// A mesh vertex v solves A v + b = 0 for the three tight planes of the region
// that owns it (two neurons lo, hi and the output f = iso), each plane an
// affine function of the weights through the masked chain with the region's
// activation pattern held fixed.  With g = dL/dv and lambda = A^{-T} g,
//   dL/dP_i = -lambda_i [v; 1]   for each tight plane P_i = (a_i, c_i).
// These plane gradients are pulled back through the chain
//   P_{l+1,j} = sum_k W_l[j,k] m_k P_{l,k} + (0, 0, 0, b_l[j]),
// giving per region G_{l+1} (gradients of layer l+1 planes) and the masked
// inputs P_l; then dW_l = sum_regions G_{l+1}^T P_l is one GEMM over all
// regions (K = 4 x regions) and db_l the sum of the c components.
// The reference has no differentiation; this follows from the implicit
// function theorem applied to the reference's own vertex definition
// (backend/inc/kernel.h:555-573 solves the same 3x3 system).
#include "am_adjoint.h"
#include "am_chain.h"
#include "am_device.h"

namespace am {

__global__ __launch_bounds__(64) void k_adjoint_fp64(NetDev net, AdjointIO io) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int L = net.L;
  double* Px = reinterpret_cast<double*>(smem_raw);
  double* Py = Px + L;
  double* Pz = Py + L;
  double* Pc = Pz + L;
  double* Gx = Pc + L;
  double* Gy = Gx + L;
  double* Gz = Gy + L;
  double* Gc = Gz + L;
  uint32_t* smask = reinterpret_cast<uint32_t*>(Gc + L);
  __shared__ double s_eq[4];
  __shared__ double g_eq[4];
  __shared__ int s_any;

  const int lane = threadIdx.x;
  const int b = blockIdx.x;
  if (b >= io.count) return;
  const int r = io.r0 + b;
  const int vm = io.vm;
  const int nv = io.info[(size_t)r * AM_INFO_STRIDE + 2];

  if (lane == 0) s_any = 0;
  __syncthreads();
  if (lane < vm && lane < nv && nv >= 3 && io.vfirst[(size_t)r * vm + lane]) s_any = 1;
  __syncthreads();
  const bool active = s_any != 0;

  for (int w = lane; w < net.Wd; w += 64) smask[w] = io.reg_mask[(size_t)r * net.Wd + w];
  for (int g = lane; g < L; g += 64) { Gx[g] = 0.0; Gy[g] = 0.0; Gz[g] = 0.0; Gc[g] = 0.0; }
  if (lane < 4) g_eq[lane] = 0.0;
  __syncthreads();
  if (active) {
    chain_fp64(net, smask, Px, Py, Pz, Pc, s_eq);
    if (lane == 0) s_eq[3] = s_eq[3] - net.iso;
    __syncthreads();
    if (lane == 0) {
      const double eq[4] = {s_eq[0], s_eq[1], s_eq[2], s_eq[3]};
      for (int k = 0; k < nv; ++k) {
        const size_t q = (size_t)r * vm + k;
        if (!io.vfirst[q]) continue;
        const int lo = io.vpair[q * 2 + 0], hi = io.vpair[q * 2 + 1];
        const int id = io.vid[q];
        const double g3[3] = {io.gradv[(size_t)id * 3 + 0], io.gradv[(size_t)id * 3 + 1],
                              io.gradv[(size_t)id * 3 + 2]};
        double p1[4], p2[4];
        if (lo < L) { p1[0] = Px[lo]; p1[1] = Py[lo]; p1[2] = Pz[lo]; p1[3] = Pc[lo]; }
        else { const int e = lo - L; p1[0] = net.ew[e * 3]; p1[1] = net.ew[e * 3 + 1]; p1[2] = net.ew[e * 3 + 2]; p1[3] = net.eb[e]; }
        if (hi < L) { p2[0] = Px[hi]; p2[1] = Py[hi]; p2[2] = Pz[hi]; p2[3] = Pc[hi]; }
        else { const int e = hi - L; p2[0] = net.ew[e * 3]; p2[1] = net.ew[e * 3 + 1]; p2[2] = net.ew[e * 3 + 2]; p2[3] = net.eb[e]; }
        double v[3];
        solve3(p1, p2, eq, v);
        // lambda = A^{-T} g: rows of A^T are the columns of A = [p1; p2; eq]
        const double t1[4] = {p1[0], p2[0], eq[0], -g3[0]};
        const double t2[4] = {p1[1], p2[1], eq[1], -g3[1]};
        const double t3[4] = {p1[2], p2[2], eq[2], -g3[2]};
        double lam[3];
        solve3(t1, t2, t3, lam);
        const double vh[4] = {v[0], v[1], v[2], 1.0};
        if (lo < L) { Gx[lo] -= lam[0] * vh[0]; Gy[lo] -= lam[0] * vh[1]; Gz[lo] -= lam[0] * vh[2]; Gc[lo] -= lam[0]; }
        if (hi < L) { Gx[hi] -= lam[1] * vh[0]; Gy[hi] -= lam[1] * vh[1]; Gz[hi] -= lam[1] * vh[2]; Gc[hi] -= lam[1]; }
        g_eq[0] -= lam[2] * vh[0]; g_eq[1] -= lam[2] * vh[1]; g_eq[2] -= lam[2] * vh[2]; g_eq[3] -= lam[2];
      }
    }
    __syncthreads();
  }
  // pull back layer by layer, staging G_{li+1} and the masked inputs P_li
  for (int li = net.D; li >= 1; --li) {
    const int nin = net.nodes[li], nout = net.nodes[li + 1], in0 = net.off[li - 1];
    const bool last = li == net.D;
    const int out0 = last ? 0 : net.off[li];
    double* gs = io.Gst[li] + (size_t)b * 4 * nout;
    double* ps = io.Pst[li] + (size_t)b * 4 * nin;
    for (int j = lane; j < nout; j += 64) {
      if (!active) { gs[j] = 0.0; gs[nout + j] = 0.0; gs[2 * nout + j] = 0.0; gs[3 * nout + j] = 0.0; continue; }
      gs[j] = last ? g_eq[0] : Gx[out0 + j];
      gs[nout + j] = last ? g_eq[1] : Gy[out0 + j];
      gs[2 * nout + j] = last ? g_eq[2] : Gz[out0 + j];
      gs[3 * nout + j] = last ? g_eq[3] : Gc[out0 + j];
    }
    for (int k = lane; k < nin; k += 64) {
      const bool m = active && mbit(smask, in0 + k);
      ps[k] = m ? Px[in0 + k] : 0.0;
      ps[nin + k] = m ? Py[in0 + k] : 0.0;
      ps[2 * nin + k] = m ? Pz[in0 + k] : 0.0;
      ps[3 * nin + k] = m ? Pc[in0 + k] : 0.0;
    }
    if (active) {
      const double* W = net.Wrow[li];   // (nout, nin) row-major
      for (int k = lane; k < nin; k += 64) {
        if (!mbit(smask, in0 + k)) continue;
        double ax = 0.0, ay = 0.0, az = 0.0, ac = 0.0;
        for (int j = 0; j < nout; ++j) {
          const double w = W[(size_t)j * nin + k];
          const double gx = last ? g_eq[0] : Gx[out0 + j];
          const double gy = last ? g_eq[1] : Gy[out0 + j];
          const double gz = last ? g_eq[2] : Gz[out0 + j];
          const double gc = last ? g_eq[3] : Gc[out0 + j];
          ax = fmaT(w, gx, ax);
          ay = fmaT(w, gy, ay);
          az = fmaT(w, gz, az);
          ac = fmaT(w, gc, ac);
        }
        Gx[in0 + k] += ax;
        Gy[in0 + k] += ay;
        Gz[in0 + k] += az;
        Gc[in0 + k] += ac;
      }
    }
    __syncthreads();
  }
  double* g0 = io.Gst[0] + (size_t)b * 4 * net.nodes[1];
  for (int j = lane; j < net.nodes[1]; j += 64) {
    g0[j] = active ? Gx[j] : 0.0;
    g0[net.nodes[1] + j] = active ? Gy[j] : 0.0;
    g0[2 * net.nodes[1] + j] = active ? Gz[j] : 0.0;
    g0[3 * net.nodes[1] + j] = active ? Gc[j] : 0.0;
  }
}

// dW[j][k] += sum over kk < K of G[kk][j] * P[kk][k]  (G: [K][M], P: [K][N]),
// 16 x 16 output tile per workgroup, K split over blockIdx.z, FP64.
__global__ __launch_bounds__(256) void k_dw_gemm(const double* G, const double* P, int K, int M, int N,
                                                 double* dW, int kchunk) {
  __shared__ double sg[32][16];
  __shared__ double sp[32][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int j = blockIdx.y * 16 + ty, k = blockIdx.x * 16 + tx;
  const int k0 = blockIdx.z * kchunk, k1 = min(K, k0 + kchunk);
  double acc = 0.0;
  for (int kb = k0; kb < k1; kb += 32) {
    for (int i = threadIdx.x; i < 32 * 16; i += 256) {
      const int rr = i >> 4, cc = i & 15;
      const int kk = kb + rr;
      const int jj = blockIdx.y * 16 + cc, kk2 = blockIdx.x * 16 + cc;
      sg[rr][cc] = (kk < k1 && jj < M) ? G[(size_t)kk * M + jj] : 0.0;
      sp[rr][cc] = (kk < k1 && kk2 < N) ? P[(size_t)kk * N + kk2] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < 32; ++rr) acc = fmaT(sg[rr][ty], sp[rr][tx], acc);
    __syncthreads();
  }
  if (j < M && k < N) atomicAdd(&dW[(size_t)j * N + k], acc);
}

// out[j] += sum over b of G[b][q][j] for q = qsel (bias / first-layer columns).
__global__ void k_col_sum(const double* G, int count, int M, int qsel, double* out, int stride_out, int bchunk) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const int b0 = blockIdx.y * bchunk, b1 = min(count, b0 + bchunk);
  double acc = 0.0;
  for (int b = b0; b < b1; ++b) acc += G[((size_t)b * 4 + qsel) * M + j];
  atomicAdd(&out[(size_t)j * stride_out], acc);
}

}  // namespace am
