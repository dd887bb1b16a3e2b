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

  const int item0 = (int)(io.item_off[r] & 0xffffffffu);
  if (lane == 0) s_any = 0;
  __syncthreads();
  if (lane < vm && lane < nv && nv >= 3 && io.vfirst[item0 + lane]) s_any = 1;
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
        if (!io.vfirst[item0 + k]) continue;
        const int lo = io.vpair[q * 2 + 0], hi = io.vpair[q * 2 + 1];
        const int id = io.vid[item0 + k];
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

// The raw (unsigned) plane of constraint g from the tiled chain's store of
// one region: hidden layer 1 is (W0 | b0); deeper layers are stored negated
// where the neuron is active; extra planes are (ew | eb).
__device__ __forceinline__ void raw_plane64(const NetDev& net, const uint32_t* mask, const double* src, int Cs, int g,
                                            double* p) {
  if (g < net.nodes[1]) {
    p[0] = net.W0[g * 3 + 0]; p[1] = net.W0[g * 3 + 1]; p[2] = net.W0[g * 3 + 2]; p[3] = net.b[0][g];
  } else if (g < net.L) {
    const bool on = mbit(mask, g);
    for (int q = 0; q < 4; ++q) {
      const double v = src[(size_t)q * Cs + g];
      p[q] = on ? -v : v;
    }
  } else {
    const int e = g - net.L;
    p[0] = net.ew[e * 3 + 0]; p[1] = net.ew[e * 3 + 1]; p[2] = net.ew[e * 3 + 2]; p[3] = net.eb[e];
  }
}

// The direct terms of one region (one wave): lambda for each vertex the
// region owns, one vertex per lane, the plane gradients -lambda_i [v; 1]
// summed in vertex order by lane 0 (the order of k_adjoint_fp64).  Written
// out: G of the output plane, the sums as a list per distinct tight hidden
// neuron, and the masked inputs P.  The pull-back through the weights is left
// to the GEMMs over all regions, and k_adj_scatter adds the lists to it.
__global__ __launch_bounds__(64) void k_adjoint_direct(NetDev net, AdjointIO io, const double* planes,
                                                       const double* eqb, int Cs) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int L = net.L;
  double* G = reinterpret_cast<double*>(smem_raw);   // [4][L]
  int* mark = reinterpret_cast<int*>(G + 4 * L);      // [L]: neuron already in the list
  __shared__ int s_list[2 * AM_VM_MAX];
  __shared__ int s_nd;
  __shared__ double s_c[AM_VM_MAX][12];
  __shared__ int s_lohi[AM_VM_MAX][2];
  __shared__ int s_own[AM_VM_MAX];
  __shared__ double g_eq[4];
  const int lane = threadIdx.x;
  const int b = blockIdx.x;
  if (b >= io.count) return;
  const int r = io.r0 + b;
  const int vm = io.vm;
  const int nv = io.info[(size_t)r * AM_INFO_STRIDE + 2];
  const uint32_t* mask = io.reg_mask + (size_t)r * net.Wd;
  const double* src = planes + (size_t)b * 4 * Cs;
  for (int g = lane; g < 4 * L; g += 64) G[g] = 0.0;
  for (int g = lane; g < L; g += 64) mark[g] = 0;
  const int item0 = (int)(io.item_off[r] & 0xffffffffu);
  const bool mine = lane < vm && lane < nv && nv >= 3 && io.vfirst[item0 + lane];
  if (lane < vm) s_own[lane] = mine ? 1 : 0;
  const bool active = __any(mine);
  if (mine) {
    const size_t q = (size_t)r * vm + lane;
    const int lo = io.vpair[q * 2 + 0], hi = io.vpair[q * 2 + 1];
    const int id = io.vid[item0 + lane];
    const double g3[3] = {io.gradv[(size_t)id * 3 + 0], io.gradv[(size_t)id * 3 + 1], io.gradv[(size_t)id * 3 + 2]};
    double p1[4], p2[4];
    raw_plane64(net, mask, src, Cs, lo, p1);
    raw_plane64(net, mask, src, Cs, hi, p2);
    const double eq[4] = {eqb[(size_t)b * 4 + 0], eqb[(size_t)b * 4 + 1], eqb[(size_t)b * 4 + 2],
                          eqb[(size_t)b * 4 + 3]};
    double v[3];
    solve3(p1, p2, eq, v);
    const double t1[4] = {p1[0], p2[0], eq[0], -g3[0]};
    const double t2[4] = {p1[1], p2[1], eq[1], -g3[1]};
    const double t3[4] = {p1[2], p2[2], eq[2], -g3[2]};
    double lam[3];
    solve3(t1, t2, t3, lam);
    for (int i = 0; i < 3; ++i) {
      s_c[lane][i * 4 + 0] = lam[i] * v[0];
      s_c[lane][i * 4 + 1] = lam[i] * v[1];
      s_c[lane][i * 4 + 2] = lam[i] * v[2];
      s_c[lane][i * 4 + 3] = lam[i];
    }
    s_lohi[lane][0] = lo;
    s_lohi[lane][1] = hi;
  }
  __syncthreads();
  if (lane == 0) {
    double ge[4] = {0.0, 0.0, 0.0, 0.0};
    int nd = 0;
    for (int k = 0; k < nv && k < vm; ++k) {
      if (!s_own[k]) continue;
      const int lo = s_lohi[k][0], hi = s_lohi[k][1];
      if (lo < L) {
        for (int c = 0; c < 4; ++c) G[c * L + lo] -= s_c[k][c];
        if (!mark[lo]) { mark[lo] = 1; s_list[nd++] = lo; }
      }
      if (hi < L) {
        for (int c = 0; c < 4; ++c) G[c * L + hi] -= s_c[k][4 + c];
        if (!mark[hi]) { mark[hi] = 1; s_list[nd++] = hi; }
      }
      for (int c = 0; c < 4; ++c) ge[c] -= s_c[k][8 + c];
    }
    for (int c = 0; c < 4; ++c) g_eq[c] = ge[c];
    s_nd = nd;
  }
  __syncthreads();
  // G of the output plane (one output); the hidden layers' direct terms as a
  // list, added to the pull-back's output by k_adj_scatter
  if (lane < 4) io.Gst[net.D][(size_t)b * 4 + lane] = active ? g_eq[lane] : 0.0;
  const int nd = s_nd;
  if (lane == 0) io.dn[b] = nd;
  for (int j = lane; j < nd; j += 64) {
    const int g = s_list[j];
    io.dg[(size_t)b * io.dcap + j] = g;
    for (int c = 0; c < 4; ++c) io.dv[((size_t)b * io.dcap + j) * 4 + c] = G[c * L + g];
  }
  // masked inputs of linear layers 1..D (zero for a region owning no vertex:
  // its G rows are zero)
  for (int li = 1; li <= net.D; ++li) {
    const int n = net.nodes[li], in0 = net.off[li - 1];
    double* ps = io.Pst[li] + (size_t)b * 4 * n;
    for (int k = lane; k < n; k += 64) {
      double p[4] = {0.0, 0.0, 0.0, 0.0};
      if (active && mbit(mask, in0 + k)) raw_plane64(net, mask, src, Cs, in0 + k, p);
      for (int c = 0; c < 4; ++c) ps[(size_t)c * n + k] = p[c];
    }
  }
}

// C (M x N) from A (M x K) and B (K x N), FP64: 64 x 64 outputs per block of
// 256 threads, 4 x 4 per thread (rows 4 ty .. 4 ty + 3, columns 2 tx + {0, 1}
// and 32 + 2 tx + {0, 1}), K in steps of 16 through LDS, summed in ascending k.
//   TA: element (m, k) of A at A[k * lda + m], else at A[m * lda + k];
//   B:  element (k, n) at B[k * ldb + n].
//   MODE 0 (pull-back): rows are (region, component), region m / 4 of the
//     tile's mask rows; C[m][n] = acc where that region's neuron n0 + n is
//     active, else +0.  K whole (one block in z).
//   MODE 1 (weight gradient): K split over blockIdx.z in chunks of kchunk;
//     block z stores its partial product in slice z of C (C + z * M * ldc),
//     which k_slice_sum then adds in ascending z: no atomics, so the sums and
//     the gradients are the same from run to run.
template <bool TA, int MODE>
__global__ __launch_bounds__(256) void k_adj_gemm(const double* __restrict__ A, int lda, const double* __restrict__ B,
                                                  int ldb, double* C, int ldc, int M, int N, int K, int kchunk,
                                                  const uint32_t* mask, int Wd, int n0) {
  constexpr int SA = 66;   // row stride: column stores of A spread over the banks, rows stay 16-byte aligned
  __shared__ __align__(16) double sa[16][SA];
  __shared__ __align__(16) double sb[16][64];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int m0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const int k0 = blockIdx.z * kchunk, k1 = min(K, k0 + kchunk);
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  // the next k block's operands are loaded into registers while this one is
  // multiplied from LDS
  double ra[4], rb[4];
  auto fetch = [&](int kb) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = tid + 256 * e;
      const int ka = TA ? (i >> 6) : (i & 15), ma = TA ? (i & 63) : (i >> 4);
      const int gk = kb + ka, gm = m0 + ma;
      ra[e] = (gk < k1 && gm < M) ? (TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk]) : 0.0;
      const int gk2 = kb + (i >> 6), gn = c0 + (i & 63);
      rb[e] = (gk2 < k1 && gn < N) ? B[(size_t)gk2 * ldb + gn] : 0.0;
    }
  };
  if (k0 < k1) fetch(k0);
  for (int kb = k0; kb < k1; kb += 16) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = tid + 256 * e;
      const int ka = TA ? (i >> 6) : (i & 15), ma = TA ? (i & 63) : (i >> 4);
      sa[ka][ma] = ra[e];
      sb[i >> 6][i & 63] = rb[e];
    }
    __syncthreads();
    if (kb + 16 < k1) fetch(kb + 16);
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      const double2 a01 = *reinterpret_cast<const double2*>(&sa[kk][4 * ty]);
      const double2 a23 = *reinterpret_cast<const double2*>(&sa[kk][4 * ty + 2]);
      const double2 b01 = *reinterpret_cast<const double2*>(&sb[kk][2 * tx]);
      const double2 b23 = *reinterpret_cast<const double2*>(&sb[kk][32 + 2 * tx]);
      const double a[4] = {a01.x, a01.y, a23.x, a23.y};
      const double bv[4] = {b01.x, b01.y, b23.x, b23.y};
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaT(a[i], bv[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + 4 * ty + i;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = c0 + (j >> 1) * 32 + 2 * tx + (j & 1);
      if (n >= N) continue;
      if (MODE == 0) {
        C[(size_t)m * ldc + n] = mbit(mask + (size_t)(m >> 2) * Wd, n0 + n) ? acc[i][j] : 0.0;
      } else {
        C[((size_t)blockIdx.z * M + m) * ldc + n] = acc[i][j];
      }
    }
  }
}

template __global__ void k_adj_gemm<false, 0>(const double*, int, const double*, int, double*, int, int, int, int, int,
                                              const uint32_t*, int, int);
template __global__ void k_adj_gemm<true, 1>(const double*, int, const double*, int, double*, int, int, int, int, int,
                                             const uint32_t*, int, int);

// The same products on the FP64 matrix cores: each wave owns a 32 x 32
// quarter of the 64 x 64 tile as 2 x 2 blocks of v_mfma_f64_16x16x4_f64
// (A fragment: lane l holds A[l % 16][l / 16]; B: B[l / 16][l % 16]; result:
// element v of lane l is row l / 16 + 4 v of column l % 16, as measured on
// gfx950 with one-hot operands).  Operands
// pass through LDS as in k_adj_gemm; TA, MODE and the arguments are the same.
typedef double am_double4 __attribute__((ext_vector_type(4)));

template <bool TA, int MODE>
__global__ __launch_bounds__(256) void k_adj_mfma(const double* __restrict__ A, int lda, const double* __restrict__ B,
                                                  int ldb, double* C, int ldc, int M, int N, int K, int kchunk,
                                                  const uint32_t* mask, int Wd, int n0) {
  constexpr int SA = 66;
  __shared__ __align__(16) double sa[16][SA];
  __shared__ __align__(16) double sb[16][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int m0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const int k0 = blockIdx.z * kchunk, k1 = min(K, k0 + kchunk);
  am_double4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = am_double4{0.0, 0.0, 0.0, 0.0};
  double ra[4], rb[4];
  auto fetch = [&](int kb) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = tid + 256 * e;
      const int ka = TA ? (i >> 6) : (i & 15), ma = TA ? (i & 63) : (i >> 4);
      const int gk = kb + ka, gm = m0 + ma;
      ra[e] = (gk < k1 && gm < M) ? (TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk]) : 0.0;
      const int gk2 = kb + (i >> 6), gn = c0 + (i & 63);
      rb[e] = (gk2 < k1 && gn < N) ? B[(size_t)gk2 * ldb + gn] : 0.0;
    }
  };
  const int l16 = lane & 15, l4 = lane >> 4;
  if (k0 < k1) fetch(k0);
  for (int kb = k0; kb < k1; kb += 16) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = tid + 256 * e;
      const int ka = TA ? (i >> 6) : (i & 15), ma = TA ? (i & 63) : (i >> 4);
      sa[ka][ma] = ra[e];
      sb[i >> 6][i & 63] = rb[e];
    }
    __syncthreads();
    if (kb + 16 < k1) fetch(kb + 16);
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int kr = ks * 4 + l4;
      const double a0 = sa[kr][wr + l16], a1 = sa[kr][wr + 16 + l16];
      const double b0 = sb[kr][wc + l16], b1 = sb[kr][wc + 16 + l16];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = c0 + wc + 16 * j + l16;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wr + 16 * i + l4 + 4 * v;
        if (m >= M || n >= N) continue;
        if (MODE == 0) {
          C[(size_t)m * ldc + n] = mbit(mask + (size_t)(m >> 2) * Wd, n0 + n) ? acc[i][j][v] : 0.0;
        } else {
          C[((size_t)blockIdx.z * M + m) * ldc + n] = acc[i][j][v];
        }
      }
    }
}

template __global__ void k_adj_mfma<false, 0>(const double*, int, const double*, int, double*, int, int, int, int, int,
                                              const uint32_t*, int, int);
template __global__ void k_adj_mfma<true, 1>(const double*, int, const double*, int, double*, int, int, int, int, int,
                                             const uint32_t*, int, int);

// Adds the direct terms of the hidden layer with neurons o .. o + n - 1 to
// its pull-back output G [count][4][n] (k_adjoint_direct's lists, one thread
// per entry; a region lists each neuron once, so no two threads meet).
// With G = mask (G_next W) this is the per-region kernel's direct + pull-back,
// as a + b and b + a are the same floating-point sum.
__global__ void k_adj_scatter(const int* dn, const int* dg, const double* dv, int count, int cap, int o, int n,
                              double* G) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = (int)(t / cap), j = (int)(t - (size_t)b * cap);
  if (b >= count || j >= dn[b]) return;
  const int g = dg[(size_t)b * cap + j] - o;
  if (g < 0 || g >= n) return;
  for (int c = 0; c < 4; ++c) G[((size_t)b * 4 + c) * n + g] += dv[((size_t)b * cap + j) * 4 + c];
}

// part[y][j] = sum over the 256 regions b of block y of G[b][qsel][j]: 64
// columns per block, four partial sums per column added in LDS in a fixed
// order.  k_slice_sum adds the blocks' partials.
__global__ __launch_bounds__(256) void k_col_sum(const double* G, int count, int M, int qsel, double* part) {
  __shared__ double ps[4][64];
  const int jx = threadIdx.x & 63, by = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + jx;
  const int b0 = blockIdx.y * 256, b1 = min(count, b0 + 256);
  double acc = 0.0;
  if (j < M)
    for (int b = b0 + by; b < b1; b += 4) acc += G[((size_t)b * 4 + qsel) * M + j];
  ps[by][jx] = acc;
  __syncthreads();
  if (by == 0 && j < M) part[(size_t)blockIdx.y * M + j] = ((ps[0][jx] + ps[1][jx]) + ps[2][jx]) + ps[3][jx];
}

// All four components at once (the first layer): part[y][q][j] = sum over the
// 256 regions b of block y of G[b][q][j].
__global__ __launch_bounds__(256) void k_col_sum4(const double* G, int count, int M, double* part) {
  __shared__ double ps[4][4][64];
  const int jx = threadIdx.x & 63, by = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + jx;
  const int b0 = blockIdx.y * 256, b1 = min(count, b0 + 256);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (j < M)
    for (int b = b0 + by; b < b1; b += 4)
      for (int q = 0; q < 4; ++q) acc[q] += G[((size_t)b * 4 + q) * M + j];
  for (int q = 0; q < 4; ++q) ps[by][q][jx] = acc[q];
  __syncthreads();
  if (by == 0 && j < M)
    for (int q = 0; q < 4; ++q)
      part[((size_t)blockIdx.y * 4 + q) * M + j] = ((ps[0][q][jx] + ps[1][q][jx]) + ps[2][q][jx]) + ps[3][q][jx];
}

// out[i * ostride] += the sum over slices z < nslice of part[z][i] (slice z at
// part + z * sstride), i < n: eight running sums over z = u mod 8, u = 0..7,
// then added pairwise in a fixed order (the loads of eight slices are in
// flight at once; the result does not depend on timing).
__global__ __launch_bounds__(256) void k_slice_sum(const double* part, int nslice, size_t sstride, int n, double* out,
                                                   int ostride) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int z = 0;
  for (; z + 8 <= nslice; z += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] += part[(size_t)(z + u) * sstride + i];
  }
  for (int u = 0; z + u < nslice; ++u) a[u] += part[(size_t)(z + u) * sstride + i];
  const double s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  out[(size_t)i * ostride] = out[(size_t)i * ostride] + s;
}

}  // namespace am
