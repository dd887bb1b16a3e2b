// One FP64 Newton step for mesh vertices a reduced-precision walk solved from
// its own planes.
//
// A mesh vertex v of region r is the solution of three equations that are
// affine in x with r's activation pattern held fixed: the pre-activations of
// its two tight neurons a and b are zero and f(x) = iso.  The walk solves them
// from the reduced planes (FP64 solve, plane coefficients from FP32 sums) and
// keeps J^-1, the inverse of that 3x3 system, per vertex.  Here each vertex
// takes one step
//   v' = v - J^-1 [s_a z_a(v), s_b z_b(v), f(v) - iso],
// the residuals from a forward pass of the network in FP64 under r's mask
// (s = -1 for an active neuron: the walk's planes are signed).  The step
// leaves an error of about |J^-1 dJ| |v - v*| (dJ the reduced planes' error,
// ~1e-6 relative) on top of the FP64 residual's rounding, so the vertex carries
// FP64-solve accuracy without FP64 planes (backend/inc/kernel.h:555-573 solves
// the same system from FP64 planes).
//
// One workgroup of 8 waves takes 32 vertices (one column each); wave w owns
// output rows 32 w .. 32 w + 31 of every layer, as two 16-row blocks by two
// 16-column blocks.  A layer's input (masked) lives in LDS; the products run
// on the FP64 matrix cores, the A fragments (weights) streamed from memory a
// few steps ahead.  The LDS (about 70 KB) lets two workgroups share a CU, so
// one's products overlap the other's layer epilogue.  The output layer (one
// row) is a VALU dot product split over the waves.
#include "am_device.h"
#include "am_kernels.h"
#include "am_newton.h"

namespace am {

typedef double am_f64x4 __attribute__((ext_vector_type(4)));

// Element (k, c) of a layer's input in LDS: rows of AM_NT doubles, the odd
// rows' halves swapped.  A 64-bit LDS read serves 32 lanes per cycle; the
// product's B reads put rows k and k + 1 (k even) in one such half-wave, which
// the swap sends to different banks; every other access walks along a row.
__device__ __forceinline__ int xat(int k, int c) { return k * AM_NT + (c ^ ((k & 1) << 4)); }

__global__ __launch_bounds__(AM_NEWTON_THREADS) void k_newton_fp64(NetDev net, ChainIO cio, NewtonIO io) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* X = reinterpret_cast<double*>(smem_raw);                             // [AM_CHAIN_MAXW][AM_NT]
  // [AM_NT][Ws]: an odd row stride, so lanes reading one word of different
  // columns' masks hit different banks
  uint32_t* smask = reinterpret_cast<uint32_t*>(X + (size_t)AM_CHAIN_MAXW * AM_NT);
  constexpr int NPART = AM_NEWTON_THREADS / AM_NT;   // row groups of the VALU layers
  __shared__ double s_v[AM_NT][3];
  __shared__ double s_z[AM_NT][2];
  __shared__ int s_ab[AM_NT][2];
  __shared__ int s_rk[AM_NT][2];
  __shared__ double s_part[NPART][AM_NT];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int base = blockIdx.x * AM_NT;
  const int total = io.count ? *io.count : 0;
  const int n = min(AM_NT, total - base);
  if (n <= 0) return;
  const int Wd = net.Wd, vm = net.vm, Ws = Wd | 1;
  if (tid < AM_NT) {
    int r = 0, k = 0, a = -1, b = -1;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    if (tid < n) {
      const int item = io.items[base + tid];
      r = io.item_region[item];
      k = item - (int)(io.roff[r] & 0xffffffffu);
      const double* v = io.verts + ((size_t)r * vm + k) * 3;
      v0 = v[0]; v1 = v[1]; v2 = v[2];
      a = io.vpair[((size_t)r * vm + k) * 2 + 0];
      b = io.vpair[((size_t)r * vm + k) * 2 + 1];
    }
    s_rk[tid][0] = r; s_rk[tid][1] = k;
    s_ab[tid][0] = a; s_ab[tid][1] = b;
    s_v[tid][0] = v0; s_v[tid][1] = v1; s_v[tid][2] = v2;
    s_z[tid][0] = 0.0; s_z[tid][1] = 0.0;
  }
  __syncthreads();
  for (int idx = tid; idx < AM_NT * Wd; idx += AM_NEWTON_THREADS) {
    const int c = idx / Wd, w = idx - c * Wd;
    smask[c * Ws + w] = c < n ? io.reg_mask[(size_t)s_rk[c][0] * Wd + w] : 0u;
  }
  __syncthreads();
  // hidden layer 1: z_j = W0[j] . v + b0[j]; thread: column c, rows part, part + NPART, ...
  {
    const int c = tid % AM_NT, part = tid / AM_NT;
    const uint32_t* mk = smask + (size_t)c * Ws;
    const int ca = s_ab[c][0], cb = s_ab[c][1];
    const int n1 = net.nodes[1];
    const double x0 = s_v[c][0], x1 = s_v[c][1], x2 = s_v[c][2];
    // four rows at a time, their weight loads in flight together
    constexpr int U = 4;
    for (int j0 = part; j0 < n1; j0 += U * NPART) {
      double w[U][4];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + u * NPART;
        const bool in = j < n1;
        w[u][0] = in ? net.W0[j * 3 + 0] : 0.0;
        w[u][1] = in ? net.W0[j * 3 + 1] : 0.0;
        w[u][2] = in ? net.W0[j * 3 + 2] : 0.0;
        w[u][3] = in ? net.b[0][j] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + u * NPART;
        if (j < n1) {
          const double z = __builtin_fma(w[u][2], x2, __builtin_fma(w[u][1], x1, __builtin_fma(w[u][0], x0, w[u][3])));
          if (j == ca) s_z[c][0] = z;
          if (j == cb) s_z[c][1] = z;
          X[xat(j, c)] = mbit(mk, j) ? z : 0.0;
        }
      }
    }
  }
  __syncthreads();
  // Hidden layers 2 .. D on the FP64 matrix cores: wave w computes rows row0 ..
  // row0 + 31 for the 32 vertices with v_mfma_f64_16x16x4f64 per 4 inputs (A:
  // lane l holds W[row + l % 16][k0 + l / 16]; B: X[k0 + l / 16][16 cb + l % 16];
  // result element v of lane l is row l / 16 + 4 v, column l % 16 of the block).
  const int l16 = lane & 15, l4 = lane >> 4;
  for (int li = 1; li < net.D; ++li) {
    const int nin = net.nodes[li], nout = net.nodes[li + 1], ld = cio.noutp[li];
    const int row0 = wave * 32;
    // row blocks with rows (W^T holds ld >= nout columns, a multiple of 16, zero past nout)
    const int nrb = row0 >= nout ? 0 : (row0 + 16 >= ld ? 1 : 2);
    am_f64x4 acc[2][2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = am_f64x4{0.0, 0.0, 0.0, 0.0};
    if (nrb > 0) {
      // W^T is zero past row nin (and up to AM_CHAIN_MAXW), so k may run to a multiple of 4
      const double* wa = cio.Wp[li] + (size_t)l4 * ld + row0 + l16;
      const int rb1 = nrb > 1 ? 16 : 0;   // the second block's offset (re-reads the first when absent)
      const int nk = (nin + 3) & ~3;
      constexpr int PF = 4;   // k steps of A fragments in flight
      double a0[PF], a1[PF];
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const bool in = 4 * u < nk;
        a0[u] = in ? wa[(size_t)(4 * u) * ld] : 0.0;
        a1[u] = in ? wa[(size_t)(4 * u) * ld + rb1] : 0.0;
      }
      for (int k0 = 0; k0 < nk; k0 += 4 * PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int kk = k0 + 4 * u;
          if (kk < nk) {
            const int kr = kk + l4;
            const double b0 = kr < nin ? X[xat(kr, l16)] : 0.0;
            const double b1 = kr < nin ? X[xat(kr, 16 + l16)] : 0.0;
            const double x0 = a0[u], x1 = a1[u];
            const int kn = kk + 4 * PF;
            a0[u] = kn < nk ? wa[(size_t)kn * ld] : 0.0;
            a1[u] = kn < nk ? wa[(size_t)kn * ld + rb1] : 0.0;
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, b1, acc[0][1], 0, 0, 0);
            if (nrb > 1) {
              acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, b0, acc[1][0], 0, 0, 0);
              acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, b1, acc[1][1], 0, 0, 0);
            }
          }
        }
      }
    }
    // the epilogue's biases, loaded while the other waves finish their products
    double bias[2][4];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = row0 + 16 * rb + l4 + 4 * v;
        bias[rb][v] = rb < nrb && j < nout ? cio.bp[li][j] : 0.0;
      }
    __syncthreads();   // every wave has read this layer's input
    const int out0 = net.off[li];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      if (rb >= nrb) continue;
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        const int c = 16 * cb + l16;
        const int ac = s_ab[c][0], bc = s_ab[c][1];
        const uint32_t* mc = smask + (size_t)c * Ws;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int j = row0 + 16 * rb + l4 + 4 * v;
          if (j < nout) {
            const double z = acc[rb][cb][v] + bias[rb][v];
            const int g = out0 + j;
            if (g == ac) s_z[c][0] = z;
            if (g == bc) s_z[c][1] = z;
            X[xat(j, c)] = mbit(mc, g) ? z : 0.0;
          }
        }
      }
    }
    __syncthreads();
  }
  // output layer (one row): f = w . X + b, the k range split over NPART row groups
  {
    const int li = net.D, nin = net.nodes[li], ld = cio.noutp[li];
    const int c = tid % AM_NT, part = tid / AM_NT;
    const double* w = cio.Wp[li];
    double acc = 0.0;
    for (int k = part; k < nin; k += NPART) acc = __builtin_fma(w[(size_t)k * ld], X[xat(k, c)], acc);
    s_part[part][c] = acc;
  }
  __syncthreads();
  if (tid >= n) return;
  double f = cio.bp[net.D][0];
  for (int q = 0; q < NPART; ++q) f += s_part[q][tid];
  // residuals of the signed planes of the vertex's system and the step
  const uint32_t* mk = smask + (size_t)tid * Ws;
  const int r = s_rk[tid][0], k = s_rk[tid][1];
  const double v[3] = {s_v[tid][0], s_v[tid][1], s_v[tid][2]};
  double res[3];
  for (int i = 0; i < 2; ++i) {
    const int g = s_ab[tid][i];
    if (g < net.L) {
      const double z = s_z[tid][i];
      res[i] = mbit(mk, g) ? -z : z;
    } else {
      const int e = g - net.L;
      res[i] = __builtin_fma(net.ew[e * 3 + 2], v[2], __builtin_fma(net.ew[e * 3 + 1], v[1],
                             __builtin_fma(net.ew[e * 3 + 0], v[0], net.eb[e])));
    }
  }
  res[2] = f - net.iso;
  const uint16_t* jv = io.jinv + ((size_t)r * vm + k) * 9;
  double* out = io.verts + ((size_t)r * vm + k) * 3;
  for (int x = 0; x < 3; ++x) {
    const double d = (double)bf16_value(jv[x]) * res[0] + (double)bf16_value(jv[3 + x]) * res[1] +
                     (double)bf16_value(jv[6 + x]) * res[2];
    out[x] = v[x] - d;
  }
}

// Every other use of a refined vertex (a vertex use of a record with the
// Newton data whose key's first use is elsewhere) takes the first use's
// coordinates, so all uses of a vertex agree.
__global__ void k_newton_scatter(NewtonIO io, const uint32_t* table, const int* slot, const unsigned* vflag,
                                 const int* info, int N, int vm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || vflag[i]) return;
  const int r = io.item_region[i];
  if (info[(size_t)r * AM_INFO_STRIDE + 7] != 1) return;
  const int owner = (int)table[slot[i]];
  const int ro = io.item_region[owner];
  const int k = i - (int)(io.roff[r] & 0xffffffffu), ko = owner - (int)(io.roff[ro] & 0xffffffffu);
  const double* src = io.verts + ((size_t)ro * vm + ko) * 3;
  double* dst = io.verts + ((size_t)r * vm + k) * 3;
  dst[0] = src[0];
  dst[1] = src[1];
  dst[2] = src[2];
}

}  // namespace am
