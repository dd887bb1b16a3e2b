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
// One workgroup of 16 waves takes 64 vertices (one column each), wave w owns
// output rows 16 w .. 16 w + 15 of every layer.  A layer's input (masked) lives
// in LDS; the products run on the FP64 matrix cores, the A fragments (weights)
// streamed from memory a few steps ahead.
#include "am_device.h"
#include "am_kernels.h"
#include "am_newton.h"

namespace am {

typedef double am_f64x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(AM_NEWTON_THREADS) void k_newton_fp64(NetDev net, ChainIO cio, NewtonIO io) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* X = reinterpret_cast<double*>(smem_raw);                             // [AM_CHAIN_MAXW][AM_NT]
  uint32_t* smask = reinterpret_cast<uint32_t*>(X + (size_t)AM_CHAIN_MAXW * AM_NT);   // [AM_NT][Wd]
  __shared__ double s_v[AM_NT][3];
  __shared__ double s_z[AM_NT][2];
  __shared__ int s_ab[AM_NT][2];
  __shared__ int s_rk[AM_NT][2];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int base = blockIdx.x * AM_NT;
  const int total = io.count ? *io.count : 0;
  const int n = min(AM_NT, total - base);
  if (n <= 0) return;
  const int Wd = net.Wd, vm = net.vm;
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
    smask[idx] = c < n ? io.reg_mask[(size_t)s_rk[c][0] * Wd + w] : 0u;
  }
  __syncthreads();
  const uint32_t* mk = smask + (size_t)lane * Wd;
  const int ca = s_ab[lane][0], cb = s_ab[lane][1];
  // hidden layer 1: z_j = W0[j] . v + b0[j]
  {
    const int n1 = net.nodes[1];
    const double x0 = s_v[lane][0], x1 = s_v[lane][1], x2 = s_v[lane][2];
    for (int jj = 0; jj < AM_NROW; ++jj) {
      const int j = wave * AM_NROW + jj;
      if (j >= n1) break;
      const double z = __builtin_fma(net.W0[j * 3 + 2], x2, __builtin_fma(net.W0[j * 3 + 1], x1,
                                     __builtin_fma(net.W0[j * 3 + 0], x0, net.b[0][j])));
      if (j == ca) s_z[lane][0] = z;
      if (j == cb) s_z[lane][1] = z;
      X[(size_t)j * AM_NT + lane] = mbit(mk, j) ? z : 0.0;
    }
  }
  __syncthreads();
  // Layers 2 .. D + 1 on the FP64 matrix cores: wave w computes rows
  // row0 .. row0 + 15 for all 64 vertices as four 16 x 16 blocks of
  // v_mfma_f64_16x16x4f64 per 4 inputs (A: lane l holds W[row0 + l % 16][k0 + l / 16];
  // B: X[k0 + l / 16][16 cb + l % 16]; result element v of lane l is row l / 16 + 4 v,
  // column l % 16 of the block).  The A fragments of the next steps are in flight
  // while a step's products run.
  __shared__ double s_f[AM_NT];
  const int l16 = lane & 15, l4 = lane >> 4;
  for (int li = 1; li <= net.D; ++li) {
    const int nin = net.nodes[li], nout = net.nodes[li + 1], ld = cio.noutp[li];
    const int row0 = wave * AM_NROW;
    const bool out_layer = li == net.D;
    am_f64x4 acc[AM_NT / 16];
#pragma unroll
    for (int cb = 0; cb < AM_NT / 16; ++cb) acc[cb] = am_f64x4{0.0, 0.0, 0.0, 0.0};
    if (row0 < nout) {
      // W^T is zero past row nin (and up to AM_CHAIN_MAXW), so k may run to a multiple of 4
      const double* wa = cio.Wp[li] + (size_t)l4 * ld + row0 + l16;
      const int nk = (nin + 3) & ~3;
      constexpr int PF = 4;   // A fragments in flight
      double a[PF];
#pragma unroll
      for (int u = 0; u < PF; ++u) a[u] = 4 * u < nk ? wa[(size_t)(4 * u) * ld] : 0.0;
      for (int k0 = 0; k0 < nk; k0 += 4 * PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int kk = k0 + 4 * u;
          if (kk < nk) {
            const int kr = kk + l4;
            double b[AM_NT / 16];
#pragma unroll
            for (int cb = 0; cb < AM_NT / 16; ++cb) b[cb] = kr < nin ? X[(size_t)kr * AM_NT + 16 * cb + l16] : 0.0;
            const double av = a[u];
            const int kn = kk + 4 * PF;
            a[u] = kn < nk ? wa[(size_t)kn * ld] : 0.0;
#pragma unroll
            for (int cb = 0; cb < AM_NT / 16; ++cb)
              acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b[cb], acc[cb], 0, 0, 0);
          }
        }
      }
    }
    __syncthreads();   // every wave has read this layer's input
    if (row0 < nout) {
      const int out0 = out_layer ? 0 : net.off[li];
#pragma unroll
      for (int cb = 0; cb < AM_NT / 16; ++cb) {
        const int c = 16 * cb + l16;
        const int ac = s_ab[c][0], bc = s_ab[c][1];
        const uint32_t* mc = smask + (size_t)c * Wd;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int j = row0 + l4 + 4 * v;
          if (j < nout) {
            const double z = acc[cb][v] + cio.bp[li][j];
            if (out_layer) {
              s_f[c] = z;   // nout == 1: row 0
            } else {
              const int g = out0 + j;
              if (g == ac) s_z[c][0] = z;
              if (g == bc) s_z[c][1] = z;
              X[(size_t)j * AM_NT + c] = mbit(mc, g) ? z : 0.0;
            }
          }
        }
      }
    }
    __syncthreads();
  }
  const double f = s_f[lane];
  if (wave != 0 || lane >= n) return;
  // residuals of the signed planes of the vertex's system and the step
  const int r = s_rk[lane][0], k = s_rk[lane][1];
  const double v[3] = {s_v[lane][0], s_v[lane][1], s_v[lane][2]};
  double res[3];
  for (int i = 0; i < 2; ++i) {
    const int g = s_ab[lane][i];
    if (g < net.L) {
      const double z = s_z[lane][i];
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
