// Reduced-precision plane path.
//
// Face kernel: for a tile of AM_GT regions the masked affine chain
//   P_{l+1} = W_l diag(m_l) P_l   (columns: x, y, z and the bias column c)
// is one GEMM per layer, W_l[n_out x n_in] x X[n_in x 4*AM_GT], on MFMA with
// X staged in LDS (backend/inc/process.h:18-193 does the same chain as
// per-region cuBLAS calls).  Two arithmetic forms:
//   * FP32: v_mfma_f32_16x16x4f32 over the inputs active in some region of the
//     tile only (the others are +0 in every column); A rows are gathered from
//     a k-major copy of W^T, X is staged already compacted;
//   * split-BF16: w = wh + wl, x = xh + xl (bf16 parts), y = wh.xh + wh.xl + wl.xh
//     with v_mfma_f32_16x16x32_bf16 and FP32 accumulation.
// Beside each plane the FP32 form writes an estimate sigma of its error: the
// variance of the accumulated rounding is propagated through the chain,
// V' = W^2 m (V + ck |P|^2), as one more small product inside the k loop (bf16
// MFMA; its A operand is the square of the weight fragments the loop already
// holds; the estimate needs a few bits only).
// The checked walk bounds each plane by lam sigma plus a few ulps instead of
// one relative bound for every plane of the net.
// The signed planes go to HBM, one SoA record per region, and the checked FP32
// walk reads them back into LDS.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "am_fast.h"
#include "am_scan.h"

namespace am {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));

#define AM_XS (AM_GN + 4)   // fp32 LDS row stride (floats)
#define AM_GRBW 4           // 16-row output blocks per wave (layers up to 256 outputs)

struct GemmIO {
  const uint32_t* reg_mask;
  int first;        // region id of local index 0
  int count;        // regions in this chunk
  float* planes;    // [count][4][Cs]
  float* eq;        // [count][4]
};

__device__ __forceinline__ uint16_t bf16_rn(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
// The same rounding (to nearest even, finite x) in one instruction (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t bf16_hw(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }

// Input column of the error-estimate product for an input plane with
// variance V and squared norm n2 (active in its region): V + ck n2, capped to
// stay finite in bf16.
__device__ __forceinline__ uint16_t sig_input(float V, float n2, float ck) {
  return bf16_hw(fminf(V + ck * n2, 1e30f));
}

// Where the error-estimate input of list position p and region t sits in the
// row paddings (bf16 units from the first padding): the B operand layout of
// v_mfma_f32_16x16x16_bf16 for the k loop's step kb = p >> 2 group.  Lane
// (q, t) of the step group starting at list position 16 g reads positions
// 16 g + 4 i + q, i = 0..3, as one 8-byte word: row 16 g + 4 q + t / 2, half t % 2.
// A group's slots lie in its own 16 rows.
__device__ __forceinline__ int sig_slot(int p, int t) {
  return ((p & ~15) + 4 * (p & 3) + (t >> 1)) * 2 * AM_XS + (t & 1) * 4 + ((p & 15) >> 2);
}

// Stage layer-1 planes: X = masked rows of W0 (and b0); signed planes to HBM.
// FP32: X holds only the tile's active inputs (kl1, nkp1 entries).
template <int PREC>
__device__ void stage_layer1(const FastNetDev& f, const GemmIO& io, const uint32_t* mk, int r0, int nreg,
                             float* Xs, uint16_t* Xh, uint16_t* Xl, int KS, const uint16_t* kl1, int nkp1,
                             uint16_t* Ypad) {
  const int tid = threadIdx.x;
  const int n1 = f.nodes[1];
  if (PREC == 0) {
    // X rows = the tile's active inputs of layer 1, in list order (+0 past the
    // list): a thread takes list position i, reads its row of (W0 | b0) once
    // and writes the row's 8 regions as float4s; with Ypad, also the first
    // error-estimate product's input (hidden layer 1's planes are exact: V = 0)
    for (int i = tid; i < nkp1; i += 256) {
      const int k = kl1[i];
      float4 w = {0.f, 0.f, 0.f, 0.f};
      if (k < n1) w = float4{f.W0[k * 3 + 0], f.W0[k * 3 + 1], f.W0[k * 3 + 2], f.bias[0][k]};
      const uint16_t xs = Ypad ? sig_input(0.f, w.x * w.x + w.y * w.y + w.z * w.z + w.w * w.w, f.ck[1]) : 0;
#pragma unroll
      for (int t = 0; t < AM_GT; ++t) {
        const bool on = k < n1 && mbit(mk + t * f.Wd, k);
        *reinterpret_cast<float4*>(Xs + i * AM_XS + 4 * t) = on ? w : float4{0.f, 0.f, 0.f, 0.f};
        if (Ypad) Ypad[sig_slot(i, t)] = on ? xs : (uint16_t)0;
      }
    }
  }
  for (int idx = tid; PREC != 0 && idx < n1 * AM_GN; idx += 256) {
    const int k = idx / AM_GN, col = idx - k * AM_GN, t = col >> 2, q = col & 3;
    const float val = q < 3 ? f.W0[k * 3 + q] : f.bias[0][k];
    const float x = mbit(mk + t * f.Wd, k) ? val : 0.f;
    {
      const uint16_t h = bf16_rn(x);
      Xh[col * KS + k] = h;
      Xl[col * KS + k] = bf16_rn(x - bf16_to_f(h));
    }
  }
  // signed planes to HBM: thread = constraint (consecutive threads, consecutive addresses)
  // (hidden layer 1's planes are rows of W0 and b0, exact: sigma 0)
  for (int k = tid; k < n1; k += 256) {
    const float v4[4] = {f.W0[k * 3 + 0], f.W0[k * 3 + 1], f.W0[k * 3 + 2], f.bias[0][k]};
    for (int t = 0; t < nreg; ++t) {
      const bool on = mbit(mk + t * f.Wd, k);
      float* dst = io.planes + (size_t)(r0 + t) * AM_PCOMP * f.Cs + k;
#pragma unroll
      for (int q = 0; q < 4; ++q) dst[(size_t)q * f.Cs] = on ? -v4[q] : v4[q];
      if (f.sig) dst[(size_t)4 * f.Cs] = 0.f;
    }
  }
  for (int idx = tid; idx < AM_GT * AM_PCOMP * f.E; idx += 256) {
    const int col = idx / f.E, e = idx - col * f.E, t = col / AM_PCOMP, q = col - t * AM_PCOMP;
    if (t < nreg && (q < 4 || f.sig))
      io.planes[((size_t)(r0 + t) * AM_PCOMP + q) * f.Cs + f.L + e] = q < 3 ? f.ew[e * 3 + q] : (q == 3 ? f.eb[e] : 0.f);
  }
}


template <int PREC>
__global__ __launch_bounds__(256, PREC == 0 ? 4 : 1) void k_gemm(FastNetDev f, GemmIO io) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int KS = f.Kpad + 8;   // bf16 column stride (halves)
  // One buffer holds a layer's input during its k loop and, after a barrier,
  // its outputs (Ys, [Kpad][AM_XS] floats) that become the next input:
  // fp32 X at list positions ([Kpad][AM_XS]); bf16 Xh/Xl ([AM_GN][KS] halves each).
  float* Ys = reinterpret_cast<float*>(smem_raw);
  float* Xs = Ys;
  uint16_t* Xh = reinterpret_cast<uint16_t*>(smem_raw);
  uint16_t* Xl = Xh + (size_t)AM_GN * KS;
  const size_t xbytes = PREC == 0 ? (size_t)f.Kpad * AM_XS * 4
                                  : std::max((size_t)f.Kpad * AM_XS * 4, (size_t)2 * AM_GN * KS * 2);
  uint32_t* mk = reinterpret_cast<uint32_t*>(smem_raw + xbytes);
  const int KLF = f.Kpad;                                            // entries of one layer's input list
  uint16_t* klist = reinterpret_cast<uint16_t*>(mk + AM_GT * f.Wd);  // fp32: [D][KLF]
  // Error estimates (FP32 form only): the product's input, element (k, t)
  // for region t < AM_GT, bf16, in the 16-byte padding of row k of the X / Y
  // buffer (columns AM_GN .., unused by the planes); after the product has
  // read it, the same slot holds the output V (k, t), bf16.
  constexpr bool SIG = PREC == 0;
  uint16_t* Ypad = reinterpret_cast<uint16_t*>(Ys) + 2 * AM_GN;
  auto xq_at = [&](int k, int t) -> uint16_t& { return Ypad[k * 2 * AM_XS + t]; };
  __shared__ int s_nk[AM_MAX_LINEAR];
  const int r0 = blockIdx.x * AM_GT;
  const int nreg = min(AM_GT, io.count - r0);

  for (int idx = tid; idx < AM_GT * f.Wd; idx += 256) {
    const int t = idx / f.Wd, w = idx - t * f.Wd;
    mk[idx] = t < nreg ? io.reg_mask[(size_t)(io.first + r0 + t) * f.Wd + w] : 0u;
  }
  // (FP32: every LDS element a step reads is written first -- list positions
  // by the staging and the compaction, parked rows below n_out -- so no clear)
  if (PREC != 0)
    for (int i = tid; i < 2 * AM_GN * KS; i += 256) Xh[i] = 0;
  __syncthreads();
  if (PREC == 0) {
    // Inputs of linear layer li that are active in some region of the tile,
    // ascending, padded to a multiple of 16 with n_li (the +0 row of W^T and,
    // through the staging below, of X; at most Kpad entries).  Inactive inputs
    // are +0 in every column, so the GEMM over the list equals the full one.
    for (int li = wave + 1; li <= f.D; li += 4) {
      const int nin = f.nodes[li], in0 = li == 1 ? 0 : f.off[li - 1];
      uint16_t* kl = klist + (size_t)(li - 1) * KLF;
      int base = 0;
      for (int kb = 0; kb < nin; kb += 64) {
        const int k = kb + lane;
        bool act = false;
        if (k < nin)
          for (int t = 0; t < AM_GT; ++t) act = act || mbit(mk + t * f.Wd, in0 + k);
        const unsigned long long bal = __ballot(act);
        if (act) kl[base + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u))] = (uint16_t)k;
        base += __popcll(bal);
      }
      const int nkp = (base + 15) & ~15;
      for (int i = base + lane; i < nkp; i += 64) kl[i] = (uint16_t)nin;
      if (lane == 0) s_nk[li] = nkp;
    }
    __syncthreads();
  }
  stage_layer1<PREC>(f, io, mk, r0, nreg, Xs, Xh, Xl, KS, klist, PREC == 0 ? s_nk[1] : 0, SIG ? Ypad : nullptr);
  __syncthreads();

  for (int li = 1; li <= f.D; ++li) {
    const int nout = f.nodes[li + 1];
    const int nrb = f.nrb[li], nkb = f.nkb[li];
    const float* bias = f.bias[li];
    // Row blocks rb = wave + 4 u of this wave (n_out <= 64 * AM_GRBW).  The k
    // loop is outermost, so each B fragment read from LDS feeds every row block.
    f32x4 accs[AM_GRBW][2];
    f32x4 accq[AM_GRBW];   // FP32: the error-estimate product V = W^2 . (input) of the row blocks
#pragma unroll
    for (int u = 0; u < AM_GRBW; ++u) {
      accs[u][0] = f32x4{0.f, 0.f, 0.f, 0.f};
      accs[u][1] = f32x4{0.f, 0.f, 0.f, 0.f};
      accq[u] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    const int nu = (nrb - wave + 3) / 4;   // row blocks of this wave (uniform; <= 0: none)
    if (nu <= 0) {
    } else if (PREC == 0) {
      // A[row][kb*4 + (lane >> 4)] = W^T[list entry][row] (k-major, rows padded to
      // 16), through a buffer resource (32-bit byte offsets); row block u's
      // column is 64 u floats past row block 0's
      const int nkl = s_nk[li] >> 2;   // a multiple of 4
      const uint16_t* kl = klist + (size_t)(li - 1) * KLF + (lane >> 4);
      const int ld4 = nrb * 16 * 4;
      const __amdgpu_buffer_rsrc_t wres = __builtin_amdgcn_make_buffer_rsrc(
          (void*)f.Wp[li], (short)0, (f.nodes[li] + 1) * ld4, 0x00020000);
      const int cw = (wave * 16 + (lane & 15)) * 4;
      const float* xb = Xs + (lane >> 4) * AM_XS + (lane & 15);
      // error-estimate operand of lane (q, t): 4 positions of its group as one
      // 8-byte word (sig_slot; columns 8..15 read region t - 8's, their
      // results are not used)
      const uint2* sb = reinterpret_cast<const uint2*>(Ypad + sig_slot(lane >> 4, lane & 7));
      for (int kb = 0; kb < nkl; kb += 4) {
        int ko[4];
        float bx[4][2];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          ko[e] = (int)kl[(kb + e) * 4] * ld4 + cw;
          bx[e][0] = xb[(kb + e) * 4 * AM_XS];
          bx[e][1] = xb[(kb + e) * 4 * AM_XS + 16];
        }
        uint2 bq = {0u, 0u};
        if (SIG) bq = sb[2 * kb * AM_XS];   // the group's 16 rows start 4 kb rows (AM_XS / 2 words each) in
#pragma unroll
        for (int u = 0; u < AM_GRBW; ++u) {
          if (u < nu) {
            float av[4];
#pragma unroll
            for (int e = 0; e < 4; ++e)
              av[e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(wres, ko[e] + 256 * u, 0, 0));
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              accs[u][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[e], bx[e][0], accs[u][0], 0, 0, 0);
              accs[u][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[e], bx[e][1], accs[u][1], 0, 0, 0);
            }
            if (SIG) {
              // squares rounded to bf16 (v_cvt_pk_bf16_f32, to nearest even)
              const bf16x4 a = {(__bf16)(av[0] * av[0]), (__bf16)(av[1] * av[1]), (__bf16)(av[2] * av[2]),
                                (__bf16)(av[3] * av[3])};
              accq[u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s16x4, a),
                                                                  __builtin_bit_cast(s16x4, bq), accq[u], 0, 0, 0);
            }
          }
        }
      }
    } else {
      const int kq = 8 * (lane >> 4);
      const uint16_t* xh0 = Xh + (lane & 15) * KS + kq;
      const uint16_t* xl0 = Xl + (lane & 15) * KS + kq;
      const uint16_t* xh1 = xh0 + 16 * KS;
      const uint16_t* xl1 = xl0 + 16 * KS;
      const uint4* Ah = reinterpret_cast<const uint4*>(f.Wh[li]) + lane;
      const uint4* Al = reinterpret_cast<const uint4*>(f.Wl[li]) + lane;
      for (int kb = 0; kb < nkb; ++kb) {
        const bf16x8 vbh0 = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(xh0 + kb * 32));
        const bf16x8 vbl0 = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(xl0 + kb * 32));
        const bf16x8 vbh1 = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(xh1 + kb * 32));
        const bf16x8 vbl1 = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(xl1 + kb * 32));
#pragma unroll
        for (int u = 0; u < AM_GRBW; ++u) {
          if (u < nu) {
            const size_t at = ((size_t)(wave + 4 * u) * nkb + kb) * 64;
            const bf16x8 vah = __builtin_bit_cast(bf16x8, Ah[at]), val_ = __builtin_bit_cast(bf16x8, Al[at]);
            f32x4 acc0 = accs[u][0], acc1 = accs[u][1];
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vah, vbh0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vah, vbh1, acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vah, vbl0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vah, vbl1, acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(val_, vbh0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(val_, vbh1, acc1, 0, 0, 0);
            accs[u][0] = acc0;
            accs[u][1] = acc1;
          }
        }
      }
    }
    // the park's biases, loaded while the other waves finish their products
    float bv[AM_GRBW][4];
#pragma unroll
    for (int u = 0; u < AM_GRBW; ++u)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int rb = wave + 4 * u, row = rb * 16 + 4 * (lane >> 4) + i;
        bv[u][i] = rb < nrb && row < nout ? bias[row] : 0.f;
      }
    // the outputs go over X, so every wave must be done reading it
    __syncthreads();
#pragma unroll
    for (int u = 0; u < AM_GRBW; ++u) {
      const int rb = wave + 4 * u;
      if (rb >= nrb) continue;
      const f32x4 acc0 = accs[u][0], acc1 = accs[u][1];
      const int c0 = lane & 15;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = rb * 16 + 4 * (lane >> 4) + i;
        if (row < nout) {
          const float b = bv[u][i];
          Ys[row * AM_XS + c0] = acc0[i] + ((c0 & 3) == 3 ? b : 0.f);
          Ys[row * AM_XS + 16 + c0] = acc1[i] + (((16 + c0) & 3) == 3 ? b : 0.f);
          // V of plane row, region c0, in row row's padding
          if (SIG && c0 < AM_GT) xq_at(row, c0) = bf16_hw(fminf(fmaxf(accq[u][i], 0.f) + f.ck[li] * b * b, 1e30f));
        }
      }
    }
    __syncthreads();
    if (li < f.D) {
      const int out0 = f.off[li];
      // signed planes to HBM: thread = constraint (consecutive threads, consecutive addresses)
      for (int j = tid; j < nout; j += 256) {
        for (int t = 0; t < nreg; ++t) {
          const bool on = mbit(mk + t * f.Wd, out0 + j);
          float* dst = io.planes + (size_t)(r0 + t) * AM_PCOMP * f.Cs + out0 + j;
          const float4 y = *reinterpret_cast<const float4*>(Ys + j * AM_XS + 4 * t);
          dst[0] = on ? -y.x : y.x;
          dst[f.Cs] = on ? -y.y : y.y;
          dst[2 * (size_t)f.Cs] = on ? -y.z : y.z;
          dst[3 * (size_t)f.Cs] = on ? -y.w : y.w;
          if (SIG && f.sig) dst[4 * (size_t)f.Cs] = 1.02f * __builtin_sqrtf(bf16_to_f(xq_at(j, t)));
        }
      }
      if (PREC == 0) {
        // next layer's X, in place: position i holds output kl2[i] masked (+0
        // for the padding entries).  The list ascends, so position i reads row
        // kl2[i] >= i: in chunks of 128 positions, each chunk reads (rows at or
        // past its first position) before a barrier and writes after it, and
        // no chunk overwrites a row a later chunk reads.
        const uint16_t* kl2 = klist + (size_t)li * KLF;
        const int nkp2 = s_nk[li + 1];
        // thread (region t, position p0): region t's four columns of positions
        // p0, p0 + 32, p0 + 64, p0 + 96 of each chunk, as one float4 each (one
        // list, mask and row read per four values); the next product's input
        // rides along, m (V + ck |P|^2) from the same float4 (sig_slot places
        // it in the rows of the position's own 16-row group, so the chunk rule
        // holds for it too)
        constexpr int PP = 256 / AM_GT;   // positions per pass of the block
        constexpr int CH = 128 / PP;      // passes per chunk
        int tq = tid;
        asm volatile("" : "+v"(tq));       // keeps these offsets from being hoisted out of the layer loop
        const int t = tq & (AM_GT - 1), p0 = tq / AM_GT;
        const uint32_t* mt = mk + t * f.Wd;
        const float ckn = f.ck[li + 1];
        for (int c0 = 0; c0 < nkp2; c0 += 128) {
          float4 xv[CH];
          uint16_t qv[CH];
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            const int i = c0 + c * PP + p0;
            xv[c] = float4{0.f, 0.f, 0.f, 0.f};
            qv[c] = 0;
            if (i < nkp2) {
              const int k = kl2[i];
              if (k < nout && mbit(mt, out0 + k)) {
                const float4 y = *reinterpret_cast<const float4*>(Ys + k * AM_XS + 4 * t);
                xv[c] = y;
                if (SIG) qv[c] = sig_input(bf16_to_f(xq_at(k, t)), y.x * y.x + y.y * y.y + y.z * y.z + y.w * y.w, ckn);
              }
            }
          }
          __syncthreads();
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            const int i = c0 + c * PP + p0;
            if (i < nkp2) {
              *reinterpret_cast<float4*>(Xs + i * AM_XS + 4 * t) = xv[c];
              if (SIG) Ypad[sig_slot(i, t)] = qv[c];
            }
          }
        }
      }
      if (PREC != 0) {
        // next layer's Xh/Xl (masked outputs split into bf16 parts, rows up to
        // a multiple of 32), over the outputs: every thread reads its values,
        // then a barrier, then the writes
        const int rows = (nout + 31) & ~31;
        constexpr int XPT = 256 * AM_GN / 256;   // values per thread for rows <= 256
        int tq = tid;
        asm volatile("" : "+v"(tq));             // keeps these offsets from being hoisted out of the layer loop
        const int col = tq & (AM_GN - 1), t = col >> 2, j0 = tq / AM_GN;
        float xv[XPT];
#pragma unroll
        for (int c = 0; c < XPT; ++c) {
          const int j = j0 + c * (256 / AM_GN);
          xv[c] = (j < nout && mbit(mk + t * f.Wd, out0 + j)) ? Ys[j * AM_XS + col] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < XPT; ++c) {
          const int j = j0 + c * (256 / AM_GN);
          if (j < rows) {
            const uint16_t h = bf16_rn(xv[c]);
            Xh[col * KS + j] = h;
            Xl[col * KS + j] = bf16_rn(xv[c] - bf16_to_f(h));
          }
        }
      }
    } else {
      if (tid < AM_GN) {
        const int t = tid >> 2, q = tid & 3;
        if (t < nreg) {
          float v = Ys[tid];
          if (q == 3) v = v - f.iso;
          io.eq[(size_t)(r0 + t) * AM_EQS + q] = v;
        }
      }
      // the face plane's estimate
      if (SIG && f.sig && tid < nreg)
        io.eq[(size_t)(r0 + tid) * AM_EQS + 4] = 1.02f * __builtin_sqrtf(bf16_to_f(xq_at(0, tid)));
    }
    __syncthreads();
  }
}

// Checked FP32 walk of one region whose planes the face kernel wrote.
__global__ __launch_bounds__(64) void k_walk_f32(NetDev net, RegionIO io, WalkParams wp, const float* planes,
                                                 const float* eqbuf, int Cs_hbm, int chunk_first) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = net.C;
  const int Cs = io.Cs, N2 = io.Csort;
  float* Px = reinterpret_cast<float*>(smem_raw);
  float* Py = Px + Cs;
  float* Pz = Py + Cs;
  float* Pc = Pz + Cs;
  float* Pe = Pc + Cs;
  float* keys = Pe + Cs;
  int* order = reinterpret_cast<int*>(keys + N2);
  uint32_t* used = reinterpret_cast<uint32_t*>(order + N2);
  uint32_t* smask = used + ((C + 31) >> 5);
  const int lane = threadIdx.x;
  const int b = blockIdx.x;          // index within the chunk
  if (b >= io.count) return;
  const int rid = chunk_first + b;
  const float* src = planes + (size_t)b * AM_PCOMP * Cs_hbm;
  const bool with_sigma = wp.lam > 0.f;
  for (int g = lane; g < C; g += 64) {
    Px[g] = src[g];
    Py[g] = src[Cs_hbm + g];
    Pz[g] = src[2 * Cs_hbm + g];
    Pc[g] = src[3 * Cs_hbm + g];
    if (with_sigma) Pe[g] = src[4 * Cs_hbm + g];
  }
  for (int w = lane; w < net.Wd; w += 64) smask[w] = io.reg_mask[(size_t)rid * net.Wd + w];
  for (int w = lane; w < ((C + 31) >> 5); w += 64) used[w] = 0u;
  const float* eb = eqbuf + (size_t)b * AM_EQS;
  const float eq[4] = {eb[0], eb[1], eb[2], eb[3]};
  const float p[3] = {(float)io.reg_point[(size_t)rid * 3 + 0], (float)io.reg_point[(size_t)rid * 3 + 1],
                      (float)io.reg_point[(size_t)rid * 3 + 2]};
  __syncthreads();
  walk_region<float, true>(net, io, wp, rid, rid - io.first, Px, Py, Pz, Pc, keys, order, used, smask, eq, p,
                           with_sigma ? Pe : nullptr, with_sigma ? eb[4] : 0.f);
}

// Ray-shooting FP32 walk: LDS holds only the planes and the mask; NW waves
// per region.
template <int NW>
__global__ __launch_bounds__(64 * NW, 24 / NW) void k_walk_ray_f32(NetDev net, RegionIO io, WalkParams wp, const float* planes,
                                                          const float* eqbuf, int Cs_hbm, int chunk_first) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int C = net.C;
  const int Cs = io.Cs;
  float4* Q = reinterpret_cast<float4*>(smem_raw);   // (x, y, z, c) per constraint
  float* Pe = reinterpret_cast<float*>(Q + Cs);
  uint32_t* smask = reinterpret_cast<uint32_t*>(Pe + Cs);
  const WalkRecs rec = carve_walk_recs(
      reinterpret_cast<unsigned char*>(Q) + (((size_t)Cs * 20 + (size_t)net.Wd * 4 + 15) & ~(size_t)15), net.vm, true);
  const int lane = threadIdx.x;
  const int b = blockIdx.x;
  if (b >= io.count) return;
  const int rid = chunk_first + b;
  const int start_in = io.reg_start[rid];
  const float* src = planes + (size_t)b * AM_PCOMP * Cs_hbm;
  const bool with_sigma = wp.lam > 0.f;
  for (int g = lane; g < C; g += 64 * NW) {
    Q[g] = float4{src[g], src[Cs_hbm + g], src[2 * Cs_hbm + g], src[3 * Cs_hbm + g]};
    if (with_sigma) Pe[g] = src[4 * Cs_hbm + g];
  }
  for (int w = lane; w < net.Wd; w += 64 * NW) smask[w] = io.reg_mask[(size_t)rid * net.Wd + w];
  const float* eb = eqbuf + (size_t)b * AM_EQS;
  const float eq[4] = {eb[0], eb[1], eb[2], eb[3]};
  const float p[3] = {(float)io.reg_point[(size_t)rid * 3 + 0], (float)io.reg_point[(size_t)rid * 3 + 1],
                      (float)io.reg_point[(size_t)rid * 3 + 2]};
  __syncthreads();
  const PlaneView<float> P{reinterpret_cast<const float*>(Q), Cs, 0, nullptr, nullptr, smask,
                           with_sigma ? Pe : nullptr};
  walk_ray<float, NW>(net, io, wp, rid, rid - io.first, P, smask, eq, p, rec, start_in, with_sigma ? eb[4] : 0.f);
}

// The level's regions the walk left to FP64 in one launch: their ids (first +
// level index, in region order, through a one-pass scan of 2048-region
// tiles) and count, and the counters: undecided regions added to acc[0],
// their reasons to acc[1 + w], the decided ones with a vertex left to FP64 to
// acc[1 + AM_WHY_COUNT] (one atomic per wave and counter).
constexpr int kRedoThreads = 256, kRedoItems = 8, kRedoTile = kRedoThreads * kRedoItems;
__global__ __launch_bounds__(kRedoThreads) void k_redo_list(const unsigned* flags, int n, int first, int* acc,
                                                           RedoList out, int tiles) {
  __shared__ int s_tile, s_excl;
  __shared__ int s_wave[kRedoThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63;
  if (tid == 0) s_tile = (int)atomicAdd(&out.state[tiles], 1ull);
  __syncthreads();
  const int t = s_tile;
  const int base = t * kRedoTile + tid * kRedoItems;
  unsigned f[kRedoItems];
  if (base + kRedoItems <= n) {
    const uint4* q = reinterpret_cast<const uint4*>(flags + base);
    const uint4 a = q[0], b = q[1];
    f[0] = a.x; f[1] = a.y; f[2] = a.z; f[3] = a.w;
    f[4] = b.x; f[5] = b.y; f[6] = b.z; f[7] = b.w;
  } else {
#pragma unroll
    for (int it = 0; it < kRedoItems; ++it) f[it] = base + it < n ? flags[base + it] : 0u;
  }
  int sum = 0, und = 0, v64 = 0;
  int why[AM_WHY_COUNT];
#pragma unroll
  for (int w = 0; w < AM_WHY_COUNT; ++w) why[w] = 0;
#pragma unroll
  for (int it = 0; it < kRedoItems; ++it) {
    const unsigned x = f[it];
    sum += (x & (AM_UNDECIDED | AM_VERT64)) ? 1 : 0;
    if (x & AM_UNDECIDED) {
      und += 1;
#pragma unroll
      for (int w = 0; w < AM_WHY_COUNT; ++w) why[w] += (x & AM_WHY_BIT(w)) ? 1 : 0;
    } else if (x & AM_VERT64) {
      v64 += 1;
    }
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    und += __shfl_xor(und, d, 64);
    v64 += __shfl_xor(v64, d, 64);
  }
  if (lane == 0 && und) atomicAdd(acc, und);
  if (lane == 0 && v64) atomicAdd(acc + 1 + AM_WHY_COUNT, v64);
  if (__ballot(und != 0)) {
#pragma unroll
    for (int w = 0; w < AM_WHY_COUNT; ++w) {
      int k = why[w];
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) k += __shfl_xor(k, d, 64);
      if (lane == 0 && k) atomicAdd(acc + 1 + w, k);
    }
  }
  int tot;
  const int off = block_excl_scan<int, kRedoThreads>(sum, s_wave, &tot);
  if (tid < 64) {
    const int excl = lookback_1pass(out.state, tiles, t, tot, out.count);
    if (lane == 0) s_excl = excl;
  }
  __syncthreads();
  int run = s_excl + off;
#pragma unroll
  for (int it = 0; it < kRedoItems; ++it)
    if (f[it] & (AM_UNDECIDED | AM_VERT64)) out.ids[run++] = first + base + it;
}

size_t redo_state_words(int count) { return (size_t)std::max(1, (count + kRedoTile - 1) / kRedoTile) + 2; }

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

#define FCHK(x)                                                                   \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") +   \
                                                   hipGetErrorString(e_) + " in " #x); \
  } while (0)

enum Slot {
  S_W0 = 0, S_EW, S_EB, S_PLANES, S_EQ,
  S_BIAS0 = 8,      // 8..23: biases per linear layer
  S_WP0 = 24,       // 24..39: operand buffers per linear layer (fp32, or hi+lo packed)
};

// Buffer s with at least `bytes`; grows geometrically (levels grow run by run).
void* slot(FastBuffers& b, int s, size_t bytes) {
  if (b.bytes[s] < bytes) {
    const size_t cap = std::max<size_t>({bytes, 2 * b.bytes[s], 256});
    if (b.mem[s]) (void)hipFree(b.mem[s]);
    b.mem[s] = nullptr;
    b.bytes[s] = 0;
    FCHK(hipMalloc(&b.mem[s], cap));
    b.bytes[s] = cap;
  }
  return b.mem[s];
}

// host copy as float of a device-readable array (ordered after earlier work on st)
std::vector<float> to_float(const void* src, size_t n, int dtype, hipStream_t st) {
  std::vector<float> out(n);
  if (n == 0) return out;
  if (dtype == 1) {
    std::vector<double> d(n);
    FCHK(hipMemcpyAsync(d.data(), src, n * 8, hipMemcpyDefault, st));
    FCHK(hipStreamSynchronize(st));
    for (size_t i = 0; i < n; ++i) out[i] = (float)d[i];
  } else {
    FCHK(hipMemcpyAsync(out.data(), src, n * 4, hipMemcpyDefault, st));
    FCHK(hipStreamSynchronize(st));
  }
  return out;
}

uint16_t host_bf16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float host_bf16_f(uint16_t h) {
  uint32_t u = uint32_t(h) << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}

size_t gemm_lds(const FastNetDev& d, int prec) {
  // one buffer for a layer's input and its outputs (k_gemm; FP32: the error
  // estimates in its row paddings), the masks, the fp32 input lists
  const size_t ys = (size_t)d.Kpad * AM_XS * 4;
  const size_t xs = prec == 0 ? ys : std::max(ys, (size_t)2 * AM_GN * (d.Kpad + 8) * 2);
  const size_t kls = prec == 0 ? (size_t)d.D * d.Kpad * 2 : 0;
  return xs + (size_t)AM_GT * d.Wd * 4 + kls + 16;
}

}  // namespace

void fast_prepare(FastNet& f, FastBuffers& b, const NetDev& net, int precision, const void* const* weights,
                  const void* const* biases, int dtype, hipStream_t st) {
  FastNetDev& d = f.dev;
  std::memset(&d, 0, sizeof(d));
  f.precision = precision;
  for (int li = 1; li < net.D; ++li)
    if (net.nskip[li] > 0) throw std::runtime_error("skip connections run on the FP64 path only");
  for (int li = 1; li <= net.D; ++li)
    if (net.nodes[li + 1] > 64 * AM_GRBW)
      throw std::runtime_error("the reduced-precision forms take layers up to " + std::to_string(64 * AM_GRBW) +
                               " wide; use fp64 or fp64ray");
  d.D = net.D; d.L = net.L; d.E = net.E; d.C = net.C; d.Wd = net.Wd;
  d.Cs = (net.C + 63) & ~63;
  int maxw = 0;
  for (int i = 0; i <= net.D + 1; ++i) {
    d.nodes[i] = net.nodes[i];
    maxw = std::max(maxw, net.nodes[i]);
  }
  for (int i = 0; i <= net.D; ++i) d.off[i] = net.off[i];
  d.Kpad = (maxw + 31) & ~31;
  d.iso = (float)net.iso;
  const int prec = precision == 1 ? 1 : 0;
  // W0, biases, extras: FP32 copies built on the device (weights and biases
  // are device-readable here; the extras are the engine's FP64 copies)
  const int f64 = dtype == 1;
  {
    const int n0 = net.nodes[1] * 3;
    float* p = (float*)slot(b, S_W0, (size_t)n0 * 4);
    launch_transpose_pad(weights[0], f64, n0, 1, p, n0, 1, st);
    d.W0 = p;
  }
  for (int li = 0; li <= net.D; ++li) {
    const int nb = net.nodes[li + 1];
    float* p = (float*)slot(b, S_BIAS0 + li, (size_t)nb * 4);
    launch_transpose_pad(biases[li], f64, nb, 1, p, nb, 1, st);
    d.bias[li] = p;
  }
  {
    float* pw = (float*)slot(b, S_EW, (size_t)std::max(1, net.E * 3) * 4);
    float* pb = (float*)slot(b, S_EB, (size_t)std::max(1, net.E) * 4);
    if (net.E) {
      launch_transpose_pad(net.ew, 1, net.E * 3, 1, pw, net.E * 3, 1, st);
      launch_transpose_pad(net.eb, 1, net.E, 1, pb, net.E, 1, st);
    }
    d.ew = pw;
    d.eb = pb;
  }
  FCHK(hipGetLastError());
  // MFMA A operands of every layer after the first
  for (int li = 1; li <= net.D; ++li) {
    const int nin = net.nodes[li], nout = net.nodes[li + 1];
    const int nrb = (nout + 15) / 16;
    d.nrb[li] = nrb;
    if (prec == 0) {
      // W^T, k-major: [nin + 1][nrb * 16]; row nin (and rows past nout) are +0
      d.nkb[li] = (nin + 3) / 4;
      const int ld = nrb * 16;
      float* p = (float*)slot(b, S_WP0 + li, (size_t)(nin + 1) * ld * 4);
      launch_transpose_pad(weights[li], f64, nout, nin, p, ld, nin + 1, st);
      FCHK(hipGetLastError());
      d.Wp[li] = p;
    } else {
      auto w = to_float(weights[li], (size_t)nout * nin, dtype, st);
      const int nkb = (nin + 31) / 32;
      d.nkb[li] = nkb;
      const size_t n = (size_t)nrb * nkb * 64 * 8;
      std::vector<uint16_t> hi(2 * n, 0);
      for (int rb = 0; rb < nrb; ++rb)
        for (int kb = 0; kb < nkb; ++kb)
          for (int l = 0; l < 64; ++l)
            for (int e = 0; e < 8; ++e) {
              const int row = rb * 16 + (l & 15), k = kb * 32 + 8 * (l >> 4) + e;
              float x = (row < nout && k < nin) ? w[(size_t)row * nin + k] : 0.f;
              const uint16_t h = host_bf16(x);
              const size_t at = (((size_t)rb * nkb + kb) * 64 + l) * 8 + e;
              hi[at] = h;
              hi[n + at] = host_bf16(x - host_bf16_f(h));
            }
      uint16_t* p = (uint16_t*)slot(b, S_WP0 + li, hi.size() * 2);
      FCHK(hipMemcpy(p, hi.data(), hi.size() * 2, hipMemcpyHostToDevice));
      d.Wh[li] = p;
      d.Wl[li] = p + n;
    }
  }
  // error estimates (FP32 form): the variance per squared
  // term of each layer's sums, 4 u^2 (K + 2) for K = n_in products and the
  // bias: u^2 per rounding and u^2 K for the accumulation, times 4 because the
  // errors measured against FP64 on the matrix cores' accumulation order run
  // about 3x those of a blocked CPU sum.  The split-BF16 form carries no
  // estimates: its 16-bit products leave most regions undecidable even with
  // per-plane bounds, so the product would cost more than it saves.
  if (prec == 0) {
    const double u2 = std::ldexp(1.0, -48);
    for (int li = 1; li <= net.D; ++li) d.ck[li] = (float)(4.0 * u2 * (net.nodes[li] + 2));
    d.sig = 1;
  }
  const size_t lds = gemm_lds(d, prec);
  if (lds > 65536) {
    if (prec == 0)
      FCHK(hipFuncSetAttribute((const void*)k_gemm<0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    else
      FCHK(hipFuncSetAttribute((const void*)k_gemm<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  // FP32: plane error bounds lam sigma per plane plus 8 u of its largest
  // coefficient (the walk's own FP32 evaluations), lam = 3: the largest error
  // measured against FP64 on 1.5M planes of the 4x256 net was 0.7 sigma.
  // Vertex estimates (to keep a region's reduced vertices, reduced_vertices = 1,
  // or to skip a vertex's Newton step, reduced_vertices = 3) use 1 sigma.
  // Split-BF16: one relative bound for every plane, 2e-3 of its largest
  // coefficient (about 4x the largest error measured), and every vertex
  // re-solved from FP64 planes.
  const int vm = f.vertex_mode;
  const float lam_v = (vm == 1 || vm == 3) ? 1.f : (vm == 2 ? 0.f : -1.f);
  if (prec == 0) f.wp = WalkParams{4.8e-7f, 1e-3f, 1e3f, 3.f, lam_v, vm == 3 ? 1 : 0};
  else f.wp = WalkParams{2e-3f, 2e-2f, 1e2f, 0.f, -1.f, 0};
  f.ready = true;
}

namespace {
constexpr int kLevelChunk = 1 << 16;   // regions per GEMM + walk launch pair
int walk_sort_size(const NetDev& net) {
  int Csort = AM_TOPK;   // the walk's top-64 order lives in keys/order[0, 64)
  while (Csort < net.C) Csort <<= 1;
  return Csort;
}
size_t walk_scan_lds(const NetDev& net) {
  return (size_t)5 * net.C * 4 + (size_t)walk_sort_size(net) * 8 + (size_t)((net.C + 31) / 32) * 4 +
         (size_t)net.Wd * 4 + 64;
}
size_t walk_ray_lds(const NetDev& net) {
  return (((size_t)net.C * 20 + (size_t)net.Wd * 4 + 15) & ~(size_t)15) + walk_rec_bytes(net.vm, true);
}
}  // namespace

void fast_level_prepare(FastNet& f, FastBuffers& b, const NetDev& net, int count) {
  const FastNetDev& d = f.dev;
  const int chunk = std::min(count, kLevelChunk);
  (void)slot(b, S_PLANES, (size_t)chunk * AM_PCOMP * d.Cs * 4);
  (void)slot(b, S_EQ, (size_t)chunk * AM_EQS * 4);
  const size_t wlds = walk_scan_lds(net);
  if (wlds > 65536)
    FCHK(hipFuncSetAttribute((const void*)k_walk_f32, hipFuncAttributeMaxDynamicSharedMemorySize, (int)wlds));
  const size_t rlds = walk_ray_lds(net);
  if (rlds > 65536) {
    FCHK(hipFuncSetAttribute((const void*)k_walk_ray_f32<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)rlds));
    FCHK(hipFuncSetAttribute((const void*)k_walk_ray_f32<4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)rlds));
    FCHK(hipFuncSetAttribute((const void*)k_walk_ray_f32<8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)rlds));
  }
}

void fast_level(FastNet& f, FastBuffers& b, const NetDev& net, const FastLevel& lvl, hipStream_t st,
                StageTimer& timer, FastTimes& t, unsigned* flags, int* acc, const RedoList& redo, bool prepared) {
  const FastNetDev& d = f.dev;
  const int prec = f.precision == 1 ? 1 : 0;
  t.launches = 0;
  if (!prepared) fast_level_prepare(f, b, net, lvl.count);
  const int CH = kLevelChunk;
  const int chunk = std::min(lvl.count, CH);
  float* planes = (float*)slot(b, S_PLANES, (size_t)chunk * AM_PCOMP * d.Cs * 4);
  float* eqb = (float*)slot(b, S_EQ, (size_t)chunk * AM_EQS * 4);
  const int Csort = walk_sort_size(net);
  const int CsL = net.C;
  const size_t wlds = walk_scan_lds(net);
  const size_t rlds = walk_ray_lds(net);
  const size_t glds = gemm_lds(d, prec);

  for (int c0 = 0; c0 < lvl.count; c0 += CH) {
    const int n = std::min(CH, lvl.count - c0);
    GemmIO g;
    g.reg_mask = lvl.reg_mask;
    g.first = lvl.first + c0;
    g.count = n;
    g.planes = planes;
    g.eq = eqb;
    const size_t m0 = timer.mark(st);
    if (prec == 0)
      hipLaunchKernelGGL(k_gemm<0>, dim3((n + AM_GT - 1) / AM_GT), dim3(256), glds, st, d, g);
    else
      hipLaunchKernelGGL(k_gemm<1>, dim3((n + AM_GT - 1) / AM_GT), dim3(256), glds, st, d, g);
    FCHK(hipGetLastError());
    const size_t m1 = timer.mark(st);
    timer.span(m0, m1, T_GEMM);

    RegionIO io;
    io.reg_mask = lvl.reg_mask;
    io.reg_start = lvl.reg_start;
    io.reg_point = lvl.reg_point;
    io.reg_info = lvl.reg_info;
    io.reg_tab = lvl.reg_tab;
    io.reg_verts = lvl.reg_verts;
    io.reg_vpair = lvl.reg_vpair;
    io.cand_mask = lvl.cand_mask;
    io.cand_start = lvl.cand_start;
    io.cand_point = lvl.cand_point;
    io.cand_hash = lvl.cand_hash;
    io.reg_jinv = lvl.reg_jinv;
    io.first = lvl.first;          // candidate slots / flags are level-relative
    io.count = n;
    io.Cs = CsL;
    io.Csort = Csort;
    io.flags = flags;
    io.ids = nullptr;
    if (f.walk_mode == 0 && f.ray_waves == 2)
      hipLaunchKernelGGL(k_walk_ray_f32<2>, dim3(n), dim3(128), rlds, st, net, io, f.wp, planes, eqb, d.Cs,
                         lvl.first + c0);
    else if (f.walk_mode == 0 && f.ray_waves == 8)
      hipLaunchKernelGGL(k_walk_ray_f32<8>, dim3(n), dim3(512), rlds, st, net, io, f.wp, planes, eqb, d.Cs,
                         lvl.first + c0);
    else if (f.walk_mode == 0)
      hipLaunchKernelGGL(k_walk_ray_f32<4>, dim3(n), dim3(256), rlds, st, net, io, f.wp, planes, eqb, d.Cs,
                         lvl.first + c0);
    else
      hipLaunchKernelGGL(k_walk_f32, dim3(n), dim3(64), wlds, st, net, io, f.wp, planes, eqb, d.Cs,
                         lvl.first + c0);
    FCHK(hipGetLastError());
    timer.span(m1, timer.mark(st), T_WALK);
    t.launches += 1;
  }
  const int tiles = (int)redo_state_words(lvl.count) - 2;
  hipLaunchKernelGGL(k_redo_list, dim3(tiles), dim3(kRedoThreads), 0, st, (const unsigned*)flags, lvl.count,
                     lvl.first, acc, redo, tiles);
  FCHK(hipGetLastError());
}

void fast_planes(FastNet& f, const uint32_t* reg_mask, int n, float* planes, float* eq, hipStream_t st) {
  const int prec = f.precision == 1 ? 1 : 0;
  GemmIO g;
  g.reg_mask = reg_mask;
  g.first = 0;
  g.count = n;
  g.planes = planes;
  g.eq = eq;
  const size_t glds = gemm_lds(f.dev, prec);
  if (prec == 0)
    hipLaunchKernelGGL(k_gemm<0>, dim3((n + AM_GT - 1) / AM_GT), dim3(256), glds, st, f.dev, g);
  else
    hipLaunchKernelGGL(k_gemm<1>, dim3((n + AM_GT - 1) / AM_GT), dim3(256), glds, st, f.dev, g);
  FCHK(hipGetLastError());
  FCHK(hipStreamSynchronize(st));
}

void fast_release(FastBuffers& b) {
  for (int i = 0; i < 64; ++i) {
    if (b.mem[i]) (void)hipFree(b.mem[i]);
    b.mem[i] = nullptr;
    b.bytes[i] = 0;
  }
}

}  // namespace am
