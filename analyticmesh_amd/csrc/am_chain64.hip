// FP64 planes of many regions at once, bit-identical to the one-region chain.
//
// The masked affine chain (backend/inc/process.h:18-193) is a GEMM per layer:
// Y = W_li * X with X = diag(m) P_{li-1}, one column per plane component (x, y,
// z, c) of each region.  A workgroup takes AM_CT regions = 64 columns, one per
// lane, and AM_CHAIN_MAXW / AM_CR waves, each owning AM_CR output rows.  Every
// lane runs the whole k loop in ascending order with one fma per step, so each
// output is the same sequence of roundings as the canonical chain:
// accumulators start at (0, 0, 0, b_j + 0) and take fma(W[j][k], m_k ? P_k : +0).
// An inactive k adds fma(w, +0, acc) = acc exactly (no accumulator is ever -0:
// each starts at +0 or a nonzero bias, and a sum of opposite-signed zeros
// rounds to +0), so zero-filling the inactive inputs reproduces skipping them.
//
// Operand delivery: X[k][lane] is a private LDS read per lane; the wave's 16
// weights W[k][row0..row0+15] arrive as one coalesced vector load (lane l
// holds row l & 15) and each fma takes its row's weight through a DPP row
// broadcast, so a weight costs no extra instruction.  Weights and X of the next
// AM_CPF k steps are in flight while a step's fmas run.  After the k loop, once every wave is
// done reading X, the waves park their rows in LDS; the tile is then written
// to HBM signed (a.x + c <= 0 inside) with consecutive threads on consecutive
// constraints, and masked in place as the next layer's input.
//
// k_walk_fp64 is the polygon walk of the region kernel on those planes.
#include "am_device.h"
#include "am_kernels.h"
#include "am_walk.h"

namespace am {


// acc[j] = fma(w from lane j of this lane's 16-lane row, x, acc[j]), j < 16:
// v_fmac_f64 with a row_newbcast DPP source is one fused multiply-add per
// output.  The leading s_nop covers the DPP read-after-VALU-write hazard on w.
__device__ __forceinline__ void chain_step16(double (&acc)[AM_CR], double w, double x) {
  static_assert(AM_CR == 16, "one DPP row broadcast per output row");
  asm volatile(
      "s_nop 1\n"
      "v_fmac_f64_dpp %0, %16, %17 row_newbcast:0 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %1, %16, %17 row_newbcast:1 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %2, %16, %17 row_newbcast:2 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %3, %16, %17 row_newbcast:3 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %4, %16, %17 row_newbcast:4 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %5, %16, %17 row_newbcast:5 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %6, %16, %17 row_newbcast:6 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %7, %16, %17 row_newbcast:7 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %8, %16, %17 row_newbcast:8 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %9, %16, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %10, %16, %17 row_newbcast:10 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %11, %16, %17 row_newbcast:11 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %12, %16, %17 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %13, %16, %17 row_newbcast:13 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %14, %16, %17 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
      "v_fmac_f64_dpp %15, %16, %17 row_newbcast:15 row_mask:0xf bank_mask:0xf\n"
      : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]), "+v"(acc[5]), "+v"(acc[6]),
        "+v"(acc[7]), "+v"(acc[8]), "+v"(acc[9]), "+v"(acc[10]), "+v"(acc[11]), "+v"(acc[12]), "+v"(acc[13]),
        "+v"(acc[14]), "+v"(acc[15])
      : "v"(w), "v"(x));
}

// 16 consecutive k indices (uint16, 32-byte aligned in LDS) into 8 uniform words.
__device__ __forceinline__ void load_k16(const uint16_t* src, uint32_t (&k)[AM_CPF / 2]) {
  static_assert(AM_CPF == 16, "two 16-byte reads per group");
  const uint4 a = reinterpret_cast<const uint4*>(src)[0];
  const uint4 b = reinterpret_cast<const uint4*>(src)[1];
  k[0] = __builtin_amdgcn_readfirstlane(a.x); k[1] = __builtin_amdgcn_readfirstlane(a.y);
  k[2] = __builtin_amdgcn_readfirstlane(a.z); k[3] = __builtin_amdgcn_readfirstlane(a.w);
  k[4] = __builtin_amdgcn_readfirstlane(b.x); k[5] = __builtin_amdgcn_readfirstlane(b.y);
  k[6] = __builtin_amdgcn_readfirstlane(b.z); k[7] = __builtin_amdgcn_readfirstlane(b.w);
}

// Plane g of a region (x, y, z, c) as the chain defines it.  Hidden layer 1's
// planes are the rows of (W0 | b0), negated where the neuron is active; the
// chain does not store them, they are rebuilt here from the weights and the
// region's mask.  The others come from the chain's record at src.
__device__ __forceinline__ void plane64(const NetDev& net, const uint32_t* mask, const double* src, int Cs_hbm, int g,
                                        double* p) {
  if (g < net.nodes[1]) {
    const bool on = mbit(mask, g);
    const double v[4] = {net.W0[g * 3 + 0], net.W0[g * 3 + 1], net.W0[g * 3 + 2], net.b[0][g]};
    for (int q = 0; q < 4; ++q) p[q] = on ? -v[q] : v[q];
  } else {
    for (int q = 0; q < 4; ++q) p[q] = src[(size_t)q * Cs_hbm + g];
  }
}

// One region's planes into LDS, threads t0, t0 + nt, ...
__device__ __forceinline__ void load_planes64(const NetDev& net, const uint32_t* mask, const double* src, int Cs_hbm,
                                              double* Px, double* Py, double* Pz, double* Pc, int t0, int nt) {
  for (int g = t0; g < net.C; g += nt) {
    double p[4];
    plane64(net, mask, src, Cs_hbm, g, p);
    Px[g] = p[0];
    Py[g] = p[1];
    Pz[g] = p[2];
    Pc[g] = p[3];
  }
}

// The planes of one tile: regions r0 .. r0 + nreg - 1 of lv (ids[r0 + t] or
// first + r0 + t), written at planes slots r0 + t.  io supplies the region
// store, the plane stride and the layer operands.
// MF: the layer products on the FP64 matrix cores (four 16 x 16 blocks of
// v_mfma_f64_16x16x4f64 per wave and 4 list entries; A: lane l holds
// W[row0 + l % 16][list entry i0 + l / 16], B: X[i0 + l / 16][16 cb + l % 16];
// result element v of lane l is row row0 + l / 16 + 4 v, column 16 cb + l % 16),
// the bias added at the end: the planes equal the sequential-fma ones up to the
// rounding of a different summation order, so the FP64 form (records bit for
// bit) keeps MF = false.
typedef double am_f64x4 __attribute__((ext_vector_type(4)));

template <bool MF>
__device__ __forceinline__ void chain_tile_body(const NetDev& net, const ChainIO& io, const ChainLevel& lv, int r0,
                                                int nreg) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // LDS: X, one row per position of the layer's active-input list (rows past
  // the list are read by the operand prefetch only), each layer's list of
  // active inputs, the tile's masks
  double* Xs = reinterpret_cast<double*>(smem_raw);                                   // [AM_CXR][AM_CXS]
  uint16_t* klist = reinterpret_cast<uint16_t*>(Xs + (size_t)AM_CXR * AM_CXS + 1);     // [D][AM_KLS], 16-B aligned
  uint32_t* mk = reinterpret_cast<uint32_t*>(klist + (size_t)net.D * AM_KLS);          // [AM_CT][Wd]
  __shared__ int s_nk[AM_MAX_LINEAR];
  for (int idx = tid; idx < AM_CT * net.Wd; idx += AM_CHAIN_THREADS) {
    const int t = idx / net.Wd, w = idx - t * net.Wd;
    uint32_t v = 0u;
    if (t < nreg) {
      const int rid = lv.ids ? lv.ids[r0 + t] : lv.first + r0 + t;
      v = io.reg_mask[(size_t)rid * net.Wd + w];
    }
    mk[idx] = v;
  }
  __syncthreads();
  // Active inputs of each layer for the whole tile: input k of linear layer li
  // is skipped when it is inactive in all AM_CT regions (its X row is +0 in
  // every column, and fma(w, +0, acc) = acc).  Wave li-1 lists layer li's in
  // ascending order and pads the list with the +0 row AM_CHAIN_MAXW up to a
  // multiple of AM_CPF, plus 2 AM_CPF more entries for the operand pipelines.
  if (wave < net.D) {
    const int li = wave + 1;
    const int nin = net.nodes[li], in0 = li == 1 ? 0 : net.off[li - 1];
    uint16_t* kl = klist + (size_t)wave * AM_KLS;
    int base = 0;
    for (int kb = 0; kb < nin; kb += 64) {
      const int k = kb + lane;
      bool act = false;
      if (k < nin)
        for (int t = 0; t < AM_CT; ++t) act = act || mbit(mk + t * net.Wd, in0 + k);
      const unsigned long long bal = __ballot(act);
      if (act) kl[base + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u))] = (uint16_t)k;
      base += __popcll(bal);
    }
    const int nkp = (base + AM_CPF - 1) / AM_CPF * AM_CPF;
    for (int i = base + lane; i < nkp + 2 * AM_CPF; i += 64) kl[i] = (uint16_t)AM_CHAIN_MAXW;
    if (lane == 0) s_nk[li] = nkp;
  }
  __syncthreads();
  const size_t Cs = (size_t)io.Cs;
  // hidden node 1: the rows of W0, unsigned and masked into X at their list
  // positions (their signed planes are not stored: readers rebuild them, plane64)
  const int n1 = net.nodes[1];
  for (int idx = tid; idx < s_nk[1] * AM_CN; idx += AM_CHAIN_THREADS) {
    const int i = idx / AM_CN, col = idx - i * AM_CN, t = col >> 2, q = col & 3;
    const int k = klist[i];
    double x = 0.0;
    if (k < n1 && mbit(mk + t * net.Wd, k)) x = q < 3 ? net.W0[k * 3 + q] : net.b[0][k];
    Xs[i * AM_CXS + col] = x;
  }
  for (int idx = tid; idx < AM_CN * net.E; idx += AM_CHAIN_THREADS) {
    const int col = idx / net.E, e = idx - col * net.E, t = col >> 2, q = col & 3;
    if (t < nreg) lv.planes[((size_t)(r0 + t) * 4 + q) * Cs + net.L + e] = q < 3 ? net.ew[e * 3 + q] : net.eb[e];
  }
  __syncthreads();
  const int q_l = lane & 3;           // this lane's column is x, y, z or c of region lane >> 2
  for (int li = 1; li <= net.D; ++li) {
    const int nout = net.nodes[li + 1], noutp = io.noutp[li];
    const int row0 = wave * AM_CR;
    // A layer narrower than one row block (the output layer): wave j < nout
    // computes row j alone, the same fma sequence with the weight as a scalar
    // operand.  Lane l first loads the weights of list positions l, l + 64, ...
    const bool narrow = nout < AM_CR;
    const bool active = !narrow && row0 < nout;   // wave-uniform
    double acc[AM_CR];
    am_f64x4 acc4[AM_CN / 16];
    double acc_n = 0.0;
    if (narrow && wave < nout) {
      const uint16_t* kl = klist + (size_t)(li - 1) * AM_KLS;
      const int nkp = s_nk[li];
      const double* wr = io.Wp[li] + wave;
      double wv[AM_CHAIN_MAXW / 64];
#pragma unroll
      for (int c = 0; c < AM_CHAIN_MAXW / 64; ++c) {
        const int i = lane + 64 * c;
        wv[c] = i < nkp ? wr[(size_t)kl[i] * noutp] : 0.0;
      }
      double a = q_l == 3 ? io.bp[li][wave] + 0.0 : 0.0;
#pragma unroll
      for (int c = 0; c < AM_CHAIN_MAXW / 64; ++c) {
        const int lo = __double2loint(wv[c]), hi = __double2hiint(wv[c]);
        const int n = min(64, nkp - 64 * c);
        for (int t = 0; t < n; ++t) {
          const double w = __hiloint2double(__builtin_amdgcn_readlane(hi, t), __builtin_amdgcn_readlane(lo, t));
          a = __builtin_fma(w, Xs[(64 * c + t) * AM_CXS + lane], a);
        }
      }
      acc_n = a;
    }
    if (MF && active) {
      const uint16_t* kl = klist + (size_t)(li - 1) * AM_KLS;
      const int nkp = s_nk[li];   // a multiple of AM_CPF; AM_CPF more entries name the +0 row
      const int l16 = lane & 15, l4 = lane >> 4;
      const double* wcol = io.Wp[li] + row0 + l16;
#pragma unroll
      for (int cb = 0; cb < AM_CN / 16; ++cb) acc4[cb] = am_f64x4{0.0, 0.0, 0.0, 0.0};
      // A fragments in flight: one list group ahead (two measured slower)
      constexpr int PF = AM_CPF / 4;
      double a[PF];
#pragma unroll
      for (int u = 0; u < PF; ++u) a[u] = wcol[(size_t)kl[4 * u + l4] * noutp];
      for (int i0 = 0; i0 < nkp; i0 += AM_CPF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int ii = i0 + 4 * u;
          double b[AM_CN / 16];
#pragma unroll
          for (int cb = 0; cb < AM_CN / 16; ++cb) b[cb] = Xs[(size_t)(ii + l4) * AM_CXS + 16 * cb + l16];
          const double av = a[u];
          a[u] = wcol[(size_t)kl[ii + AM_CPF + l4] * noutp];
#pragma unroll
          for (int cb = 0; cb < AM_CN / 16; ++cb)
            acc4[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b[cb], acc4[cb], 0, 0, 0);
        }
      }
    }
    if (!MF && active) {
      const double* bias = io.bp[li] + row0;
#pragma unroll
      for (int j = 0; j < AM_CR; ++j) acc[j] = q_l == 3 ? bias[j] + 0.0 : 0.0;
      // lane l carries W[k][row0 + (l & 15)]; each fma takes row j's weight by a
      // DPP row broadcast of that register
      // Weights through a buffer resource: row k's byte offset k * noutp * 8 is
      // uniform (the scalar offset), the lane's element row0 + (lane & 15) the
      // vector offset, so a weight load needs no vector address arithmetic.
      // Wp holds rows 0..AM_CHAIN_MAXW (the last is the +0 padding row).
      const __amdgpu_buffer_rsrc_t wres = __builtin_amdgcn_make_buffer_rsrc(
          (void*)io.Wp[li], (short)0, (AM_CHAIN_MAXW + 1) * noutp * 8, 0x00020000);
      const int woff = (row0 + (lane & 15)) * 8;
      const int wrow = noutp * 8;
      // k runs over the tile's active inputs in ascending order; X row i holds
      // list entry i (padding entries name the +0 weight row and have +0 X:
      // fma(+0, +0, acc) = acc).  Operands of the next AM_CPF steps are in
      // flight while a step's fmas run, and the k indices of the group after
      // are read a group ahead.
      const uint16_t* kl = klist + (size_t)(li - 1) * AM_KLS;
      const int nkp = s_nk[li];
      double wq[AM_CPF], xq[AM_CPF];
      uint32_t kn[AM_CPF / 2];   // k indices of the group after next, two per word (uniform: SGPRs)
      load_k16(kl, kn);
#pragma unroll
      for (int u = 0; u < AM_CPF; ++u) {
        const int k = (int)((kn[u >> 1] >> ((u & 1) * 16)) & 0xffffu);
        wq[u] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(wres, woff, k * wrow, 0));
        xq[u] = Xs[u * AM_CXS + lane];
        __builtin_amdgcn_sched_barrier(0);
      }
      load_k16(kl + AM_CPF, kn);
      for (int i = 0; i < nkp; i += AM_CPF) {
        const double* xnext = Xs + (size_t)(i + AM_CPF) * AM_CXS + lane;   // rows of the next group
        uint32_t kk[AM_CPF / 2];
#pragma unroll
        for (int u = 0; u < AM_CPF / 2; ++u) kk[u] = kn[u];
        // indices of the group after next (past the end: unused entries of the list)
        load_k16(kl + i + 2 * AM_CPF, kn);
#pragma unroll
        for (int u = 0; u < AM_CPF; ++u) {
          chain_step16(acc, wq[u], xq[u]);
          // refill the slot just used, AM_CPF steps ahead; the scheduling
          // barrier keeps the loads in slot order, so each step waits only
          // for its own (oldest) load
          const int k = (int)((kk[u >> 1] >> ((u & 1) * 16)) & 0xffffu);
          wq[u] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(wres, woff, k * wrow, 0));
          xq[u] = xnext[u * AM_CXS];
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    __syncthreads();                  // every wave is done reading this layer's input
    // park the layer's outputs in LDS (unsigned, unmasked), row = output index
    if (MF && active) {
      const int l16 = lane & 15, l4 = lane >> 4;
      const bool cbias = (l16 & 3) == 3;   // columns 16 cb + l16: component c
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = row0 + l4 + 4 * v;
        if (row < nout) {
          const double b = cbias ? io.bp[li][row] : 0.0;
#pragma unroll
          for (int cb = 0; cb < AM_CN / 16; ++cb) Xs[row * AM_CXS + 16 * cb + l16] = acc4[cb][v] + b;
        }
      }
    }
    if (!MF && active) {
#pragma unroll
      for (int j = 0; j < AM_CR; ++j) {
        const int row = row0 + j;
        if (row < nout) Xs[row * AM_CXS + lane] = acc[j];
      }
    }
    if (narrow && wave < nout) Xs[wave * AM_CXS + lane] = acc_n;
    __syncthreads();
    if (li < net.D) {
      // signed planes to HBM, consecutive threads on consecutive constraints
      // (thread: row jt, columns ct, ct + AM_CHAIN_THREADS / 256, ...)
      const int out0 = net.off[li];
      const int jt = tid & 255, ct = tid >> 8;
      for (int col = ct; col < AM_CN; col += AM_CHAIN_THREADS / 256) {
        const int t = col >> 2, q = col & 3;
        if (t >= nreg) break;
        double* dst = lv.planes + ((size_t)(r0 + t) * 4 + q) * Cs + out0;
        for (int j = jt; j < nout; j += 256) {
          const double y = Xs[j * AM_CXS + col];
          dst[j] = mbit(mk + t * net.Wd, out0 + j) ? -y : y;
        }
      }
      // the next layer's input at its list positions: position i holds output
      // kl2[i] masked (+0 for padding entries).  The list ascends, so position
      // i reads row kl2[i] >= i: in two halves of positions, each half reads
      // (rows at or past its first position) before a barrier and writes
      // (its own positions) after it, and no half overwrites a row a later
      // half still reads.
      const uint16_t* kl2 = klist + (size_t)li * AM_KLS;
      const int nkp2 = s_nk[li + 1];
      constexpr int PPT = AM_CHAIN_THREADS / AM_CN;                 // positions per pass of the block
      constexpr int XH = AM_CHAIN_MAXW / PPT / 2;                   // passes per half
      // (an empty asm keeps the thread's offsets from being hoisted out of the
      // layer loop, where they would stay live through the k loop)
      int tq = tid;
      asm volatile("" : "+v"(tq));
      const int col = tq & (AM_CN - 1), t = col >> 2, p0 = tq / AM_CN;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        double xv[XH];
#pragma unroll
        for (int c = 0; c < XH; ++c) {
          const int i = (h * XH + c) * PPT + p0;
          xv[c] = 0.0;
          if (i < nkp2) {
            const int k = kl2[i];
            if (k < nout && mbit(mk + t * net.Wd, out0 + k)) xv[c] = Xs[k * AM_CXS + col];
          }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < XH; ++c) {
          const int i = (h * XH + c) * PPT + p0;
          if (i < nkp2) Xs[i * AM_CXS + col] = xv[c];
        }
      }
    } else if (tid < AM_CN && (tid >> 2) < nreg) {
      const double y = Xs[tid];
      lv.eq[(size_t)(r0 + (tid >> 2)) * 4 + (tid & 3)] = (tid & 3) == 3 ? y - net.iso : y;
    }
    __syncthreads();
  }
}

// The matrix-core chain on half tiles: 8 regions (32 columns) per workgroup of
// 8 waves, each wave two 16-row blocks (rows 32 w .. 32 w + 31) by two 16-column
// blocks, so one A fragment feeds two products and one B fragment two.  Its
// LDS (about 75 KB) lets two workgroups share a CU: one runs its products while
// the other parks, stores and compacts a layer.  Planes as chain_tile_body.
__device__ __forceinline__ void chain_tile_half(const NetDev& net, const ChainIO& io, const ChainLevel& lv, int r0,
                                                int nreg) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  double* Xs = reinterpret_cast<double*>(smem_raw);                                   // [AM_CXR][AM_HXS]
  uint16_t* klist = reinterpret_cast<uint16_t*>(Xs + (size_t)AM_CXR * AM_HXS + 1);     // [D][AM_KLS]
  uint32_t* mk = reinterpret_cast<uint32_t*>(klist + (size_t)net.D * AM_KLS);          // [AM_HT][Wd]
  __shared__ int s_nk[AM_MAX_LINEAR];
  for (int idx = tid; idx < AM_HT * net.Wd; idx += AM_HALF_THREADS) {
    const int t = idx / net.Wd, w = idx - t * net.Wd;
    uint32_t v = 0u;
    if (t < nreg) {
      const int rid = lv.ids ? lv.ids[r0 + t] : lv.first + r0 + t;
      v = io.reg_mask[(size_t)rid * net.Wd + w];
    }
    mk[idx] = v;
  }
  __syncthreads();
  // active-input lists as in chain_tile_body (waves take layers in turn)
  for (int li = wave + 1; li <= net.D; li += AM_HALF_THREADS / 64) {
    const int nin = net.nodes[li], in0 = li == 1 ? 0 : net.off[li - 1];
    uint16_t* kl = klist + (size_t)(li - 1) * AM_KLS;
    int base = 0;
    for (int kb = 0; kb < nin; kb += 64) {
      const int k = kb + lane;
      bool act = false;
      if (k < nin)
        for (int t = 0; t < AM_HT; ++t) act = act || mbit(mk + t * net.Wd, in0 + k);
      const unsigned long long bal = __ballot(act);
      if (act) kl[base + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u))] = (uint16_t)k;
      base += __popcll(bal);
    }
    const int nkp = (base + AM_CPF - 1) / AM_CPF * AM_CPF;
    for (int i = base + lane; i < nkp + AM_CPF; i += 64) kl[i] = (uint16_t)AM_CHAIN_MAXW;
    if (lane == 0) s_nk[li] = nkp;
  }
  __syncthreads();
  const size_t Cs = (size_t)io.Cs;
  const int n1 = net.nodes[1];
  // layer-1 input rows: a thread takes list position i and half h of the tile's
  // regions, reads its row of (W0 | b0) once and writes 4 regions x 4 values
  {
    static_assert(AM_HALF_THREADS == 512 && AM_CHAIN_MAXW <= 256, "two halves of 256 positions");
    const int h = tid >> 8;
    for (int i = tid & 255; i < s_nk[1]; i += 256) {
      const int k = klist[i];
      double w[4] = {0.0, 0.0, 0.0, 0.0};
      if (k < n1) {
        w[0] = net.W0[k * 3 + 0];
        w[1] = net.W0[k * 3 + 1];
        w[2] = net.W0[k * 3 + 2];
        w[3] = net.b[0][k];
      }
#pragma unroll
      for (int tt = 0; tt < AM_HT / 2; ++tt) {
        const int t = h * (AM_HT / 2) + tt;
        const bool on = k < n1 && mbit(mk + t * net.Wd, k);
#pragma unroll
        for (int q = 0; q < 4; ++q) Xs[i * AM_HXS + 4 * t + q] = on ? w[q] : 0.0;
      }
    }
  }
  for (int idx = tid; idx < AM_HN * net.E; idx += AM_HALF_THREADS) {
    const int col = idx / net.E, e = idx - col * net.E, t = col >> 2, q = col & 3;
    if (t < nreg) lv.planes[((size_t)(r0 + t) * 4 + q) * Cs + net.L + e] = q < 3 ? net.ew[e * 3 + q] : net.eb[e];
  }
  __syncthreads();
  const int l16 = lane & 15, l4 = lane >> 4;
  for (int li = 1; li <= net.D; ++li) {
    const int nout = net.nodes[li + 1], noutp = io.noutp[li];
    const int row0 = wave * 32;
    // row blocks of this wave with rows (W^T holds noutp >= nout columns, zero past nout)
    const int nrb = row0 >= nout ? 0 : (row0 + 16 >= noutp ? 1 : 2);
    am_f64x4 acc[2][2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = am_f64x4{0.0, 0.0, 0.0, 0.0};
    if (nrb > 0) {
      const uint16_t* kl = klist + (size_t)(li - 1) * AM_KLS;
      const int nkp = s_nk[li];
      const double* wcol = io.Wp[li] + row0 + l16;
      const int rb1 = nrb > 1 ? 16 : 0;   // the second block's column offset (re-reads the first when absent)
      constexpr int PF = AM_CPF / 4;
      double a0[PF], a1[PF];
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const double* w = wcol + (size_t)kl[4 * u + l4] * noutp;
        a0[u] = w[0];
        a1[u] = w[rb1];
      }
      for (int i0 = 0; i0 < nkp; i0 += AM_CPF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int ii = i0 + 4 * u;
          const double b0 = Xs[(size_t)(ii + l4) * AM_HXS + l16];
          const double b1 = Xs[(size_t)(ii + l4) * AM_HXS + 16 + l16];
          const double x0 = a0[u], x1 = a1[u];
          const double* w = wcol + (size_t)kl[ii + AM_CPF + l4] * noutp;
          a0[u] = w[0];
          a1[u] = w[rb1];
          acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, b0, acc[0][0], 0, 0, 0);
          acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, b1, acc[0][1], 0, 0, 0);
          if (nrb > 1) {
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, b1, acc[1][1], 0, 0, 0);
          }
        }
      }
    }
    // the park's biases (column c of each region: lanes with l16 % 4 == 3),
    // loaded while the other waves finish their products
    const bool cbias = (l16 & 3) == 3;
    double bias[2][4];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = row0 + 16 * rb + l4 + 4 * v;
        bias[rb][v] = cbias && rb < nrb && row < nout ? io.bp[li][row] : 0.0;
      }
    __syncthreads();                  // every wave is done reading this layer's input
    {
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        if (rb >= nrb) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = row0 + 16 * rb + l4 + 4 * v;
          if (row < nout) {
            const double b = bias[rb][v];
            Xs[row * AM_HXS + l16] = acc[rb][0][v] + b;
            Xs[row * AM_HXS + 16 + l16] = acc[rb][1][v] + b;
          }
        }
      }
    }
    __syncthreads();
    if (li < net.D) {
      const int out0 = net.off[li];
      const int jt = tid & 255, ct = tid >> 8;
      for (int col = ct; col < AM_HN; col += AM_HALF_THREADS / 256) {
        const int t = col >> 2, q = col & 3;
        if (t >= nreg) break;
        double* dst = lv.planes + ((size_t)(r0 + t) * 4 + q) * Cs + out0;
        for (int j = jt; j < nout; j += 256) {
          const double y = Xs[j * AM_HXS + col];
          dst[j] = mbit(mk + t * net.Wd, out0 + j) ? -y : y;
        }
      }
      const uint16_t* kl2 = klist + (size_t)li * AM_KLS;
      const int nkp2 = s_nk[li + 1];
      constexpr int PPT = AM_HALF_THREADS / AM_HN;                  // positions per pass of the block
      constexpr int XH = AM_CHAIN_MAXW / PPT / 2;                   // passes per half
      int tq = tid;
      asm volatile("" : "+v"(tq));
      const int col = tq & (AM_HN - 1), t = col >> 2, p0 = tq / AM_HN;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        double xv[XH];
#pragma unroll
        for (int c = 0; c < XH; ++c) {
          const int i = (h * XH + c) * PPT + p0;
          xv[c] = 0.0;
          if (i < nkp2) {
            const int k = kl2[i];
            if (k < nout && mbit(mk + t * net.Wd, out0 + k)) xv[c] = Xs[k * AM_HXS + col];
          }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < XH; ++c) {
          const int i = (h * XH + c) * PPT + p0;
          if (i < nkp2) Xs[i * AM_HXS + col] = xv[c];
        }
      }
    } else if (tid < AM_HN && (tid >> 2) < nreg) {
      const double y = Xs[tid];
      lv.eq[(size_t)(r0 + (tid >> 2)) * 4 + (tid & 3)] = (tid & 3) == 3 ? y - net.iso : y;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(AM_CHAIN_THREADS) void k_chain_tile_fp64(NetDev net, ChainIO io) {
  if (io.zero && blockIdx.x == 0 && threadIdx.x == 0) *io.zero = 0;
  const int r0 = blockIdx.x * AM_CT;
  const int cnt = launch_count(io);
  if (r0 >= cnt) return;
  const ChainLevel lv{io.first, cnt, io.ids, io.planes, io.eq};
  chain_tile_body<false>(net, io, lv, r0, min(AM_CT, cnt - r0));
}

__global__ __launch_bounds__(AM_HALF_THREADS) void k_chain_tile_mfma(NetDev net, ChainIO io) {
  if (io.zero && blockIdx.x == 0 && threadIdx.x == 0) *io.zero = 0;
  const int r0 = blockIdx.x * AM_HT;
  const int cnt = launch_count(io);
  if (r0 >= cnt) return;
  const ChainLevel lv{io.first, cnt, io.ids, io.planes, io.eq};
  chain_tile_half(net, io, lv, r0, min(AM_HT, cnt - r0));
}

// Many networks in one launch: block b takes tiles[b] = (network, r0, regions);
// the tile's regions are lv.first + r0 + t, all of one network.
template <bool MF>
__global__ __launch_bounds__(AM_CHAIN_THREADS) void k_chain_tile_fp64_multi(const NetDev* nets, const ChainIO* cios,
                                                                           ChainLevel lv, const int4* tiles) {
  const int4 tl = tiles[blockIdx.x];
  chain_tile_body<MF>(nets[tl.x], cios[tl.x], lv, tl.y, tl.z);
}
template __global__ void k_chain_tile_fp64_multi<false>(const NetDev*, const ChainIO*, ChainLevel, const int4*);
template __global__ void k_chain_tile_fp64_multi<true>(const NetDev*, const ChainIO*, ChainLevel, const int4*);

// The candidate scan of one region (rid) on the planes at slot sl.
__device__ __forceinline__ void walk_fp64_body(const NetDev& net, const RegionIO& io, int rid, int sl,
                                               const double* planes, const double* eqb, int Cs_hbm) {
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
  for (int w = lane; w < net.Wd; w += 64) smask[w] = io.reg_mask[(size_t)rid * net.Wd + w];
  for (int w = lane; w < ((C + 31) >> 5); w += 64) used[w] = 0u;
  load_planes64(net, io.reg_mask + (size_t)rid * net.Wd, planes + (size_t)sl * 4 * Cs_hbm, Cs_hbm, Px, Py, Pz,
                Pc, lane, 64);
  double p[3];
  p[0] = io.reg_point[(size_t)rid * 3 + 0];
  p[1] = io.reg_point[(size_t)rid * 3 + 1];
  p[2] = io.reg_point[(size_t)rid * 3 + 2];
  const double eq[4] = {eqb[(size_t)sl * 4 + 0], eqb[(size_t)sl * 4 + 1], eqb[(size_t)sl * 4 + 2],
                        eqb[(size_t)sl * 4 + 3]};
  __syncthreads();
  const WalkParams wp{0.f, 0.f, 0.f};
  walk_region<double, false>(net, io, wp, rid, rid - io.first, Px, Py, Pz, Pc, keys, order, used, smask, eq, p);
  __syncthreads();   // this region's LDS is free for the next
}

__global__ __launch_bounds__(64) void k_walk_fp64(NetDev net, RegionIO io, int rid0, const int* count,
                                                  const int* slots, const double* planes, const double* eqb,
                                                  int Cs_hbm) {
  const int total = count ? min(*count, io.count) : io.count;
  // block-stride over the regions (one iteration per block when the grid covers them)
  for (int b = blockIdx.x; b < total; b += gridDim.x) {
    const int sl = slots ? slots[b] : b;     // planes slot of this region
    const int rid = io.ids ? io.ids[sl] : rid0 + sl;
    walk_fp64_body(net, io, rid, sl, planes, eqb, Cs_hbm);
  }
}

// Network of a region in a batch of networks: the last word of its mask.
__device__ __forceinline__ int region_net(const RegionIO& io, int Wd, int rid) {
  return (int)io.reg_mask[(size_t)rid * Wd + Wd - 1];
}

// LDS layout of the candidate scan for a network of C constraints (the host
// sizes the launch for the widest network of the batch).
__device__ __forceinline__ RegionIO scan_layout(const RegionIO& io, int C) {
  RegionIO r = io;
  int n2 = AM_TOPK;
  while (n2 < C) n2 <<= 1;
  r.Cs = C;
  r.Csort = n2;
  return r;
}

__global__ __launch_bounds__(64) void k_walk_fp64_multi(const NetDev* nets, RegionIO io, int Wd, int rid0,
                                                        const int* count, const int* slots, const double* planes,
                                                        const double* eqb, int Cs_hbm) {
  const int total = count ? min(*count, io.count) : io.count;
  for (int b = blockIdx.x; b < total; b += gridDim.x) {
    const int sl = slots ? slots[b] : b;
    const int rid = rid0 + sl;
    const NetDev& net = nets[region_net(io, Wd, rid)];
    walk_fp64_body(net, scan_layout(io, net.C), rid, sl, planes, eqb, Cs_hbm);
  }
}

// Checked ray-shooting walk on the FP64 planes of k_chain_tile_fp64.  The
// planes are the canonical FP64 planes, so the only uncertainty is the
// rounding of the walk's own evaluations; regions it cannot certify are
// flagged (io.flags, level-relative) for the candidate scan.
template <int NW>
__device__ __forceinline__ void walk_ray64_body(const NetDev& net, const RegionIO& io, int rid, int b,
                                                const WalkParams& wp, const double* planes, const double* eqb,
                                                int Cs_hbm) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // LDS holds the planes past hidden layer 1 (index g - n1); layer 1's are
  // rebuilt from W0 and b0 where read (PlaneView), which leaves LDS for more
  // regions per CU
  const int n1 = net.nodes[1], Cl = net.C - n1;
  double2* Q = reinterpret_cast<double2*>(smem_raw);   // (x, y) pairs, then (z, c) pairs
  uint32_t* smask = reinterpret_cast<uint32_t*>(Q + 2 * Cl);
  const WalkRecs rec = carve_walk_recs(reinterpret_cast<unsigned char*>(Q + 2 * Cl) + ((net.Wd * 4 + 15) & ~15),
                                       net.vm, false);
  const int lane = threadIdx.x;
  const int start_in = io.reg_start[rid];
  for (int w = lane; w < net.Wd; w += 64 * NW) smask[w] = io.reg_mask[(size_t)rid * net.Wd + w];
  const double* src = planes + (size_t)b * 4 * Cs_hbm + n1;
  for (int i = lane; i < Cl; i += 64 * NW) {
    Q[i] = double2{src[i], src[Cs_hbm + i]};
    Q[Cl + i] = double2{src[2 * (size_t)Cs_hbm + i], src[3 * (size_t)Cs_hbm + i]};
  }
  double p[3];
  p[0] = io.reg_point[(size_t)rid * 3 + 0];
  p[1] = io.reg_point[(size_t)rid * 3 + 1];
  p[2] = io.reg_point[(size_t)rid * 3 + 2];
  const double eq[4] = {eqb[(size_t)b * 4 + 0], eqb[(size_t)b * 4 + 1], eqb[(size_t)b * 4 + 2],
                        eqb[(size_t)b * 4 + 3]};
  __syncthreads();
  const PlaneView<double> P{reinterpret_cast<const double*>(Q), Cl, n1, net.W0, net.b[0], smask};
  walk_ray<double, NW>(net, io, wp, rid, rid - io.first, P, smask, eq, p, rec, start_in);
}

// NW waves per region (the records do not depend on NW: every reduction breaks
// ties by constraint index and each serial step runs on one wave).
// With `left` (the counter zeroed before the launch), a region the walk
// leaves undecided puts its chunk-local slot on the candidate scan's list and
// counts itself and its reasons in left.run (as k_collect_slots would).
template <int NW>
__global__ __launch_bounds__(64 * NW, 24 / NW) void k_walk_ray64(NetDev net, RegionIO io, int rid0, WalkParams wp,
                                                        const double* planes, const double* eqb, int Cs_hbm,
                                                        int only_undecided, SlotList left) {
  if ((int)blockIdx.x >= launch_count(io)) return;
  const int rid = io.ids ? io.ids[blockIdx.x] : rid0 + blockIdx.x;
  // a region an earlier walk decided keeps its records (its flag is 0 and stays 0)
  if (only_undecided && !(io.flags[rid - io.first] & AM_UNDECIDED)) return;
  walk_ray64_body<NW>(net, io, rid, blockIdx.x, wp, planes, eqb, Cs_hbm);
  // thread 0 wrote the region's flag (finish_region)
  if (left.cnt && threadIdx.x == 0) {
    const unsigned f = io.flags[rid - io.first];
    if (f & AM_UNDECIDED) {
      left.slots[atomicAdd(left.cnt, 1)] = (int)blockIdx.x;
      atomicAdd(left.run, 1);
      for (int w = 0; w < AM_WHY_COUNT; ++w)
        if (f & AM_WHY_BIT(w)) atomicAdd(left.run + 1 + w, 1);
    }
  }
}

template <int NW>
__global__ __launch_bounds__(64 * NW, 24 / NW) void k_walk_ray64_multi(const NetDev* nets, RegionIO io, int Wd, int rid0,
                                                              WalkParams wp, const double* planes, const double* eqb,
                                                              int Cs_hbm) {
  if ((int)blockIdx.x >= io.count) return;
  const int rid = rid0 + blockIdx.x;
  walk_ray64_body<NW>(nets[region_net(io, Wd, rid)], io, rid, blockIdx.x, wp, planes, eqb, Cs_hbm);
}

template __global__ void k_walk_ray64<2>(NetDev, RegionIO, int, WalkParams, const double*, const double*, int, int,
                                         SlotList);
template __global__ void k_walk_ray64<4>(NetDev, RegionIO, int, WalkParams, const double*, const double*, int, int,
                                         SlotList);
template __global__ void k_walk_ray64<8>(NetDev, RegionIO, int, WalkParams, const double*, const double*, int, int,
                                         SlotList);
template __global__ void k_walk_ray64_multi<2>(const NetDev*, RegionIO, int, int, WalkParams, const double*,
                                               const double*, int);
template __global__ void k_walk_ray64_multi<4>(const NetDev*, RegionIO, int, int, WalkParams, const double*,
                                               const double*, int);
template __global__ void k_walk_ray64_multi<8>(const NetDev*, RegionIO, int, int, WalkParams, const double*,
                                               const double*, int);

// Vertices of a region whose polygon (the tight-constraint pair of every
// vertex, its order and orientation) a reduced-precision walk certified: each
// is the intersection of the FP64 planes of its pair with the face plane,
// solved as the FP64 ray walk solves it (new constraint, constraint walked
// along, face plane), so the coordinates carry FP64 rounding only.
__global__ __launch_bounds__(64) void k_vertex_fp64(NetDev net, RegionIO io, int rid0, const double* planes,
                                                    const double* eqb, int Cs_hbm) {
  const int b = blockIdx.x;
  if (b >= launch_count(io)) return;
  const int rid = io.ids ? io.ids[b] : rid0 + b;
  if (io.flags && (io.flags[rid - io.first] & AM_UNDECIDED)) return;   // left to the FP64 walk
  const int vm = net.vm;
  const int nv = io.reg_info[(size_t)rid * AM_INFO_STRIDE + 2];
  const double* src = planes + (size_t)b * 4 * Cs_hbm;
  const uint32_t* mask = io.reg_mask + (size_t)rid * net.Wd;
  const double eq[4] = {eqb[(size_t)b * 4 + 0], eqb[(size_t)b * 4 + 1], eqb[(size_t)b * 4 + 2],
                        eqb[(size_t)b * 4 + 3]};
  for (int i = threadIdx.x; i < nv && i < vm; i += 64) {
    const int r = io.reg_vpair[((size_t)rid * vm + i) * 2 + 0];
    const int s = io.reg_vpair[((size_t)rid * vm + i) * 2 + 1];
    if (r < 0 || r >= net.C || s < 0 || s >= net.C) continue;
    double pr[4], ps[4];
    plane64(net, mask, src, Cs_hbm, r, pr);
    plane64(net, mask, src, Cs_hbm, s, ps);
    double v[3];
    solve3_rcp(pr, ps, eq, v);
    double* out = io.reg_verts + ((size_t)rid * vm + i) * 3;
    out[0] = v[0];
    out[1] = v[1];
    out[2] = v[2];
  }
}

// Chunk-local slots of the regions k_walk_ray64 left undecided: region of
// slot b is ids[b] or rid0 + b; its flag is level-relative.
__global__ void k_collect_slots(const unsigned* flags, int n, const int* dn, int dbase, const int* ids, int rid0,
                                int lvl_first, int* cnt, int* slots, int* run) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= (dn ? min(n, *dn - dbase) : n)) return;
  const int rid = ids ? ids[b] : rid0 + b;
  const unsigned f = flags[rid - lvl_first];
  if (f & AM_UNDECIDED) {
    slots[atomicAdd(cnt, 1)] = b;
    atomicAdd(run, 1);
    for (int w = 0; w < AM_WHY_COUNT; ++w)
      if (f & AM_WHY_BIT(w)) atomicAdd(run + 1 + w, 1);
  }
}

}  // namespace am
