// Reduced-precision plane path: the masked affine chain of a tile of regions
// as real GEMMs on MFMA (FP32, or split-BF16 with three products), a checked
// FP32 walk, and a list of the regions whose predicates it could not decide
// (the engine re-runs those through the FP64 kernel).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "am_kernels.h"
#include "am_timer.h"
#include "am_walk.h"

namespace am {

#define AM_GT 8                 // regions per GEMM workgroup
#define AM_PCOMP 5              // plane record components in HBM: x, y, z, c, sigma
#define AM_EQS 8                // floats per region of the face-plane record (x, y, z, c, sigma, pad)
#define AM_GN (4 * AM_GT)       // GEMM columns: x, y, z, c of each region

struct FastNetDev {
  int D, L, E, C, Wd;
  int Cs;                       // stride of one plane component array in HBM (C rounded up to 64)
  int Kpad;                     // LDS row count (largest width rounded up to 32)
  int nodes[AM_MAX_LINEAR + 1];
  int off[AM_MAX_LINEAR + 1];
  int nrb[AM_MAX_LINEAR];       // 16-row blocks of linear layer li
  int nkb[AM_MAX_LINEAR];       // 4-wide (fp32) or 32-wide (bf16) k blocks of layer li
  const float* W0;              // (n1, 3)
  const float* bias[AM_MAX_LINEAR];
  const float* Wp[AM_MAX_LINEAR];       // fp32 W^T, k-major [n_in + 1][nrb * 16], row n_in = +0
  const uint16_t* Wh[AM_MAX_LINEAR];    // bf16 hi A operands, [nrb][nkb][64 lanes][8]
  const uint16_t* Wl[AM_MAX_LINEAR];    // bf16 lo A operands
  // Error estimates of the planes: per region and plane j of layer li + 1 the
  // variance V_j = sum_k W_jk^2 m_k (V_k + ck[li] |P_k|^2) + ck[li] b_j^2 (V_k
  // of the layer's input planes, 0 for hidden layer 1's exact planes), a bf16
  // MFMA product inside the FP32 k loop (A: the squared weight fragments).
  float ck[AM_MAX_LINEAR];      // variance per squared term of layer li's sums
  int sig;                      // 1: compute the estimates (planes component 4, eq[4])
  const float* ew;
  const float* eb;
  float iso;
};

// Forms: 0 FP32 MFMA, 1 split-BF16 MFMA.
struct FastNet {
  int precision = 0;
  int walk_mode = 0;            // 0: ray-shooting walk, 1: the reference's candidate scan (checked)
  int ray_waves = 4;            // waves per region of the ray-shooting walk (2, 4, 8)
  int vertex_mode = 3;          // FP32 form: 0 every vertex re-solved from FP64 planes; 1 vertices solved
                                // from the reduced planes kept where their error estimate is within 1e-5;
                                // 2 vertices solved from the reduced planes, then one FP64 Newton step
                                // after the march (the walk keeps their inverse systems); 3 (default)
                                // as 2, the step only for vertices whose estimate is not within 1e-5
  bool ready = false;
  FastNetDev dev{};
  WalkParams wp{};
};

struct FastBuffers {
  void* mem[64] = {nullptr};
  size_t bytes[64] = {0};
};

struct FastLevel {
  int first, count;
  const uint32_t* reg_mask;
  const int* reg_start;
  const double* reg_point;
  int* reg_info;
  int* reg_tab;
  double* reg_verts;
  int* reg_vpair;
  uint32_t* cand_mask;
  int* cand_start;
  double* cand_point;
  uint64_t* cand_hash;
  uint16_t* reg_jinv = nullptr;   // optional: the walk's inverse vertex systems (Newton vertices)
};

struct FastTimes {
  int launches;
};

// Weights in MFMA operand order; `weights`/`biases` are device-readable
// (float or double as `dtype`), the layouts are built on stream st.
void fast_prepare(FastNet& f, FastBuffers& b, const NetDev& net, int precision, const void* const* weights,
                  const void* const* biases, int dtype, hipStream_t st);
// The regions of a level that a reduced walk left to FP64: ids[0, *count)
// (device), and the tile state of the one-pass scan that lists them
// (redo_state_words(level size) words, zero before the first use; every
// launch leaves it zero again).
struct RedoList {
  int* ids;
  int* count;
  unsigned long long* state;
};
size_t redo_state_words(int count);

// Planes + checked walk for one level; the walk marks each region in
// flags[0, lvl.count) (AM_UNDECIDED plus reason bits, AM_VERT64, or 0 when it
// wrote the region's records).  The last launch lists the regions marked
// undecided or AM_VERT64 into `redo` (in region order) and adds the undecided
// ones to acc[0], their reasons to acc[1 + w] and the AM_VERT64 ones to
// acc[1 + AM_WHY_COUNT] on the device; nothing here waits.
// Kernel times go to `timer` (T_GEMM, T_WALK).
// With `prepared`, fast_level_prepare(f, b, net, lvl.count) has run before
// (then fast_level only enqueues launches: it can be captured into a graph).
void fast_level(FastNet& f, FastBuffers& b, const NetDev& net, const FastLevel& lvl, hipStream_t st,
                StageTimer& timer, FastTimes& t, unsigned* flags, int* acc, const RedoList& redo,
                bool prepared = false);
// Buffers and launch attributes of fast_level for a level of `count` regions.
void fast_level_prepare(FastNet& f, FastBuffers& b, const NetDev& net, int count);
// Planes of n given regions (their masks in `reg_mask`) with the prepared
// reduced-precision form: planes [n][AM_PCOMP][Cs] f32 (x, y, z, c signed, then
// each plane's error estimate sigma), eq [n][AM_EQS] f32 (x, y, z, c, sigma).
void fast_planes(FastNet& f, const uint32_t* reg_mask, int n, float* planes, float* eq, hipStream_t st);
void fast_release(FastBuffers& b);

}  // namespace am
