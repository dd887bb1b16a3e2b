// Kernel interfaces of the marching engine (device side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "am_limits.h"

namespace am {

// Region store + the current level's candidate slots.
struct RegionIO {
  const uint32_t* reg_mask;   // [R][Wd]
  const int* reg_start;       // [R]   start plane (-1: seed)
  const double* reg_point;    // [R][3] seed point of the region
  int* reg_info;              // [R][AM_INFO_STRIDE]
  int* reg_tab;               // [R][vm + 1]
  double* reg_verts;          // [R][vm][3]
  int* reg_vpair;             // [R][vm][2]
  uint32_t* cand_mask;        // [count * vm][Wd]
  int* cand_start;            // [count * vm]   -2: no candidate in this slot
  double* cand_point;         // [count * vm][3]
  uint64_t* cand_hash;        // [count * vm] key hash sum of the candidate's mask (key_hash_sum)
  int first;                  // region id of blockIdx.x == 0
  int count;                  // regions in this launch
  int Cs;                     // stride of one SoA plane array in LDS
  int Csort;                  // power of two >= C (sort width)
  unsigned* flags;            // [count] per-region flags (bit 0: undecided in reduced precision)
  const int* ids;             // optional: region ids to process (FP64 re-decision); else first + blockIdx.x
  const int* dcount = nullptr;  // optional: ids past *dcount - dbase are not valid (a launch sized on
  int dbase = 0;                // the host for the worst case; blocks past the device count exit)
  uint16_t* reg_jinv = nullptr; // optional [R][vm][9] bf16: per vertex the inverse of its 3x3 system
                                // from the reduced planes (columns for the new, walked and face plane),
                                // for the FP64 Newton step (a record with it has info[7] = 1)
};

// Regions of a launch whose count may be known on the device only.
template <typename IO>
__device__ __forceinline__ int launch_count(const IO& io) {
  return io.dcount ? min(io.count, *io.dcount - io.dbase) : io.count;
}

// Visited-set insertion of one level's candidates.
struct InsertIO {
  const uint32_t* cand_mask;  // [n][Wd]
  const uint64_t* cand_hash;  // [n] key_hash_sum of each candidate's mask
  const int* cand_start;      // [n]  -2 = empty slot
  const uint32_t* reg_mask;   // [R][Wd]
  uint32_t* table;            // [cap]  0xFFFFFFFF empty, 0x80000000|cand pending, else region id
  int* slot;                  // [n]  table slot of a pending candidate, -1: empty or already visited
  int n;
  int Wd;
  uint32_t cap_mask;
  unsigned long long* scan_state;  // [scan_words] tile states of the level's one-pass scan, cleared here
  int scan_words;
};

__global__ void k_region_fp64(NetDev net, RegionIO io);

// Tiled FP64 chain: the planes of a tile of regions in HBM, for the walk kernel.
#define AM_CT 16                // regions per chain workgroup
#define AM_CN (4 * AM_CT)       // chain columns (one per lane): x, y, z, c of each region
#define AM_CXS (AM_CN + 1)      // LDS row stride (doubles) of the chain's layer buffer
#define AM_CR 16                // output rows per wave
#define AM_CHAIN_MAXW 256       // widest layer the tiled chain takes
#define AM_RAY_WAVES 4          // waves per region in the ray-shooting walks
#define AM_CPF 16               // k steps of chain operands in flight (k lists are padded to a multiple)
#define AM_KLS (AM_CHAIN_MAXW + 2 * AM_CPF)   // entries of one layer's active-input list
#define AM_CXR (AM_CHAIN_MAXW + AM_CPF)       // rows of the chain's layer buffer (list positions + prefetch)
#define AM_CHAIN_THREADS (64 * AM_CHAIN_MAXW / AM_CR)
// Half tiles of the matrix-core chain (k_chain_tile_mfma)
#define AM_HT 8                 // regions per workgroup
#define AM_HN (4 * AM_HT)       // columns
#define AM_HXS (AM_HN + 1)      // LDS row stride (doubles)
#define AM_HALF_THREADS 512     // 8 waves, 32 output rows each

struct ChainIO {
  const uint32_t* reg_mask;   // [R][Wd]
  const int* ids;             // optional region ids; else first + tile-local index
  int first, count;           // regions of this launch
  int Cs;                     // stride of one component array in `planes` (>= C)
  double* planes;             // [count][4][Cs]: signed constraint planes, extras appended; entries of
                              // hidden layer 1 are not written (readers rebuild them from W0, b0, mask)
  double* eq;                 // [count][4]: output plane with iso subtracted
  const double* Wp[AM_MAX_LINEAR];  // li >= 1: W_li^T padded, [n_li][noutp[li]] (zero rows past n_out)
  const double* bp[AM_MAX_LINEAR];  // li >= 1: b_li padded to noutp[li]
  int noutp[AM_MAX_LINEAR];         // n_out rounded up to AM_CR
  const int* dcount = nullptr;      // as in RegionIO: regions past *dcount - dbase are not valid
  int dbase = 0;
  int* zero = nullptr;              // set to 0 by the launch (a counter that later launches of the chunk fill)
};

// A list of chunk-local slots filled by atomics (cnt zeroed before), with
// counters of what was listed: run[0] the entries, run[1 + w] those with
// reason bit w.
struct SlotList {
  int* cnt = nullptr;
  int* slots = nullptr;
  int* run = nullptr;
};

// The regions and plane buffers of one tiled-chain launch.
struct ChainLevel {
  int first, count;           // regions first .. first + count - 1 (or ids[0 .. count))
  const int* ids;
  double* planes;             // [count][4][Cs]
  double* eq;                 // [count][4]
};

__global__ void k_chain_tile_fp64(NetDev net, ChainIO io);
// The same planes with the products on the FP64 matrix cores: equal up to the
// rounding of another summation order (the forms that certify their decisions).
// Tiles of AM_HT regions, AM_HALF_THREADS threads, LDS chain_half_lds_bytes.
__global__ void k_chain_tile_mfma(NetDev net, ChainIO io);
// A batch of networks (each region's network is the last word of its mask):
// tiles[b] = (network, first slot, regions, -) of one network.
// MF: the products on the matrix cores (k_chain_tile_mfma).
template <bool MF>
__global__ void k_chain_tile_fp64_multi(const NetDev* nets, const ChainIO* cios, ChainLevel lv, const int4* tiles);
// The walk of k_region_fp64 on planes from k_chain_tile_fp64.  Block b takes
// planes slot s = slots[b] (or b); its region is io.ids[s] when ids are given,
// else rid0 + s; candidate slots are numbered from io.first as in k_region_fp64.
// With `count`, blocks at or past *count exit (a grid sized for the worst case).
__global__ void k_walk_fp64(NetDev net, RegionIO io, int rid0, const int* count, const int* slots,
                            const double* planes, const double* eq, int Cs_hbm);
struct WalkParams;
// only_undecided: blocks whose region is not flagged AM_UNDECIDED in io.flags exit.
// NW waves per region: 2, 4 or 8.  With left.cnt, the regions the walk leaves
// undecided are listed as k_collect_slots lists them.
template <int NW>
__global__ void k_walk_ray64(NetDev net, RegionIO io, int rid0, WalkParams wp, const double* planes,
                             const double* eq, int Cs_hbm, int only_undecided, SlotList left);
// Batch-of-networks forms (region rid0 + b; its network is mask word Wd - 1;
// the scan's LDS layout follows each network's C, the launch is sized for the
// widest network).
__global__ void k_walk_fp64_multi(const NetDev* nets, RegionIO io, int Wd, int rid0, const int* count,
                                  const int* slots, const double* planes, const double* eq, int Cs_hbm);
template <int NW>
__global__ void k_walk_ray64_multi(const NetDev* nets, RegionIO io, int Wd, int rid0, WalkParams wp,
                                   const double* planes, const double* eq, int Cs_hbm);
// Chunk-local slots of the regions k_walk_ray64 left undecided (cnt[0] counts
// them); run[0] accumulates the count and run[1 + w] the reasons, per run.
// With dn, entries past *dn - dbase are not valid.
__global__ void k_collect_slots(const unsigned* flags, int n, const int* dn, int dbase, const int* ids, int rid0,
                                int lvl_first, int* cnt, int* slots, int* run);
// FP64 vertices of regions whose topology a reduced-precision walk decided:
// block b re-solves the face vertices of region io.ids[b] (or rid0 + b) from
// the FP64 planes of slot b (its two tight constraints and the face plane);
// with io.flags, regions flagged AM_UNDECIDED are skipped.
__global__ void k_vertex_fp64(NetDev net, RegionIO io, int rid0, const double* planes, const double* eq,
                              int Cs_hbm);
__global__ void k_fill_i32(int* p, size_t n, int v);

// Weight layouts on the device (am_layout.hip): dst[j * ni + i] = src[i * cols + j]
// for i < rows, j < cols, +0 elsewhere (i < ni, j < nj); src float or double
// (f64), device-readable.  A copy with padding is cols = nj = 1.
void launch_transpose_pad(const void* src, int f64, int rows, int cols, double* dst, int ni, int nj, hipStream_t st);
void launch_transpose_pad(const void* src, int f64, int rows, int cols, float* dst, int ni, int nj, hipStream_t st);

}  // namespace am
