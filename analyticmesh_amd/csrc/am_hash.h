// Visited-set and compaction kernels, and the topology fingerprint (device side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "am_kernels.h"

namespace am {

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v);
__global__ void k_fill_i32(int* p, size_t n, int v);
__global__ void k_insert(InsertIO io);

// Commit of a level's new regions: the candidates, their slots, the store.
struct CommitIO {
  const int* slot;
  const uint32_t* cand_mask;
  const int* cand_start;
  const double* cand_point;
  uint32_t* reg_mask;
  int* reg_start;
  double* reg_point;
  uint32_t* table;
  int Wd, R0;
};
#define AM_EMPTY 0xFFFFFFFFu
#define AM_PEND 0x80000000u

// A candidate is new (becomes a region) when it is the smallest of its pending
// group: its slot still holds pending | c.  Only that candidate's own commit
// rewrites the slot, so the flag reads the same before and during the commit.
__device__ __forceinline__ int first_flag(const CommitIO& io, size_t c) {
  const int s = io.slot[c];
  return s >= 0 && io.table[s] == (AM_PEND | (uint32_t)c) ? 1 : 0;
}

__device__ __forceinline__ void commit_one(const CommitIO& io, size_t c, int rank) {
  const int rid = io.R0 + rank;
  for (int w = 0; w < io.Wd; ++w) io.reg_mask[(size_t)rid * io.Wd + w] = io.cand_mask[c * io.Wd + w];
  io.reg_start[rid] = io.cand_start[c];
  io.reg_point[(size_t)rid * 3 + 0] = io.cand_point[c * 3 + 0];
  io.reg_point[(size_t)rid * 3 + 1] = io.cand_point[c * 3 + 1];
  io.reg_point[(size_t)rid * 3 + 2] = io.cand_point[c * 3 + 2];
  io.table[io.slot[c]] = (uint32_t)rid;
}

// scan source: 1 for a candidate that becomes a region
struct FirstSrc {
  CommitIO io;
  __device__ int operator()(size_t c) const { return first_flag(io, c); }
};
// scan sink that commits each new region as its rank is known
struct CommitSink {
  CommitIO io;
  __device__ void operator()(size_t c, int prefix, int v) const {
    if (v) commit_one(io, c, prefix);
  }
};
__global__ void k_commit(CommitIO io, const int* rank, int n);

// One-pass scan of the new-region flags of a level and the commit of the new
// regions (k_scan_commit).  state: [tiles] tile states (flag << 32 | count; flag
// 1: the tile's own count, 2: its inclusive prefix) and [tiles] the tile
// counter, all zero at launch (k_insert clears them).  With rank set the
// kernel only writes the ranks of the new regions (the store grows before
// k_commit); total (host-mapped) receives the level's count.
struct ScanCommitIO {
  CommitIO cio;
  unsigned long long* state;
  int n, tiles;
  int* rank;
  int* total;
};
constexpr int kScanCommitThreads = 512;
constexpr int kScanCommitItems = 8;
constexpr int kScanCommitTile = kScanCommitThreads * kScanCommitItems;   // candidates per block
inline size_t scan_commit_tiles(size_t n) { return (n + kScanCommitTile - 1) / kScanCommitTile; }
__global__ void k_scan_commit(ScanCommitIO io);
__global__ void k_hash_keys(const uint32_t* masks, int n, int Wd, uint64_t* out);
__global__ void k_rehash(const uint32_t* reg_mask, int R, int Wd, uint32_t* table, uint32_t cap_mask);
__global__ void k_fingerprint(const uint32_t* reg_mask, const int* info, const int* vpair, int R, int Wd, int Wh,
                              int vm, int L, int net, unsigned long long* out);
__global__ void k_net_hist(const uint32_t* reg_mask, int Wd, int first, int count, int* hist);
__global__ void k_net_flags(const uint32_t* reg_mask, int Wd, int R, int net, int* flags);
__global__ void k_net_scatter(const int* flags, const int* rank, int R, int* ids);

}  // namespace am
