// Visited set, frontier compaction and vertex deduplication.
//
// Visited set (replaces backend/inc/hash.h:33-67 and backend/src/states.cu:30-79):
// an open-addressing table in HBM whose slots hold either a region id or,
// while a level is being inserted, "pending | candidate index".  Keys are the
// full packed masks, compared word by word against immutable copies (the
// region store or the level's candidate buffer), so membership is exact.
// Among equal candidates of one level the lowest candidate index wins, which
// makes the next frontier's order independent of atomic timing.
//
// Vertex deduplication lives in am_combine.hip.
#include "am_device.h"
#include "am_kernels.h"
#include "am_hash.h"

namespace am {


// All loads issued before any comparison decides (no early exit), so a
// thread's key reads overlap instead of waiting on each other.
__device__ __forceinline__ bool words_equal(const uint32_t* a, const uint32_t* b, int n) {
  uint32_t diff = 0;
  for (int w = 0; w < n; ++w) diff |= a[w] ^ b[w];
  return diff == 0;
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_i32(int* p, size_t n, int v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Candidate c claims a slot (CAS) or joins the pending group of an equal key
// there, lowering the slot to the group's smallest candidate (atomicMin on
// pending | c); a key already in the store, or an empty candidate, gets slot -1.
__global__ void k_insert(InsertIO io) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= io.n) return;
  if (io.cand_start[c] == -2) { io.slot[c] = -1; return; }
  const uint32_t* key = io.cand_mask + (size_t)c * io.Wd;
  uint32_t s = (uint32_t)mix64(io.cand_hash[c]) & io.cap_mask;
  for (uint32_t probes = 0; probes <= io.cap_mask; ++probes) {
    const uint32_t prev = atomicCAS(&io.table[s], AM_EMPTY, AM_PEND | (uint32_t)c);
    if (prev == AM_EMPTY) {
      io.slot[c] = (int)s;
      return;
    }
    if (prev & AM_PEND) {
      const int o = (int)(prev & ~AM_PEND);
      if (words_equal(key, io.cand_mask + (size_t)o * io.Wd, io.Wd)) {
        atomicMin(&io.table[s], AM_PEND | (uint32_t)c);
        io.slot[c] = (int)s;
        return;
      }
    } else if (words_equal(key, io.reg_mask + (size_t)prev * io.Wd, io.Wd)) {
      io.slot[c] = -1;
      return;
    }
    s = (s + 1) & io.cap_mask;
  }
  io.slot[c] = -1;  // table full: the host sizes the table so this never happens
}

// New regions get ids R0 + rank in candidate order (rank from a stored scan).
__global__ void k_commit(CommitIO io, const int* rank, int n) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n && first_flag(io, c)) commit_one(io, c, rank[c]);
}

// Key hash sums of n candidate masks given whole (the seeds).
__global__ void k_hash_keys(const uint32_t* masks, int n, int Wd, uint64_t* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) out[c] = key_hash_sum(masks + (size_t)c * Wd, Wd);
}

// Re-insert every region id into an empty (larger) table.
__global__ void k_rehash(const uint32_t* reg_mask, int R, int Wd, uint32_t* table, uint32_t cap_mask) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  uint32_t s = (uint32_t)mix64(key_hash_sum(reg_mask + (size_t)r * Wd, Wd)) & cap_mask;
  for (uint32_t probes = 0; probes <= cap_mask; ++probes) {
    if (atomicCAS(&table[s], AM_EMPTY, (uint32_t)r) == AM_EMPTY) return;
    s = (s + 1) & cap_mask;
  }
}

}  // namespace am

namespace am {

// Topology fingerprint of the region store: order-independent sums (mod 2^64)
// of a per-row hash over three kinds of int64 rows -- the visited masks, the
// canonical faces (mask words, vertex count, the vertex pair keys lo<<32|hi
// rotated to start at the smallest, -1 past the count) and one vertex-key row
// per vertex use (lo, hi, mask without bits lo and hi).  The same rows and hash
// as analyticmesh_amd/canonical.py, so a GPU run can be compared with the
// oracle's fingerprint at full size without copying its records to the host.
struct RowHash {
  uint64_t h = 0x243F6A8885A308D3ULL;
  __device__ __forceinline__ void col(int64_t v) {
    h = (h ^ (uint64_t)v) * 0x9E3779B97F4A7C15ULL;
    h ^= h >> 31;
  }
  __device__ __forceinline__ uint64_t done() const {
    uint64_t x = (h ^ (h >> 33)) * 0xFF51AFD7ED558CCDULL;
    return x ^ (x >> 33);
  }
};

// out[0..2]: visited, faces, vertex-key sums; out[3], out[4]: faces, vertex uses;
// out[5]: regions.  Rows hash the first Wh of a mask's Wd words; with net >= 0
// only regions whose last mask word equals net count (a batch of networks).
__global__ __launch_bounds__(256) void k_fingerprint(const uint32_t* reg_mask, const int* info, const int* vpair,
                                                     int R, int Wd, int Wh, int vm, int L, int net,
                                                     unsigned long long* out) {
  __shared__ unsigned long long red[6][256];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t acc[6] = {0, 0, 0, 0, 0, 0};
  if (r < R && (net < 0 || reg_mask[(size_t)r * Wd + Wd - 1] == (uint32_t)net)) {
    const uint32_t* m = reg_mask + (size_t)r * Wd;
    acc[5] = 1;
    RowHash hv;
    for (int w = 0; w < Wh; ++w) hv.col((int64_t)m[w]);
    acc[0] = hv.done();
    const int nv = info[(size_t)r * AM_INFO_STRIDE + 2];
    if (nv >= 3) {
      const int* vp = vpair + (size_t)r * vm * 2;
      int64_t best = INT64_MAX;
      int i0 = 0;
      for (int k = 0; k < nv; ++k) {
        const int a = vp[2 * k], b = vp[2 * k + 1];
        const int64_t key = ((int64_t)min(a, b) << 32) | (int64_t)max(a, b);
        if (key < best) { best = key; i0 = k; }
      }
      RowHash hf;
      for (int w = 0; w < Wh; ++w) hf.col((int64_t)m[w]);
      hf.col(nv);
      int k = i0;
      for (int s = 0; s < vm; ++s) {
        if (s < nv) {
          const int a = vp[2 * k], b = vp[2 * k + 1];
          hf.col(((int64_t)min(a, b) << 32) | (int64_t)max(a, b));
          k = k + 1 == nv ? 0 : k + 1;
        } else {
          hf.col(-1);
        }
      }
      acc[1] = hf.done();
      acc[3] = 1;
      for (int q = 0; q < nv; ++q) {
        const int a = vp[2 * q], b = vp[2 * q + 1];
        const int lo = min(a, b), hi = max(a, b);
        RowHash hk;
        hk.col(lo);
        hk.col(hi);
        for (int w = 0; w < Wh; ++w) {
          uint32_t x = m[w];
          if (lo >= 0 && lo < L && (lo >> 5) == w) x &= ~(1u << (lo & 31));
          if (hi >= 0 && hi < L && (hi >> 5) == w) x &= ~(1u << (hi & 31));
          hk.col((int64_t)x);
        }
        acc[2] += hk.done();
      }
      acc[4] = (uint64_t)nv;
    }
  }
  for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 6) atomicAdd(&out[threadIdx.x], red[threadIdx.x][0]);
}

// Regions per network of level regions first .. first + count - 1 (network =
// last mask word).
__global__ void k_net_hist(const uint32_t* reg_mask, int Wd, int first, int count, int* hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) atomicAdd(&hist[reg_mask[(size_t)(first + i) * Wd + Wd - 1]], 1);
}

// Ids of the regions of network `net` (last mask word), in region order:
// flags for the scan, then the scatter.
__global__ void k_net_flags(const uint32_t* reg_mask, int Wd, int R, int net, int* flags) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R) flags[r] = reg_mask[(size_t)r * Wd + Wd - 1] == (uint32_t)net ? 1 : 0;
}
__global__ void k_net_scatter(const int* flags, const int* rank, int R, int* ids) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R && flags[r]) ids[rank[r]] = r;
}

}  // namespace am
