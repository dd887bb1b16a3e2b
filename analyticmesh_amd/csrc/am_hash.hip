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
#include "am_scan.h"

namespace am {


// Keys are handled by groups of AM_KG lanes: lane j of a group takes words
// j, j + AM_KG, ... of a key (16-byte quads when the key is a whole number of
// quads), so one load instruction of a wave reads AM_KG whole keys' segments,
// coalesced.
#define AM_KG 4

// Whether keys a and b (Wd words) are equal, in every lane of the group; a
// group that does not compare (cmp false, the same in all its lanes) takes
// part with no loads.  Called by the whole wave.
__device__ __forceinline__ bool group_equal(const uint32_t* a, const uint32_t* b, int Wd, int j, int g, bool cmp) {
  uint32_t diff = 0;
  if (cmp) {
    if ((Wd & 3) == 0) {
      const uint4* a4 = reinterpret_cast<const uint4*>(a);
      const uint4* b4 = reinterpret_cast<const uint4*>(b);
      for (int q = j; q < (Wd >> 2); q += AM_KG) {
        const uint4 x = a4[q], y = b4[q];
        diff |= (x.x ^ y.x) | (x.y ^ y.y) | (x.z ^ y.z) | (x.w ^ y.w);
      }
    } else {
      for (int w = j; w < Wd; w += AM_KG) diff |= a[w] ^ b[w];
    }
  }
  const unsigned long long bal = __ballot(diff != 0);
  return ((bal >> (g * AM_KG)) & ((1ull << AM_KG) - 1)) == 0;
}

// Copy of a key (Wd words) by a group, as group_equal reads it.
__device__ __forceinline__ void group_copy(uint32_t* dst, const uint32_t* src, int Wd, int j) {
  if ((Wd & 3) == 0) {
    for (int q = j; q < (Wd >> 2); q += AM_KG) reinterpret_cast<uint4*>(dst)[q] = reinterpret_cast<const uint4*>(src)[q];
  } else {
    for (int w = j; w < Wd; w += AM_KG) dst[w] = src[w];
  }
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
// A wave takes 64 consecutive candidate slots: empty ones (no face edge) are
// settled by their own lane, the others are compacted and probed 64 / AM_KG at
// a time, one group of lanes per candidate (the group's first lane does the
// atomics, the group compares keys).  Also clears the tile states of the scan
// that follows on the stream.
__global__ __launch_bounds__(256) void k_insert(InsertIO io) {
  __shared__ int s_list[4][64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int gtid = blockIdx.x * 256 + tid;
  if (gtid < io.scan_words) io.scan_state[gtid] = 0ull;
  const int c = gtid;
  bool live = false;
  if (c < io.n) {
    live = io.cand_start[c] != -2;
    if (!live) io.slot[c] = -1;
  }
  const unsigned long long bal = __ballot(live);
  const int nlive = __popcll(bal);
  if (live) s_list[wv][__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u))] = c;
  __syncthreads();
  const int j = lane & (AM_KG - 1), g = lane / AM_KG, lead = lane & ~(AM_KG - 1);
  for (int p = 0; p < nlive; p += 64 / AM_KG) {
    const bool act = p + g < nlive;
    const int cc = act ? s_list[wv][p + g] : 0;
    const uint32_t* key = io.cand_mask + (size_t)cc * io.Wd;
    uint32_t s = act ? (uint32_t)mix64(io.cand_hash[cc]) & io.cap_mask : 0u;
    bool done = !act;
    int result = -1;
    for (uint32_t probes = 0; __ballot(!done) != 0ull; ++probes) {
      uint32_t prev = 0u;
      if (!done && j == 0) {
        // read the slot first: most candidates meet a key already there, and
        // a slot only ever goes empty -> pending c -> pending c' < c within
        // the kernel, so a non-empty value read here is the key's slot state
        prev = __hip_atomic_load(&io.table[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == AM_EMPTY) prev = atomicCAS(&io.table[s], AM_EMPTY, AM_PEND | (uint32_t)cc);
      }
      prev = (uint32_t)__shfl((int)prev, lead, 64);
      bool cmp = false;
      const uint32_t* other = key;
      if (!done) {
        if (prev == AM_EMPTY) {
          result = (int)s;
          done = true;
        } else {
          cmp = true;
          other = (prev & AM_PEND) ? io.cand_mask + (size_t)(prev & ~AM_PEND) * io.Wd
                                   : io.reg_mask + (size_t)prev * io.Wd;
        }
      }
      const bool eq = group_equal(key, other, io.Wd, j, g, cmp);
      if (cmp) {
        if (eq) {
          if (prev & AM_PEND) {
            // lower the group's smallest candidate (only a smaller index can)
            if (j == 0 && (uint32_t)cc < (prev & ~AM_PEND)) atomicMin(&io.table[s], AM_PEND | (uint32_t)cc);
            result = (int)s;
          }
          done = true;
        } else if (probes >= io.cap_mask) {
          done = true;   // table full: the host sizes the table so this never happens
        } else {
          s = (s + 1) & io.cap_mask;
        }
      }
    }
    if (act && j == 0) io.slot[cc] = result;
  }
}

// Decoupled look-back scan (Merrill and Garland, "Single-pass parallel prefix
// scan with decoupled look-back", 2016) over the new-region flags, with the
// commit: a block takes the next tile of kScanCommitTile candidates in start order
// (a counter), publishes the tile's count, then its first wave adds up the
// states of the tiles before it, 64 at a time, back to the nearest inclusive
// prefix, and publishes its own.  A tile publishes its count before it waits
// and waits only on tiles that started before it, so every wait ends.  The new
// regions of the tile are then written by groups of AM_KG lanes (mask, start,
// point, and the table slot now holding the region id).
__global__ __launch_bounds__(kScanCommitThreads) void k_scan_commit(ScanCommitIO io) {
  __shared__ int s_tile, s_excl;
  __shared__ int s_wave[kScanCommitThreads / 64];
  __shared__ int s_list[kScanCommitTile];
  const int tid = threadIdx.x, lane = tid & 63;
  if (tid == 0) s_tile = (int)atomicAdd(&io.state[io.tiles], 1ull);
  __syncthreads();
  const int t = s_tile;
  const size_t base = (size_t)t * kScanCommitTile + (size_t)tid * kScanCommitItems;
  int v[kScanCommitItems];
  int sum = 0;
#pragma unroll
  for (int it = 0; it < kScanCommitItems; ++it) {
    v[it] = base + it < (size_t)io.n ? first_flag(io.cio, base + it) : 0;
    sum += v[it];
  }
  int tot;
  const int off = block_excl_scan<int, kScanCommitThreads>(sum, s_wave, &tot);
  if (tid < 64) {
    constexpr unsigned long long F1 = 1ull << 32, F2 = 2ull << 32;
    if (lane == 0)
      __hip_atomic_store(&io.state[t], (t == 0 ? F2 : F1) | (unsigned)tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int excl = 0;
    if (t > 0) {
      for (int look = t - 1;; look -= 64) {
        const int idx = look - lane;   // lane 0: the nearest tile
        unsigned long long st = idx >= 0 ? __hip_atomic_load(&io.state[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                         : F2;
        while (__ballot((st >> 32) == 0) != 0ull) {
          if ((st >> 32) == 0) {
            __builtin_amdgcn_s_sleep(1);
            st = __hip_atomic_load(&io.state[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
        const unsigned long long b2 = __ballot((st >> 32) == 2);
        const int stop = b2 ? __ffsll((long long)b2) - 1 : 63;
        int x = lane <= stop ? (int)(unsigned)st : 0;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
        excl += x;
        if (b2) break;
      }
      if (lane == 0)
        __hip_atomic_store(&io.state[t], F2 | (unsigned)(excl + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) {
      s_excl = excl;
      if (t == io.tiles - 1 && io.total) *io.total = excl + tot;
    }
  }
  __syncthreads();
  const int excl = s_excl;
  int run = off;
#pragma unroll
  for (int it = 0; it < kScanCommitItems; ++it) {
    if (v[it]) {
      if (io.rank) io.rank[base + it] = excl + run;
      else s_list[run] = (int)(base + it);
    }
    run += v[it];
  }
  if (io.rank) return;
  __syncthreads();
  const CommitIO& ci = io.cio;
  const int j = tid & (AM_KG - 1), gid = tid / AM_KG;
  constexpr int NG = kScanCommitThreads / AM_KG;   // groups of the block
  if ((ci.Wd & 3) == 0) {
    // U regions per group at a time, all their loads before any store
    constexpr int U = 4, NQ = AM_MAX_WORDS / 4 / AM_KG;
    const int nq = ci.Wd >> 2;
    const uint4* src4 = reinterpret_cast<const uint4*>(ci.cand_mask);
    uint4* dst4 = reinterpret_cast<uint4*>(ci.reg_mask);
    for (int k0 = gid; k0 < tot; k0 += NG * U) {
      uint4 mv[U][NQ];
      int cc[U], sv[U], sl[U];
      double pv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u * NG;
        cc[u] = k < tot ? s_list[k] : -1;
#pragma unroll
        for (int t = 0; t < NQ; ++t) mv[u][t] = uint4{0u, 0u, 0u, 0u};
        sv[u] = 0;
        sl[u] = 0;
        pv[u] = 0.0;
        if (cc[u] >= 0) {
#pragma unroll
          for (int t = 0; t < NQ; ++t) {
            const int q = j + t * AM_KG;
            mv[u][t] = q < nq ? src4[(size_t)cc[u] * nq + q] : uint4{0u, 0u, 0u, 0u};
          }
          if (j == 0) {
            sv[u] = ci.cand_start[cc[u]];
            sl[u] = ci.slot[cc[u]];
          } else if (j <= 3) {
            pv[u] = ci.cand_point[(size_t)cc[u] * 3 + j - 1];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (cc[u] >= 0) {
          const int rid = ci.R0 + excl + k0 + u * NG;
#pragma unroll
          for (int t = 0; t < NQ; ++t) {
            const int q = j + t * AM_KG;
            if (q < nq) dst4[(size_t)rid * nq + q] = mv[u][t];
          }
          if (j == 0) {
            ci.reg_start[rid] = sv[u];
            ci.table[sl[u]] = (uint32_t)rid;
          } else if (j <= 3) {
            ci.reg_point[(size_t)rid * 3 + j - 1] = pv[u];
          }
        }
      }
    }
  } else {
    for (int k = gid; k < tot; k += NG) {
      const int cc = s_list[k];
      const int rid = ci.R0 + excl + k;
      group_copy(ci.reg_mask + (size_t)rid * ci.Wd, ci.cand_mask + (size_t)cc * ci.Wd, ci.Wd, j);
      if (j == 0) {
        ci.reg_start[rid] = ci.cand_start[cc];
        ci.table[ci.slot[cc]] = (uint32_t)rid;
      } else if (j <= 3) {
        ci.reg_point[(size_t)rid * 3 + j - 1] = ci.cand_point[(size_t)cc * 3 + j - 1];
      }
    }
  }
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
