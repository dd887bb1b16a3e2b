// Exclusive prefix sums over device arrays (hand-written; no library scan).
//
// Three launches: each 256-thread block reduces a tile of 2048 elements into
// a partial; one block scans the partials (in tiles, carrying the running sum)
// and writes the grand total; each block then rescans its tile from its
// partial.  Within a block a thread owns 8 consecutive elements; the thread
// totals are scanned by a wave64 shuffle scan and the four wave totals through
// LDS.  Elements come from a source functor, so the scanned values need not be
// stored (flags derived from records, packed pairs of counts).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace am {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <typename T>
__device__ __forceinline__ T shfl_up_any(T v, int d) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)u, d, 64);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(u >> 32), d, 64);
    return (T)(((uint64_t)hi << 32) | lo);
  } else {
    return (T)__shfl_up((int)v, d, 64);
  }
}

// Exclusive scan of v over the block of NT threads (s_wave: NT / 64 slots);
// *total receives the block's sum.
template <typename T, int NT = kScanThreads>
__device__ __forceinline__ T block_excl_scan(T v, T* s_wave, T* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  T inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T o = shfl_up_any(inc, d);
    if (lane >= d) inc += o;
  }
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  T off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    const T x = s_wave[w];
    if (w < wid) off += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

template <typename T, typename Src>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(Src src, size_t n, T* partials) {
  __shared__ T s_wave[kScanThreads / 64];
  const size_t base = (size_t)blockIdx.x * kScanTile;
  T sum = 0;
#pragma unroll
  for (int it = 0; it < kScanItems; ++it) {
    const size_t i = base + (size_t)it * kScanThreads + threadIdx.x;
    if (i < n) sum += src(i);
  }
  T tot;
  (void)block_excl_scan(sum, s_wave, &tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// One block: exclusive scan of nb partials in place; the total goes to *total
// (device or host-mapped memory) when total is not null.
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(T* partials, size_t nb, T* total) {
  __shared__ T s_wave[kScanThreads / 64];
  T carry = 0;
  for (size_t base = 0; base < nb; base += kScanTile) {
    T v[kScanItems];
    T sum = 0;
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
      const size_t i = base + (size_t)threadIdx.x * kScanItems + it;
      v[it] = i < nb ? partials[i] : T(0);
      sum += v[it];
    }
    T tot;
    T run = carry + block_excl_scan(sum, s_wave, &tot);
#pragma unroll
    for (int it = 0; it < kScanItems; ++it) {
      const size_t i = base + (size_t)threadIdx.x * kScanItems + it;
      if (i < nb) partials[i] = run;
      run += v[it];
    }
    carry += tot;
  }
  if (total && threadIdx.x == 0) *total = carry;
}

// out[i] = partials[tile] + sum of src over the tile's elements before i.  The
// sink functor receives (i, exclusive prefix, value) for every element.
template <typename T, typename Src, typename Sink>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(Src src, size_t n, const T* partials, Sink sink) {
  __shared__ T s_wave[kScanThreads / 64];
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  T v[kScanItems];
  T sum = 0;
#pragma unroll
  for (int it = 0; it < kScanItems; ++it) {
    v[it] = base + it < n ? src(base + it) : T(0);
    sum += v[it];
  }
  T tot;
  T run = partials[blockIdx.x] + block_excl_scan(sum, s_wave, &tot);
#pragma unroll
  for (int it = 0; it < kScanItems; ++it) {
    if (base + it < n) sink(base + it, run, v[it]);
    run += v[it];
  }
}

// Decoupled look-back (Merrill and Garland, 2016) of tile t whose own total
// is tot, run by one whole wave (the block's first); returns the sum of the
// tiles before t.  state[0, tiles) hold one word per tile, flag << 32 | value
// (flag 1: the tile's own total, 2: its inclusive prefix); state[tiles] is
// the counter the blocks took their tile numbers from (in launch order, so a
// tile only waits on tiles already running) and state[tiles + 1] counts the
// tiles done looking back.  The last of those sets all tiles + 2 words back
// to zero: the next launch on the same state needs no clearing.  *total (when
// not null) receives the sum of every tile.  Values stay below 2^31.
__device__ __forceinline__ int lookback_1pass(unsigned long long* state, int tiles, int t, int tot, int* total) {
  constexpr unsigned long long F1 = 1ull << 32, F2 = 2ull << 32;
  const int lane = threadIdx.x & 63;
  if (lane == 0)
    __hip_atomic_store(&state[t], (t == 0 ? F2 : F1) | (unsigned)tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int excl = 0;
  if (t > 0) {
    for (int look = t - 1;; look -= 64) {
      const int idx = look - lane;   // lane 0: the nearest tile
      unsigned long long st =
          idx >= 0 ? __hip_atomic_load(&state[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : F2;
      while (__ballot((st >> 32) == 0) != 0ull) {
        if ((st >> 32) == 0) {
          __builtin_amdgcn_s_sleep(1);
          st = __hip_atomic_load(&state[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
      __hip_atomic_store(&state[t], F2 | (unsigned)(excl + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (lane == 0 && t == tiles - 1 && total) *total = excl + tot;
  // this tile reads no more state: the last one done clears it
  int done = 0;
  if (lane == 0) done = (int)atomicAdd(&state[tiles + 1], 1ull);
  done = __shfl(done, 0, 64);
  if (done == tiles - 1)
    for (int i = lane; i < tiles + 2; i += 64)
      __hip_atomic_store(&state[i], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

// Plain sources and sinks.
template <typename T>
struct ArraySrc {
  const T* p;
  __device__ T operator()(size_t i) const { return p[i]; }
};
template <typename T>
struct ArraySink {
  T* p;
  __device__ void operator()(size_t i, T prefix, T) const { p[i] = prefix; }
};

inline size_t scan_blocks(size_t n) { return (n + kScanTile - 1) / kScanTile; }

// Exclusive scan of src[0, n) delivered to sink; partials needs scan_blocks(n)
// elements; total (device or host-mapped) receives the sum.  n > 0.
template <typename T, typename Src, typename Sink>
void scan_launch(Src src, size_t n, Sink sink, T* partials, T* total, hipStream_t st) {
  const size_t nb = scan_blocks(n);
  hipLaunchKernelGGL((k_scan_reduce<T, Src>), dim3((unsigned)nb), dim3(kScanThreads), 0, st, src, n, partials);
  hipLaunchKernelGGL((k_scan_partials<T>), dim3(1), dim3(kScanThreads), 0, st, partials, nb, total);
  hipLaunchKernelGGL((k_scan_down<T, Src, Sink>), dim3((unsigned)nb), dim3(kScanThreads), 0, st, src, n,
                     (const T*)partials, sink);
}

}  // namespace am
