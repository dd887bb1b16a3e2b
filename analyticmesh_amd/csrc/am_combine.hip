// Mesh assembly: vertex deduplication and face indexing of the region store
// (replaces backend/inc/polymesh.h:144-346, the quantised-coordinate hash).
//
// A vertex use ("item") is one vertex of one face.  Items are numbered region
// by region after an exclusive scan of the face sizes, so the item array is
// exactly the mesh's face-index array and no slot is spent on regions without
// a face or on the unused tail of the vertex cap.  A vertex is keyed by its two
// tight constraints (lo, hi) and its region's mask with those two bits
// cleared; every region around the vertex produces the same key.  The first
// item of each key (lowest item index) gives the vertex its index and
// coordinates, so vertex numbering follows (region, vertex) order.
//
// Passes:
//   1. region scan: packed (face flag << 32 | face size) -> first item and face
//      rank of every region; the down-sweep also writes each item's region.
//   2. k_cinsert, sixteen lanes per region: they read the region's mask once
//      (coalesced); the key hash is a sum of per-word hashes, so each item's
//      hash is the region's sum with the one or two changed words swapped out.
//      Items (four lanes each) claim a table slot by CAS; an item whose key is
//      already present lowers the slot to the smallest item of that key
//      (atomicMin).
//   3. first-occurrence flags, scanned -> vertex ranks and the vertex count.
//   4. k_cassign, one thread per item: vertex index into the face-index array,
//      coordinates by the first item, face offsets by each face's first item.
//
// Scratch comes from a caller-provided workspace (CombineWork), so the same
// kernels serve the allocating entry point and the one that allocates nothing.
#include <algorithm>
#include <climits>
#include <cstdio>
#include <mutex>

#include "am_combine.h"
#include "am_device.h"
#include "am_scan.h"

namespace am {

#define AM_CEMPTY 0xFFFFFFFFu

__device__ __forceinline__ uint64_t pair_hash(int lo, int hi) {
  return mix64((uint64_t(uint32_t(lo)) << 32) | uint32_t(hi));
}

struct RegionCountSrc {
  const int* info;
  __device__ unsigned long long operator()(size_t r) const {
    const int nv = info[r * AM_INFO_STRIDE + 2];
    return nv >= 3 ? ((1ULL << 32) | (unsigned long long)nv) : 0ULL;
  }
};
__global__ void k_citem_region(const unsigned long long* roff, const int* info, int R, int* item_region) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const int nv = info[(size_t)r * AM_INFO_STRIDE + 2];
  if (nv < 3) return;
  const int i0 = (int)(roff[r] & 0xffffffffu);
  for (int k = 0; k < nv; ++k) item_region[i0 + k] = r;
}
struct FlagSrc {
  const unsigned* p;
  __device__ unsigned operator()(size_t i) const { return p[i]; }
};

// Sixteen lanes per region, sixteen regions per 256-thread block: the lanes of
// a region read its mask once (coalesced) into LDS and sum the per-word
// hashes.  The region's items then go four at a time, four lanes per item: the
// item's first lane does the atomics, and a key compare reads the other
// region's mask as the four lanes' 16-byte quads (or words), one ballot
// deciding equality.  Every lane of a wave stays in the probe loops, whose
// trip counts are wave-uniform, so the ballots and shuffles see whole groups.
#define AM_CG 16
#define AM_CK 4
__global__ __launch_bounds__(256) void k_cinsert(CombineIn in, CombineWork w) {
  __shared__ uint32_t s_mask[256 / AM_CG][AM_MAX_WORDS];
  const int lane = threadIdx.x & 63;
  const int g = threadIdx.x & (AM_CG - 1), grp = threadIdx.x / AM_CG;
  const int r = blockIdx.x * (256 / AM_CG) + grp;
  const bool live = r < in.R;
  const int nv = live ? in.info[(size_t)r * AM_INFO_STRIDE + 2] : 0;
  const bool has = live && nv >= 3;
  uint64_t hs = 0;
  if (has)
    for (int q = g; q < in.Wd; q += AM_CG) {
      const uint32_t x = in.reg_mask[(size_t)r * in.Wd + q];
      s_mask[grp][q] = x;
      hs += word_hash(x, q);
    }
  // sum over the region's 16 lanes (mod 2^64)
#pragma unroll
  for (int d = AM_CG / 2; d > 0; d >>= 1) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)hs, d, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(hs >> 32), d, 64);
    hs += ((uint64_t)hi << 32) | lo;
  }
  __syncthreads();
  const uint32_t* m = s_mask[grp];
  const int item0 = has ? (int)(w.roff[r] & 0xffffffffu) : 0;
  const int j4 = g & (AM_CK - 1), sg = g / AM_CK, lead = lane & ~(AM_CK - 1);
  const bool quads = (in.Wd & 3) == 0;
  for (int k0 = 0; __ballot(has && k0 < nv) != 0ull; k0 += AM_CG / AM_CK) {
    const int k = k0 + sg;
    const bool act = has && k < nv;
    const int item = item0 + k;
    int lo = 0, hi = 0, wl = -1, wh = -1;
    uint32_t bl = 0u, bh = 0u, s = 0u;
    if (act) {
      const int a = in.vpair[((size_t)r * in.vm + k) * 2 + 0], b = in.vpair[((size_t)r * in.vm + k) * 2 + 1];
      lo = a < b ? a : b;
      hi = a < b ? b : a;
      // key hash: the region's word-hash sum with the words holding bits lo and hi cleared
      uint64_t h = hs;
      const bool lo_in = lo >= 0 && lo < in.L, hi_in = hi >= 0 && hi < in.L;
      wl = lo_in ? lo >> 5 : -1;
      wh = hi_in ? hi >> 5 : -1;
      bl = lo_in ? 1u << (lo & 31) : 0u;
      bh = hi_in ? 1u << (hi & 31) : 0u;
      if (wl >= 0) {
        uint32_t x = m[wl] & ~bl;
        if (wh == wl) x &= ~bh;
        h += word_hash(x, wl) - word_hash(m[wl], wl);
      }
      if (wh >= 0 && wh != wl) h += word_hash(m[wh] & ~bh, wh) - word_hash(m[wh], wh);
      h = mix64(h ^ pair_hash(lo, hi));
      s = (uint32_t)h & w.cap_mask;
    }
    bool done = !act;
    for (uint32_t probes = 0; __ballot(!done) != 0ull; ++probes) {
      uint32_t prev = 0u;
      if (!done && j4 == 0) {
        // read first (a slot only goes empty -> item -> smaller item here), claim when empty
        prev = __hip_atomic_load(&w.table[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == AM_CEMPTY) prev = atomicCAS(&w.table[s], AM_CEMPTY, (uint32_t)item);
      }
      prev = (uint32_t)__shfl((int)prev, lead, 64);
      bool cmp = false;
      const uint32_t* m2 = in.reg_mask;
      if (!done) {
        if (prev == AM_CEMPTY) {
          done = true;
        } else {
          // compare keys with the slot's item (records are immutable here; any
          // item the slot holds has the key of the item that claimed it)
          const int r2 = w.item_region[prev];
          const int k2 = (int)prev - (int)(w.roff[r2] & 0xffffffffu);
          const int a2 = in.vpair[((size_t)r2 * in.vm + k2) * 2 + 0], b2 = in.vpair[((size_t)r2 * in.vm + k2) * 2 + 1];
          // when the pairs agree, bits lo and hi are the same words in both keys
          cmp = (a2 < b2 ? a2 : b2) == lo && (a2 < b2 ? b2 : a2) == hi;
          m2 = in.reg_mask + (size_t)r2 * in.Wd;
        }
      }
      uint32_t diff = 0u;
      if (cmp) {
        if (quads) {
          const uint4* m24 = reinterpret_cast<const uint4*>(m2);
          for (int q4 = j4; q4 < (in.Wd >> 2); q4 += AM_CK) {
            const uint4 y = m24[q4];
            const uint32_t yv[4] = {y.x, y.y, y.z, y.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int q = 4 * q4 + e;
              const uint32_t clr = (q == wl ? bl : 0u) | (q == wh ? bh : 0u);
              diff |= (m[q] ^ yv[e]) & ~clr;
            }
          }
        } else {
          for (int q = j4; q < in.Wd; q += AM_CK) {
            const uint32_t clr = (q == wl ? bl : 0u) | (q == wh ? bh : 0u);
            diff |= (m[q] ^ m2[q]) & ~clr;
          }
        }
      }
      const unsigned long long bal = __ballot(diff != 0u);
      const bool eq = ((bal >> lead) & ((1ull << AM_CK) - 1)) == 0ull;
      if (cmp && eq) {
        if (j4 == 0 && (uint32_t)item < prev) atomicMin(&w.table[s], (uint32_t)item);
        done = true;
      } else if (!done) {
        if (probes >= w.cap_mask) done = true;   // table full: sized so that it never is
        else s = (s + 1) & w.cap_mask;
      }
    }
    if (act && j4 == 0) w.slot[item] = (int)s;
  }
}

__global__ void k_cflag(CombineWork w, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) w.vflag[i] = w.table[w.slot[i]] == (uint32_t)i ? 1u : 0u;
}

__global__ void k_cassign(CombineIn in, CombineWork w, CombineOut out, int N, double scale, double cx, double cy,
                          double cz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int r = w.item_region[i];
  const unsigned long long ro = w.roff[r];
  const int k = i - (int)(ro & 0xffffffffu);
  const uint32_t owner = w.table[w.slot[i]];
  const int id = (int)w.vrank[owner];
  out.face_idx[i] = id;
  if (owner == (uint32_t)i) {
    const double* v = in.verts + ((size_t)r * in.vm + k) * 3;
    out.xyz[(size_t)id * 3 + 0] = v[0] * scale + cx;
    out.xyz[(size_t)id * 3 + 1] = v[1] * scale + cy;
    out.xyz[(size_t)id * 3 + 2] = v[2] * scale + cz;
  }
  if (k == 0) out.face_off[(int)(ro >> 32)] = i;
  if (i == 0) out.face_off[out.F] = N;
}

__global__ void k_cfill(uint32_t* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = AM_CEMPTY;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t combine_table_slots(long long N) {
  size_t cap = 1024;
  while (cap < (size_t)2 * (size_t)N) cap *= 2;
  return cap;
}

size_t combine_region_bytes(long long R) {
  return align256((size_t)R * 8) + align256((size_t)(scan_blocks((size_t)R) + 1) * 8) + 256;
}

size_t combine_item_bytes(long long N) {
  const size_t n = (size_t)std::max<long long>(N, 1);
  return align256(combine_table_slots(N) * 4) + 4 * align256(n * 4) + align256((scan_blocks(n) + 1) * 4) + 256;
}

void combine_carve_regions(char* base, long long R, CombineWork& w) {
  size_t o = 0;
  w.roff = (unsigned long long*)(base + o);
  o += align256((size_t)R * 8);
  w.rpart = (unsigned long long*)(base + o);
  o += align256((size_t)(scan_blocks((size_t)R) + 1) * 8);
  w.totals = (unsigned long long*)(base + o);   // [0]: packed (faces << 32 | items), [1]: vertices
}

void combine_carve_items(char* base, long long N, CombineWork& w) {
  const size_t n = (size_t)std::max<long long>(N, 1);
  const size_t cap = combine_table_slots(N);
  size_t o = 0;
  w.table = (uint32_t*)(base + o);
  o += align256(cap * 4);
  w.cap_mask = (uint32_t)(cap - 1);
  w.item_region = (int*)(base + o);
  o += align256(n * 4);
  w.slot = (int*)(base + o);
  o += align256(n * 4);
  w.vflag = (unsigned*)(base + o);
  o += align256(n * 4);
  w.vrank = (unsigned*)(base + o);
  o += align256(n * 4);
  w.ipart = (unsigned*)(base + o);
}

void combine_count(const CombineIn& in, CombineWork& w, hipStream_t st) {
  scan_launch<unsigned long long>(RegionCountSrc{in.info}, (size_t)in.R, ArraySink<unsigned long long>{w.roff},
                                  w.rpart, w.totals, st);
}

void combine_dedup(const CombineIn& in, CombineWork& w, long long N, hipStream_t st) {
  const size_t cap = (size_t)w.cap_mask + 1;
  hipLaunchKernelGGL(k_citem_region, dim3((unsigned)((in.R + 255) / 256)), dim3(256), 0, st,
                     (const unsigned long long*)w.roff, in.info, in.R, w.item_region);
  hipLaunchKernelGGL(k_cfill, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, st, w.table, cap);
  hipLaunchKernelGGL(k_cinsert, dim3((unsigned)((in.R + 256 / AM_CG - 1) / (256 / AM_CG))), dim3(256), 0, st, in, w);
  hipLaunchKernelGGL(k_cflag, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, w, (int)N);
  scan_launch<unsigned>(FlagSrc{w.vflag}, (size_t)N, ArraySink<unsigned>{w.vrank}, w.ipart,
                        (unsigned*)(w.totals + 1), st);
}

void combine_assign(const CombineIn& in, const CombineWork& w, const CombineOut& out, long long N, double scale,
                    const double* c3, hipStream_t st) {
  hipLaunchKernelGGL(k_cassign, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, in, w, out, (int)N, scale,
                     c3[0], c3[1], c3[2]);
}

// ---------------------------------------------------------------------------
// Scratch of am_combine_noalloc: one static device array per GPU, reached
// through its device address (hipGetSymbolAddress, once per device).  Calls on
// one GPU share it: overlapping calls race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[AM_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[AM_NOALLOC_MAX_DEVICES] = {nullptr};
}  // namespace

hipError_t noalloc_scratch(void** out, size_t* bytes) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= AM_NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  *bytes = sizeof(scratch);
  return hipSuccess;
}

}  // namespace am
