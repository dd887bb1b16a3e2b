// Mesh assembly kernels and their host-side launch sequence (see am_combine.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "am_limits.h"

// am_combine_noalloc: largest inputs its static scratch takes, and the scratch size
#define AM_NOALLOC_MAX_REGIONS (1 << 18)
#define AM_NOALLOC_MAX_ITEMS (1 << 20)
#define AM_NOALLOC_SCRATCH_BYTES ((size_t)28 << 20)
#define AM_NOALLOC_MAX_DEVICES 64

namespace am {

struct CombineIn {
  const uint32_t* reg_mask;  // [R][Wd]
  const int* info;           // [R][AM_INFO_STRIDE]
  const int* vpair;          // [R][vm][2]
  const double* verts;       // [R][vm][3]
  int R, Wd, vm, L;
};

// Scratch of one combine: a region part (sized by R) and an item part (sized
// by the item count N, known after combine_count).
struct CombineWork {
  unsigned long long* roff = nullptr;   // [R] exclusive scan: faces before << 32 | items before
  unsigned long long* rpart = nullptr;  // scan partials of the region scan
  unsigned long long* totals = nullptr; // [0] faces << 32 | items; [1] (low 32 bits) vertices
  uint32_t* table = nullptr;            // [cap] item index of a key's smallest item, or empty
  uint32_t cap_mask = 0;
  int* item_region = nullptr;           // [N]
  int* slot = nullptr;                  // [N] table slot of the item's key
  unsigned* vflag = nullptr;            // [N] 1: first item of its key
  unsigned* vrank = nullptr;            // [N] exclusive scan of vflag
  unsigned* ipart = nullptr;            // scan partials of the item scan
};

struct CombineOut {
  double* xyz;      // [V][3]
  int* face_off;    // [F + 1]
  int* face_idx;    // [N]
  int F;
};

size_t combine_table_slots(long long N);
size_t combine_region_bytes(long long R);
size_t combine_item_bytes(long long N);
void combine_carve_regions(char* base, long long R, CombineWork& w);
void combine_carve_items(char* base, long long N, CombineWork& w);
// region scan: w.roff and w.totals[0]
void combine_count(const CombineIn& in, CombineWork& w, hipStream_t st);
// item region map, key table, first-occurrence flags and their scan (vertex count in w.totals[1])
void combine_dedup(const CombineIn& in, CombineWork& w, long long N, hipStream_t st);
// face-index array, vertex coordinates (v * scale + center) and face offsets
void combine_assign(const CombineIn& in, const CombineWork& w, const CombineOut& out, long long N, double scale,
                    const double* c3, hipStream_t st);

// Device address and size of this GPU's static scratch (the address is looked
// up once per device, with that device current).
hipError_t noalloc_scratch(void** out, size_t* bytes);

}  // namespace am
