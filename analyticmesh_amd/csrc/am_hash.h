// Visited-set, compaction and mesh-assembly kernels (device side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "am_kernels.h"

namespace am {

struct VertexIO {
  const uint32_t* reg_mask;  // [R][Wd]
  const int* info;           // [R][AM_INFO_STRIDE]
  const int* vpair;          // [R][vm][2]
  const double* verts;       // [R][vm][3]
  uint32_t* table;           // [cap]
  int* status;               // [R*vm]
  int* winner;               // [R*vm]
  int* owner;                // [R*vm]
  int n;                     // R * vm
  int vm, L, Wd;
  uint32_t cap_mask;
};

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v);
__global__ void k_fill_i32(int* p, size_t n, int v);
__global__ void k_insert(InsertIO io);
__global__ void k_mark_first(const int* status, const int* winner, int* first, int* link, int n);
__global__ void k_commit(const int* first, const int* rank, const int* link, const int* slot,
                         const uint32_t* cand_mask, const int* cand_start, const double* cand_point,
                         uint32_t* reg_mask, int* reg_start, double* reg_point, uint32_t* table,
                         int n, int Wd, int R0);
__global__ void k_rehash(const uint32_t* reg_mask, int R, int Wd, uint32_t* table, uint32_t cap_mask);
__global__ void k_vinsert(VertexIO io);
__global__ void k_vmark_first(const int* status, const int* winner, int* first, int n);
__global__ void k_vassign(VertexIO io, const int* first, const int* rank, int* vid, double* out_xyz,
                          double scale, double cx, double cy, double cz);
__global__ void k_face_flags(const int* info, int R, int* fflag, int* fnv);
__global__ void k_face_write(const int* info, const int* fflag, const int* frank, const int* foff, const int* vid,
                             int R, int vm, int* face_off, int* face_idx);

__global__ void k_fingerprint(const uint32_t* reg_mask, const int* info, const int* vpair, int R, int Wd, int vm,
                              int L, unsigned long long* out);

}  // namespace am
