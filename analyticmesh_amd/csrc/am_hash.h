// Visited-set and compaction kernels, and the topology fingerprint (device side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "am_kernels.h"

namespace am {

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v);
__global__ void k_fill_i32(int* p, size_t n, int v);
__global__ void k_insert(InsertIO io);
__global__ void k_mark_first(const int* status, const int* winner, int* first, int* link, int n);
__global__ void k_commit(const int* first, const int* rank, const int* link, const int* slot,
                         const uint32_t* cand_mask, const int* cand_start, const double* cand_point,
                         uint32_t* reg_mask, int* reg_start, double* reg_point, uint32_t* table,
                         int n, int Wd, int R0);
__global__ void k_rehash(const uint32_t* reg_mask, int R, int Wd, uint32_t* table, uint32_t cap_mask);
__global__ void k_fingerprint(const uint32_t* reg_mask, const int* info, const int* vpair, int R, int Wd, int Wh,
                              int vm, int L, int net, unsigned long long* out);
__global__ void k_net_hist(const uint32_t* reg_mask, int Wd, int first, int count, int* hist);
__global__ void k_net_flags(const uint32_t* reg_mask, int Wd, int R, int net, int* flags);
__global__ void k_net_scatter(const int* flags, const int* rank, int R, int* ids);

}  // namespace am
