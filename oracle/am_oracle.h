// C API of the CPU restatement (test infrastructure only; see am_oracle.cpp).
#ifndef AM_ORACLE_H
#define AM_ORACLE_H
#include <stdint.h>

#define AMO_INFO_STRIDE 8
// info[r*8 + 0] start index the region was queued with (-1 = seed)
// info[r*8 + 1] start plane after the seed search (-1 = none found)
// info[r*8 + 2] number of vertices
// info[r*8 + 3] number of tab entries
// info[r*8 + 4] 1 if the vertex list was reversed by the orientation step
// info[r*8 + 5] status (AMO_ST_*)
#define AMO_ST_OK 0
#define AMO_ST_NO_START 1
#define AMO_ST_GAVE_UP 2
#define AMO_ST_BAD_START 3

#ifdef __cplusplus
extern "C" {
#endif
uint32_t amo_fnv1a(const uint8_t* d, int n);
uint32_t amo_fnv1(const uint8_t* d, int n);
void amo_get_i_j(int x, int ij_max, int* i, int* j);
void* amo_create(int n_nodes, const int* nodes, const double* const* W, const double* const* b,
                 const int* arc_table, int arc_rows, int arc_cols,
                 int n_tm, const int* tm_shapes, const double* const* tms,
                 int E, const double* ew, const double* eb,
                 int precision, int traversal, int vm, int threads);
int amo_run(void* h, int n_seeds, const double* seed_points, const uint8_t* seed_states,
            double iso, int flip, long max_regions);
long amo_num_regions(void* h);
int amo_mask_words(void* h);
int amo_num_constraints(void* h);
double amo_seconds(void* h);
long amo_tag_collisions(void* h);
long amo_num_batches(void* h);
int amo_truncated(void* h);
void amo_get(void* h, uint32_t* masks, int* info, int* tabs, double* verts, int* vpairs, double* points);
void amo_planes(void* h, const uint32_t* mask, double iso, double* A, double* eq);
void amo_destroy(void* h);
#ifdef __cplusplus
}
#endif
#endif
