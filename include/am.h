/* C-ABI of the MI355X analytic-marching engine (libam.so).
 *
 * Drop-in for the reference's native module `build.cuam`
 * (backend/src/cuam.cpp:186-217): the same five calls, with plain pointers,
 * status codes instead of exit(), and one context per caller instead of the
 * process-global singleton (backend/src/cuam_kernel.cu:22-25).
 *
 * Pointers to weights, seeds and planes may be host or device memory; the
 * engine copies them into its own device buffers during the call.  All
 * results stay in engine-owned memory until the next am_run or am_destroy.
 */
#ifndef AM_H
#define AM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct am_ctx am_ctx;
typedef int am_status;

#define AM_OK 0
#define AM_ERR_ARG 1        /* invalid argument (reference: TORCH_CHECK, cuam.cpp:69-180) */
#define AM_ERR_STATE 2      /* call out of order (reference: prints and returns, cuam.cpp:107-111) */
#define AM_ERR_HIP 3        /* HIP runtime error (reference: exit(), utilities.h:73-95) */
#define AM_ERR_CAPACITY 4   /* a configured capacity was exceeded */
#define AM_ERR_NOGPU 5      /* no usable gfx950 device */
#define AM_ERR_IO 6         /* file could not be written */

/* Arithmetic of the constraint planes and of the walk. */
#define AM_PREC_FP32 0            /* FP32 MFMA planes, FP64 re-decision of undecided regions */
#define AM_PREC_SPLIT_BF16X3 1    /* split-BF16 (3 products) MFMA planes, FP64 re-decision */
#define AM_PREC_FP64 2            /* FP64 everywhere (the reference's default dtype) */
#define AM_PREC_FP64_RAY 3        /* FP64 planes, checked FP64 ray walk, FP64 candidate scan where undecided */

/* Element type of caller buffers. */
#define AM_DTYPE_F32 0
#define AM_DTYPE_F64 1

/* Replaces Init(float_type, nodesnum, arc_table, num_extra_constraints)
 * (backend/src/cuam.cpp:58-95, backend/src/cuam_kernel.cu:128-148).
 * nodes: n_nodes widths, nodes[0] == 3, nodes[n_nodes-1] == 1, n_nodes >= 3.
 * arc_table: row-major int32 [arc_rows][arc_cols], arc_rows == n_nodes - 2,
 * row r = [count, from0, tm0, from1, tm1, ...] (backend/model.py:23-31).
 * precision: AM_PREC_*.  Fails with AM_ERR_NOGPU when no device is present. */
am_status am_create(am_ctx** out, int n_nodes, const int* nodes, const int* arc_table, int arc_rows,
                    int arc_cols, int num_extra_constraints, int precision);

/* Replaces AnalyticMarching(weights, biases, states, points, arc_tm,
 * w_extra_constraints, b_extra_constraints, iso, flip_insideout)
 * (backend/src/cuam.cpp:97-184, backend/src/cuam_kernel.cu:30-123).
 * weights[i]: (nodes[i+1], nodes[i]) row-major; biases[i]: (nodes[i+1],);
 * tms[t]: (tm_shapes[2t], tm_shapes[2t+1]) row-major, a (0, 0) shape is the
 * identity; points: (n_seeds, 3); states: (n_seeds, L) bytes, nonzero = active
 * (L = sum of hidden widths); extra_w: (E, 3); extra_b: (E,).  Element type of
 * every floating-point buffer: dtype (AM_DTYPE_*).  stream: a hipStream_t or
 * NULL.  The call returns once the mesh records are complete. */
am_status am_run(am_ctx* ctx, const void* const* weights, const void* const* biases, int n_tm,
                 const int* tm_shapes, const void* const* tms, int dtype, const void* points,
                 const uint8_t* states, int n_seeds, const void* extra_w, const void* extra_b, double iso,
                 int flip_insideout, void* stream);

/* Replaces CombineMesh(scale, center) (backend/src/cuam_kernel.cu:185-200):
 * deduplicates vertices and indexes faces; output vertex = v * scale + center. */
am_status am_combine(am_ctx* ctx, double scale, const double* center3);

/* Variant of am_combine that allocates nothing (no counterpart in the reference;
 * for callers that cannot allocate device memory: code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same mesh as
 * am_combine + am_get_mesh, written into the caller's DEVICE arrays: vertices
 * (vertex_cap x 3 double), face_offsets (face_cap >= faces + 1 ints),
 * face_indices (index_cap ints); counts3 (host) receives vertices, faces, face
 * indices.  Scratch is a static device array of 28 MiB per GPU (held until the
 * process ends), which bounds the input to 2^18 regions and 2^20 vertex uses;
 * larger meshes, or outputs too small, fail with AM_ERR_CAPACITY.  All calls on
 * one GPU share that scratch: calls that overlap on one GPU race.  The
 * context's own mesh (am_get_mesh, am_export_ply, am_adjoint) is not changed. */
am_status am_combine_noalloc(am_ctx* ctx, double scale, const double* center3, double* vertices,
                             long long vertex_cap, int* face_offsets, long long face_cap, int* face_indices,
                             long long index_cap, long long* counts3);

/* Replaces ExportMesh(file_path, is_polymesh, is_float32)
 * (backend/src/cuam_kernel.cu:205-246, backend/inc/polymesh.h:350-423):
 * binary little-endian PLY, polygons or fan triangles, float or double xyz. */
am_status am_export_ply(am_ctx* ctx, const char* path, int is_polymesh, int is_float32);

/* Replaces Destroy() (backend/src/cuam_kernel.cu:153-179). */
am_status am_destroy(am_ctx* ctx);

/* Message of the last failing call on ctx (or of a failed am_create when ctx is NULL). */
const char* am_last_error(const am_ctx* ctx);

/* Sizes after am_run / am_combine. */
am_status am_get_counts(am_ctx* ctx, long long* n_regions, long long* n_faces, long long* n_vertices,
                        long long* n_face_indices);

/* Combined mesh: vertices (V x 3 double), face_offsets (F + 1), face_indices.  The arrays may be
 * host or device memory; the mesh is copied from device memory straight into them. */
am_status am_get_mesh(am_ctx* ctx, double* vertices, int* face_offsets, int* face_indices);

/* Per-region records in visiting order (parity view): masks [R][W] u32,
 * info [R][8] (start queued, start used, vertices, tab entries, reversed,
 * status, source: 0 or 1 an FP64 walk, 2 a reduced-precision walk, then 1
 * where the FP32 form's Newton step refined the vertices, else 0), tabs
 * [R][vm+1], verts [R][vm][3], vpairs [R][vm][2],
 * points [R][3] (seed point each region was queued with).  Any pointer may be NULL. */
am_status am_get_regions(am_ctx* ctx, uint32_t* masks, int* info, int* tabs, double* verts, int* vpairs,
                         double* points);

/* Gradient of L = sum_v <dL/dv, v> with respect to every weight and bias,
 * with the activation regions held fixed (no counterpart in the reference,
 * which has no differentiation).  grad_vertices: (V, 3) in the vertex order
 * of am_get_mesh (before scale/center); grad_weights[i] / grad_biases[i]
 * receive arrays shaped like weights[i] / biases[i] of am_run, in `dtype`.
 * Requires am_combine and a network without skip connections. */
am_status am_adjoint(am_ctx* ctx, const void* grad_vertices, int dtype, void* const* grad_weights,
                     void* const* grad_biases);

/* Constraint planes of n regions (packed masks [n][W]) computed by the
 * context's reduced-precision face kernel, for accuracy checks: planes
 * [n][4][C] f32 (x, y, z, c arrays, signed so that a.x + c <= 0 inside),
 * eq [n][4] f32 (f = iso plane), and, when sigma is not NULL, the kernel's
 * error estimate of each plane, sigma [n][C + 1] f32 (the constraints', then
 * the face plane's).  Needs a prior am_run on this context.
 * No counterpart in the reference (its planes stay internal). */
am_status am_planes(am_ctx* ctx, const uint32_t* masks, int n, float* planes, float* eq, float* sigma);

/* Order-independent topology fingerprint of the last am_run (no counterpart in
 * the reference; a parity view for full-size runs): out5[0..2] are sums mod
 * 2^64 of a per-row hash over the visited masks, the canonical faces and the
 * vertex keys (one row per vertex use); out5[3] faces, out5[4] vertex uses.
 * Rows and hash are those of analyticmesh_amd/canonical.py's fingerprint(). */
am_status am_topology_fingerprint(am_ctx* ctx, unsigned long long* out5);

/* Tunables: "vm" (polygon vertex cap, 3..64, default 20 like the reference's
 * _VAR_VERT_MAX), "table_log2" (initial visited-table size), "max_regions"
 * (stop after the breadth-first level during which this many regions were
 * processed, -1 = unlimited), "region_limit" (am_run fails with
 * AM_ERR_CAPACITY once more regions than this would be stored; -1 = the
 * engine's own limit, min(2^30, (2^31 - 1) / vm)), "ray_waves" (waves per
 * region of the FP64 ray walk: 2, 4 (default) or 8; the records do not depend
 * on it), "scan_blocks" (grid of the candidate scan over the regions the ray
 * walk leaves undecided, default 2048), "graph_levels" (1: each breadth-first
 * level's launches, every form, are captured into a hipGraph and launched as
 * one; the executable graph is updated in place from level to level; the
 * per-stage kernel times of the stats are not taken then),
 * "walk" (0 ray walk, 1
 * candidate scan, for the reduced forms), "ray_waves_reduced" (waves per
 * region of the reduced forms' ray walk: 2, 4 (default) or 8; 4 measured
 * fastest on the 4x256 net, 2 and 8 give the same records), "reduced_vertices" (FP32 form: 3
 * default, each vertex solved in FP64 from the reduced planes that the face
 * kernel's error estimate does not put within 1e-5 of |v| takes one FP64
 * Newton step after the march, its residuals from an FP64 forward pass of the
 * network under its region's mask; 2 every such vertex takes the step; 0 every
 * vertex of a region the reduced walk decides is re-solved from FP64 planes;
 * 1 a region's reduced-plane vertices are kept where the estimate puts all of
 * them within 1e-5 of |v|, the other regions' are re-solved; the split-BF16
 * form always re-solves),
 * "adjoint_tiled" (1 default: the
 * adjoint on the tiled FP64 chain and GEMMs; 0 one kernel per region),
 * "adjoint_mfma" (1 default: those GEMMs on the FP64 matrix cores; 0 VALU),
 * "chain_mfma" (1 default: the FP64 plane chain of the forms that certify
 * their decisions -- the FP64 ray walk, the reduced forms' FP64 redo, the
 * batch engine's ray walk -- runs on the FP64 matrix cores, planes equal to
 * the sequential-fma chain's up to the rounding of another summation order,
 * same topology; 0: the sequential-fma chain everywhere.  The bit-exact FP64
 * form always uses the sequential-fma chain). */
am_status am_set_option(am_ctx* ctx, const char* key, long long value);

/* Timings and counters of the last am_run/am_combine; see AM_STAT_* indices. */
#define AM_STAT_RUN_MS 0        /* am_run wall time, seeds in -> records in device memory */
#define AM_STAT_REGION_MS 1     /* region kernels (HIP events) */
#define AM_STAT_INSERT_MS 2     /* visited-set insertion + compaction (HIP events) */
#define AM_STAT_GEMM_MS 3       /* plane GEMM kernels (HIP events; reduced precisions) */
#define AM_STAT_WALK_MS 4       /* walk kernels after the GEMM (HIP events; reduced precisions) */
#define AM_STAT_REDO_MS 5       /* FP64 re-decision kernels (HIP events) */
#define AM_STAT_LEVELS 6
#define AM_STAT_REGIONS 7
#define AM_STAT_CANDIDATES 8
#define AM_STAT_REDO_REGIONS 9
#define AM_STAT_COMBINE_MS 10
#define AM_STAT_REGION_LAUNCHES 11
#define AM_STAT_ADJOINT_MS 12   /* am_adjoint kernels (wall time) */
#define AM_STAT_CHAIN64_MS 13   /* FP64 tiled plane chain, all launches (HIP events) */
#define AM_STAT_WALK64_MS 14    /* FP64 walk after the tiled chain, all launches (HIP events) */
#define AM_STAT_SETUP_MS 15     /* am_run set-up: weight uploads and operand layouts (wall time) */
#define AM_STAT_WHY0 16         /* 16..27: regions the reduced-precision walk left to FP64, per reason:
                                   seed shot, start edge, plane already in polygon, sign within bound,
                                   proposed vertex infeasible, ill-conditioned system, turn within bound,
                                   vertex cap, open edge, orientation within bound, candidate scan,
                                   same vertex */
#define AM_STAT_RAY64_MS 28     /* FP64 checked ray walk (HIP events) */
#define AM_STAT_SCAN64_REGIONS 29 /* regions decided by the FP64 candidate scan */
#define AM_STAT_RAY64_UNDECIDED 30 /* regions the FP64 ray walk left to the scan */
#define AM_STAT_CHAIN64_LAUNCHES 31 /* launches of the FP64 tiled chain */
#define AM_STAT_VERT64_MS 32    /* FP64 vertex re-solve of regions the reduced-precision walk decided (HIP events) */
#define AM_STAT_VERT64_REGIONS 33 /* regions whose vertices were re-solved from FP64 planes */
#define AM_STAT_FP64_ONLY 34    /* 1: a reduced form ran on the FP64 region kernel (skips or layers over 256 wide) */
#define AM_STAT_HOST_WAITS 35   /* host waits on the device during the run */
#define AM_STAT_GRAPH_LAUNCHES 36 /* levels launched as one captured graph ("graph_levels") */
#define AM_STAT_NEWTON_MS 37    /* FP32 form: FP64 Newton step of the vertices (HIP events, with its vertex dedup) */
#define AM_STAT_NEWTON_VERTICES 38 /* vertices it refined */
#define AM_STAT_COUNT 39
am_status am_get_stats(am_ctx* ctx, double* out, int n);

/* ---- Many networks in one job (no counterpart in the reference, which marches
 * one network per Init and re-runs Init when the architecture changes,
 * backend/main.py:437-448, backend/src/cuam_kernel.cu:53-98).  All networks of a
 * batch march together: each breadth-first level of every network goes through
 * one launch of each kernel (the FP64 plane chain over tiles of one network,
 * the walks over regions of any network), with one region store and visited
 * set.  Per network the records equal those of its own am_run in the same
 * precision.  Networks: no skips, no extra planes, layers up to 256 wide. */
typedef struct am_batch am_batch;

/* n_nets networks; network i has n_nodes[i] widths nodes[i][0..n_nodes[i]).
 * precision: AM_PREC_FP64 (candidate scan, the FP64 records of am_run) or
 * AM_PREC_FP64_RAY (checked FP64 ray walk, scan where undecided). */
am_status am_batch_create(am_batch** out, int n_nets, const int* n_nodes, const int* const* nodes, int precision);

/* "vm", "table_log2", "region_limit", "ray_waves", "scan_blocks" as for am_set_option. */
am_status am_batch_set_option(am_batch* b, const char* key, long long value);

/* weights / biases: the linear layers of network 0, then of network 1, ...
 * (n_nodes[i] - 1 arrays each, shaped as for am_run); points[i] (n_seeds[i] x 3)
 * and states[i] (n_seeds[i] x L_i bytes) are network i's seeds. */
am_status am_batch_run(am_batch* b, const void* const* weights, const void* const* biases, int dtype,
                       const void* const* points, const uint8_t* const* states, const int* n_seeds, double iso,
                       int flip_insideout, void* stream);

/* Topology fingerprint of one network's regions (as am_topology_fingerprint;
 * out6[5] = its region count). */
am_status am_batch_fingerprint(am_batch* b, int net, unsigned long long* out6);

/* One network's region records in its visiting order (layouts of
 * am_get_regions with that network's mask words); n_regions receives the
 * count; with every array NULL only the count is returned. */
am_status am_batch_get_regions(am_batch* b, int net, long long* n_regions, uint32_t* masks, int* info, int* tabs,
                               double* verts, int* vpairs, double* points);

/* Vertex dedup and face indexing of all networks at once (vertices of
 * different networks never merge); counts and stats as for one context. */
am_status am_batch_combine(am_batch* b, double scale, const double* center3);
am_status am_batch_get_counts(am_batch* b, long long* n_regions, long long* n_faces, long long* n_vertices,
                              long long* n_face_indices);
am_status am_batch_get_stats(am_batch* b, double* out, int n);
const char* am_batch_last_error(const am_batch* b);
am_status am_batch_destroy(am_batch* b);

#ifdef __cplusplus
}
#endif
#endif
