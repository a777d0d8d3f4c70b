/* C ABI of the MI355X screened Poisson reconstruction library
 * (libwarp_psr_amd.so).  Plain pointers and sizes only; every entry point
 * returns 0 on success, 2 for invalid input (the reference raises ValueError,
 * message text identical), 3 for a type mismatch (TypeError), 1 otherwise;
 * psr_last_error() returns the calling thread's last message.
 *
 * Device pointers are HIP device pointers on the current device; streams are
 * hipStream_t passed as void* (NULL = default stream).  A PsrResult is used
 * from one stream at a time; independent results may run on distinct streams.
 *
 * Reference interfaces replaced (alecjacobson/warp-psr, Python over Warp):
 *   psr_params_default  -> keyword defaults of reconstruct()      src/warp_psr/reconstruction.py:137-165
 *   psr_reconstruct     -> reconstruct(positions, normals, ...)   src/warp_psr/reconstruction.py:137-349
 *   psr_evaluate        -> Reconstruction.evaluate / LevelField.evaluate
 *                                                                 src/warp_psr/reconstruction.py:129-130, field.py:150-184
 *   psr_result_iso      -> Reconstruction.iso_value               src/warp_psr/reconstruction.py:124-127
 *   psr_result_field    -> Reconstruction.device_field()          src/warp_psr/reconstruction.py:132-134,
 *                                                                 field.py:140-148, device_field.py:115-142
 *   psr_result_metadata_json -> Reconstruction.metadata          src/warp_psr/reconstruction.py:303-346
 *   psr_result_array    -> Reconstruction.diagnostics (stage arrays)
 *                                                                 src/warp_psr/reconstruction.py:303-349
 *   psr_relax_sliced    -> relax_sliced(A, b, x, schedule, depth, iters)
 *                                                                 src/warp_psr/relaxation.py:189-238
 *   psr_relax_scheduled -> relax_scheduled(A, b, x, schedule, iters, forward, z)
 *                                                                 src/warp_psr/relaxation.py:221-238,261-271
 *   psr_result_free     -> (Python garbage collection of Reconstruction)
 *   psr_field_vjp       -> DifferentiableReconstruction.evaluate backward
 *                                                                 src/warp_psr/field_adjoint.py:255-288, field.py:214-286
 *   psr_backward        -> reconstruct_differentiable backward    src/warp_psr/reconstruction_adjoint.py:89-177
 *   psr_field_vjp_inputs, psr_backward_inputs -> the same, for positions and normals
 *                                                                 src/warp_psr/reconstruction_adjoint.py:89-177
 *   PsrParams.float_operators -> the mixed precision mode (no reference counterpart;
 *                                float64 run with float32 finer-level operators)
 */
#ifndef WARP_PSR_AMD_H
#define WARP_PSR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct PsrParams {
  int depth;               /* default 6 */
  int full_depth;          /* default 3 */
  int base_depth;          /* -1 = full_depth */
  int solve_depth;         /* -1 = depth */
  int kernel_depth;        /* -1 = max(depth - 2, 0) */
  double samples_per_node; /* default 1.5 */
  double scale;            /* default 1.1 */
  double point_weight;     /* default 2.0; 0 selects PSR */
  int exact_interpolation; /* default 0 */
  int iters;               /* default 8 */
  int base_v_cycles;       /* default 1 */
  double cg_accuracy;      /* default 1e-3 */
  int keep_intermediates;  /* keep per-stage arrays for psr_result_array (default 1) */
  int float_operators;     /* 1: mixed precision -- a float64 run whose finer-level Laplacian +
                              screening operators are stored as float32 (default 0; float64 only) */
  int relax_tile;          /* finer-level sliced Gauss-Seidel: side of the (x, y) tiles that one
                              workgroup relaxes per wave (even, 8..40), 0 = library default, -1 = one
                              launch per colour and wave; every choice gives the same bits */
} PsrParams;

typedef struct PsrResult PsrResult;

/* Borrowed device view of a reconstructed field (zero-copy; valid while the
 * result lives and is not mutated).  Mirrors the reference DeviceField struct. */
typedef struct PsrField {
  const int64_t* keys;        /* sorted basis keys d<<48|x<<32|y<<16|z, all levels */
  const void* coefficients;   /* per-key level corrections (float or double) */
  const void* cumulative;     /* per-key cumulative coefficients */
  const int32_t* level_offsets; /* device, depth+2 entries: keys of level d in [off[d], off[d+1]) */
  int64_t count;
  int depth;
  int dtype;                  /* 0 float32, 1 float64 */
  double linear[9];           /* model -> cube, row major */
  double translation[3];
  double iso_value;
} PsrField;

void psr_params_default(PsrParams* p);

/* points, normals: device arrays of n x 3 values (dtype 0 = float32, 1 = float64). */
int psr_reconstruct(const void* points, const void* normals, int64_t n, int dtype, const PsrParams* params,
                    void* stream, PsrResult** out);

/* Reconstructs `batch` independent clouds concurrently: each on one of up to
 * max_concurrency (<= 0: 8) library worker streams and host threads, after
 * the caller's `stream` (which produced the inputs).  points[i]/normals[i]
 * are device arrays of counts[i] x 3 (all of `dtype`); params holds one
 * PsrParams for all clouds (params_count 1) or one per cloud.  Returns when
 * every result is complete; out[i] receives cloud i's result.  Invalid
 * parameters fail before any device work; on any failure every result is
 * freed, out[] is NULLed and the message names the cloud.  No reference
 * counterpart (the reference reconstructs one cloud per call,
 * src/warp_psr/reconstruction.py:137-349); for callers with many small clouds. */
int psr_reconstruct_batch(const void* const* points, const void* const* normals, const int64_t* counts, int batch, int dtype,
                          const PsrParams* params, int params_count, int max_concurrency, void* stream, PsrResult** out);

/* queries: device m x 3 in model space; values m, grads m x 3 (may be NULL), valid m int32. */
int psr_evaluate(const PsrResult* result, const void* queries, int64_t m, void* values, void* grads, int32_t* valid,
                 void* stream);

int psr_result_iso(const PsrResult* result, double* iso);
int psr_result_field(const PsrResult* result, PsrField* field);
/* Writes NUL-terminated JSON into buf (size bytes); *needed receives the full length + 1. */
int psr_result_metadata_json(const PsrResult* result, char* buf, size_t size, size_t* needed);
/* Host copy (owned by the result) of a named stage array, e.g. "sample_keys", "densities",
 * "normal_keys", "tree_keys"/"basis_keys" (per level), "matrix_values"/"cross_values"/"rhs"/
 * "corrections" (per solved level), "keys", "coefficients", "cumulative"; level -1 for
 * whole-run arrays.  Integer arrays are int32/int64, values in the run precision. */
int psr_result_array(PsrResult* result, const char* name, int level, const void** data, int64_t* nbytes);
void psr_result_free(PsrResult* result);

/* Sliced Gauss-Seidel (reference order) on a caller CSR operator, all device
 * pointers: keys[rows] sorted basis keys of one level, CSR offsets[rows+1],
 * columns/values (<= 27 entries per row), b[rows], x[rows] updated in place. */
int psr_relax_sliced(int dtype, const int64_t* keys, int64_t rows, const int32_t* offsets, const int32_t* columns,
                     const void* values, const void* b, void* x, int depth, int iters, void* stream);

/* 8-colour Gauss-Seidel on the same caller CSR arguments: z < 0 runs `iters`
 * sweeps over the colour groups (x&1)+2(y&1)+4(z&1) of row_schedule(keys,
 * depth), colours 0..7 (forward != 0) or 7..0; z in [0, 2**depth] relaxes slice
 * z of the sliced schedule instead, its colours in the same order (the z mode
 * of relax_scheduled).  Returns 2 for z beyond 2**depth. */
int psr_relax_scheduled(int dtype, const int64_t* keys, int64_t rows, const int32_t* offsets, const int32_t* columns,
                        const void* values, const void* b, void* x, int depth, int iters, int forward, int z, void* stream);

/* Reverse mode (first order, octree / grouping / splat depths / solver branches
 * fixed; src/warp_psr/reconstruction_adjoint.py:61-185).
 *
 * psr_field_vjp: value seeds (m) and model-gradient seeds (m x 3, either may be
 * NULL) at model-space queries -> float64 gradients w.r.t. the K concatenated
 * coefficients and cumulative coefficients (accumulated; K = PsrField.count)
 * and w.r.t. the queries (m x 3 in the run precision, accumulated; may be NULL).
 * Mirrors DifferentiableField.evaluate's backward (field_adjoint.py:62-104,
 * 255-288, field.py:214-286).
 *
 * psr_backward: coefficient / cumulative gradients (either may be NULL) and an
 * isovalue seed -> gradient w.r.t. the input normals (n x 3, accumulated);
 * `normals` are the forward's input normals.
 *
 * psr_field_vjp_inputs: psr_field_vjp plus transform_grad (4 float64 on the
 * device, accumulated, may be NULL): the loss's gradient w.r.t. the sample
 * normalisation (inv, t1, t2, t3) through the model-to-cube query mapping
 * (transforms.py:56-79), which depends on the input positions' bounds.
 *
 * psr_backward_inputs: the same seeds (with transform_grad from
 * psr_field_vjp_inputs, or NULL) -> gradients w.r.t. the input positions
 * and/or normals (n x 3 each, accumulated; one output may be NULL).  `points`
 * and `normals` are the forward's inputs.  Position gradients go through the
 * isovalue centroids, the normal splats' basis, depth and weight, the density
 * levels and aggregates, and the sample transform with its extrema owners
 * (sample_adjoint.py:16-102, splat_adjoint.py:104-188) and, in screened
 * runs, the constraint points' screening terms in each level's operator,
 * cross operator and point load (assembly_adjoint.py:44-91,
 * constraints.py:55-84), the base hierarchy's Galerkin products and the
 * normal average's weight.  Screened position gradients replay every
 * relaxation once (about the forward solve's cost again). */
int psr_field_vjp(PsrResult* result, const void* queries, int64_t m, const void* value_seed, const void* grad_seed,
                  double* coef_grad, double* cum_grad, void* query_grad, void* stream);
int psr_backward(PsrResult* result, const double* coef_grad, const double* cum_grad, double iso_seed, const void* normals,
                 int64_t n, void* grad_normals, void* stream);
int psr_field_vjp_inputs(PsrResult* result, const void* queries, int64_t m, const void* value_seed, const void* grad_seed,
                         double* coef_grad, double* cum_grad, void* query_grad, double* transform_grad, void* stream);
int psr_backward_inputs(PsrResult* result, const double* coef_grad, const double* cum_grad, double iso_seed,
                        const double* transform_grad, const void* points, const void* normals, int64_t n, void* grad_points,
                        void* grad_normals, void* stream);

/* Marching cubes on a uniform model-space grid (src/warp_psr/meshing.py:150-280,
 * the sparse backend's steps), all buffers on the device and on `stream`:
 *   psr_mesh_begin       grid of shape[0] x shape[1] x shape[2] nodes between
 *                        lower and upper (float32 spacing (upper-lower)/(n-1));
 *                        returns 2 with the reference's messages for bad grids
 *   psr_mesh_slab_nodes  node count of planes [start, start + width]
 *   psr_mesh_slab_points their model-space points (3 floats each, x-major)
 *   psr_mesh_slab_values classify the cells of planes [start, start + width)
 *                        from the caller's values at those points
 *   psr_mesh_select      count and list the crossing cells (value >= 0 is
 *                        positive); invalid != 0 if a value was not finite
 *   psr_mesh_corner_points  the 8 corner points of every retained cell, in the
 *                        corner order (000,100,110,010,001,101,111,011)
 *   psr_mesh_triangulate from the caller's values at those corners: one
 *                        vertex per crossing grid edge, triangles from the
 *                        library's case table (psr_mesh_case_table); faces
 *                        point toward increasing values
 *   psr_mesh_copy        vertices (3 floats each) and indices (3 per triangle)
 *   psr_mesh_vertex_vjp  vertex gradients -> corner-value gradients (added)
 * The corner values given to psr_mesh_triangulate must stay valid for
 * psr_mesh_vertex_vjp.  Each mesh is used from one stream at a time. */
typedef struct PsrMesh PsrMesh;
int psr_mesh_begin(const int32_t shape[3], const float lower[3], const float upper[3], void* stream, PsrMesh** out);
int psr_mesh_slab_nodes(PsrMesh* mesh, int32_t start, int32_t width, int64_t* nodes);
int psr_mesh_slab_points(PsrMesh* mesh, int32_t start, int32_t width, float* points);
int psr_mesh_slab_values(PsrMesh* mesh, int32_t start, int32_t width, const float* values);
int psr_mesh_select(PsrMesh* mesh, int64_t* retained, int32_t* invalid);
int psr_mesh_corner_points(PsrMesh* mesh, float* points);
int psr_mesh_triangulate(PsrMesh* mesh, const float* corner_values, int64_t* vertices, int64_t* triangles);
int psr_mesh_copy(PsrMesh* mesh, float* vertices, int32_t* indices);
int psr_mesh_vertex_vjp(PsrMesh* mesh, const float* corner_values, const float* grad_vertices, float* grad_values);
/* The 256-case table: counts[256] triangles per case, edges[256][36] cube
 * edges per triangle; edge k joins corners (lower, upper) =
 * (0,1),(1,2),(3,2),(0,3),(4,5),(5,6),(7,6),(4,7),(0,4),(1,5),(2,6),(3,7). */
int psr_mesh_case_table(int8_t* counts, int8_t* edges);
void psr_mesh_free(PsrMesh* mesh);

/* Tree metadata (src/warp_psr/tree_metadata.py:95-153, build_tree_metadata):
 * every retained brood of every level (the active support nodes and the
 * broods of the finer level's parents, the samples, the density splat
 * records and the normal records), in global lattice coordinates (local +
 * 2^(d+1), tree-key encoded, sorted), after two padding levels (local depths
 * -2 and -1).  Per node: flags (1|4 SPACE and Dirichlet validity, 2 Neumann
 * validity, 64 ghost children), ghost mark, parent index in the previous
 * level and 8 child indices in the next level (x bit first; -1 if absent).
 * psr_tree_metadata_copy copies one level into caller device buffers (any may
 * be NULL): keys[n], flags[n], ghost[n], parents[n], children[8 n]. */
typedef struct PsrTreeMetadata PsrTreeMetadata;
int psr_tree_metadata(PsrResult* result, void* stream, PsrTreeMetadata** out);
int psr_tree_metadata_levels(const PsrTreeMetadata* meta, int32_t* count);
int psr_tree_metadata_level(const PsrTreeMetadata* meta, int32_t index, int32_t* local_depth, int64_t* count);
int psr_tree_metadata_copy(const PsrTreeMetadata* meta, int32_t index, int64_t* keys, uint8_t* flags, uint8_t* ghost,
                           int32_t* parents, int32_t* children, void* stream);
void psr_tree_metadata_free(PsrTreeMetadata* meta);

/* First key of each level d = 0..depth+1 of sorted device keys (d<<48 | ...)
 * into the device array offsets (depth + 2 int32): the level_offsets of a
 * PsrField built by a caller from its own keys and coefficients, as the
 * reference's LevelField(keys, coefficients, ...) is (src/warp_psr/field.py:
 * 122-136; its find_key searches all keys, src/warp_psr/field.py:19-33).
 * Depth outside 0..15 fails with code 2 and the reference's message. */
int psr_field_level_offsets(const int64_t* keys, int64_t count, int depth, int32_t* offsets, void* stream);

/* The isovalue reduction _compute_isovalue (src/warp_psr/reconstruction.py:
 * 101-116) on caller device arrays: out (device, one value of dtype) =
 * sum(double(values[i] * weights[i])) / sum(double(weights[i])), one serial
 * double chain for n <= 1024, else double partials of 1024-element tiles
 * combined by halving trees, rounded to dtype.  Partial buffers come from the
 * library's pool. */
int psr_isovalue(int dtype, const void* values, const void* weights, int64_t n, void* out, void* stream);

/* psr_isovalue without any allocation: the same kernels and bits, with the
 * double partials in a file-scope device array of the library (1 MiB per GPU,
 * held until the process ends) instead of the pool.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  At most 2**26
 * values (more fails with code 2).  Every call on one GPU shares the array:
 * calls that overlap on one GPU race. */
int psr_isovalue_noalloc(int dtype, const void* values, const void* weights, int64_t n, void* out, void* stream);

/* Evaluates any PsrField view (psr_result_field) at m device queries with the
 * header-only psr_evaluate_field of warp_psr_amd_field.h -- the kernel-side
 * evaluate_field(field, point) of src/warp_psr/device_field.py:91-112 --
 * giving the same bits as psr_evaluate.  grads may be NULL. */
int psr_field_evaluate(const PsrField* field, const void* queries, int64_t m, void* values, void* grads, int32_t* valid,
                       void* stream);

const char* psr_last_error(void);
/* Frees the library's cached device blocks. */
void psr_release_cache(void);

#ifdef __cplusplus
}
#endif

#endif /* WARP_PSR_AMD_H */
