/*
 * wos_amd.h -- C ABI of the MI355X walk-on-spheres library (libwos_amd.so).
 *
 * The entry points are the operations the reference's only FFI, the pybind11
 * module wos_ext (src/python/scene.cu, src/python/wos.cu), exposes for the hot
 * path, plus the fused estimator entry points the Python layer needs.  Every
 * array argument is a device pointer (or any pointer HIP can address) that the
 * caller owns; the library copies scene geometry into buffers the scene
 * handle owns.  `stream` is a hipStream_t (NULL = default stream).  Every
 * function returns 0 on success and -1 on error; wos_last_error() then holds a
 * message.  Calls are asynchronous on `stream` unless stated otherwise.
 */
#ifndef WOS_AMD_H
#define WOS_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wos_scene wos_scene;

/* WoS(nwalks, nsteps=32, epsilon=1e-3, double_sided=False,
 *     prevent_fd_artifacts=False, use_IS_for_greens=True)
 * -- src/python/wos.cu:9-12, fields include/wos.cuh:147-152.
 * use_IS_for_greens selects how a walk with a source samples the interior
 * term: the radius from the tabulated CDF of r log(R/r) with the draw reused
 * inside its bin and pdf = pmf N / R / (2 pi) (include/wos.cuh:49-88, the
 * branch the reference's source compiles out at include/wos.cuh:121-134), or
 * uniformly in the ball (0; what the reference's CUDA walk executes). */
typedef struct {
    int   nwalks;
    int   nsteps;
    float epsilon;
    int   double_sided;
    int   prevent_fd_artifacts;
    int   use_IS_for_greens;
} wos_walk_params;

/* OursCUDA / Ours3D settings (wos/wos_grad.py:114-128, wos/wos_grad_3d.py:105-116) */
typedef struct {
    int   nwalks;
    int   nsteps;
    float epsilon;          /* primary walk */
    float epsilon2;         /* secondary walks of the normal derivative */
    float clamping;         /* angle clamp of the normal-derivative kernel */
    int   double_sided;
    int   control_variates; /* 2D only; 3D always uses the control variate */
} wos_grad_params;

/* A batch of closest-point records, dr::ClosestPointRecord<2/3>
 * (include/scene.cuh:18-58).  Any member may be NULL when not wanted.
 * p, n: n*dim floats (row-major); t: n floats (2D) or uv: 2n floats (3D). */
typedef struct {
    uint8_t *valid;
    float   *p;
    float   *n;
    float   *t;
    int     *prim_id;
    float   *contrib;
} wos_record;

/* Message of the last failed call on this thread ("" if none). */
const char *wos_last_error(void);

/* Scene(vertices, indices, values, use_bvh=False[, source_type, source_params])
 * -- src/python/scene.cu:29-33 and :69-71 (Scene3D), include/scene.cuh:330-342.
 * dim = 2 (segments, indices n_prims*2) or 3 (triangles, n_prims*3).
 * source_type: 0 none, 1 Gaussian (5 params), 2 sinusoid (5 params)
 * (include/scene.cuh:270-301), 3 Fourier sum with n_source_params = 4K:
 * (ax, ay, phase, coeff) per term, f = sum coeff sin(ax x + ay y + phase).
 * Builds vertex normals, the AABB and the LBVH on the device (always, as the
 * reference does, include/scene.cuh:437-452) and synchronises `stream`.
 * Rejects indices out of range, zero-length / zero-area primitives and scenes
 * of more than 2^26 segments (2D) or 2^25 triangles (3D). */
int wos_scene_create(int dim, const float *vertices, int n_vertices, const int *indices, int n_prims,
                     const float *values, int use_bvh, int source_type, const float *source_params,
                     int n_source_params, void *stream, wos_scene **out);
/* Launch tuning of every walk launch (no counterpart in the reference, whose
 * thrust launches are untuned): walks a wave claims per queue visit (chunk),
 * persistent workgroups per CU, BVH levels kept in LDS, and node visits per
 * lane between refills of the walk kernel (window; 0, the default, means 8
 * for 2D scenes and 16 for 3D).  Process-wide; chosen by the autotuner in the
 * Python package. */
int wos_set_tuning(int chunk, int wgs_per_cu, int top_levels, int window);
int wos_get_tuning(int *chunk, int *wgs_per_cu, int *top_levels, int *window);
/* Whether walk launches visit the query points in Morton order within the
 * scene's box (default 1) or in input order (0).  Results are identical
 * either way; the order only changes which BVH nodes the walks in flight
 * share in L2.  Process-wide; no counterpart in the reference. */
int wos_set_spatial_order(int enable);
/* Scene destructor (the reference frees its thrust buffers with the object). */
int wos_scene_destroy(wos_scene *scene);
/* Scene.values / Scene.source_params setters (def_readwrite in
 * src/python/scene.cu:38, :47): replace without rebuilding the BVH. */
int wos_scene_set_values(wos_scene *scene, const float *values, void *stream);
int wos_scene_set_source(wos_scene *scene, int source_type, const float *source_params, int n_source_params,
                         void *stream);
/* Scene.use_bvh (src/python/scene.cu:40) */
int wos_scene_set_use_bvh(wos_scene *scene, int use_bvh);
/* Read-only fields: Scene.normals (device copy, n_vertices*dim floats),
 * Scene.aabb (host floats min[3], max[3]), tree depth, dimension. */
int wos_scene_get_normals(const wos_scene *scene, float *out, void *stream);
int wos_scene_get_info(const wos_scene *scene, float *aabb_min3, float *aabb_max3, int *dim, int *n_vertices,
                       int *n_prims, int *max_depth);
/* Copies the LBVH topology (Morton-sorted primitive ids, n_prims; Morton codes,
 * n_prims; children of the n_prims-1 internal nodes, leaf j encoded as ~j) --
 * the structure lbvh::bvh builds at include/scene.cuh:452. */
int wos_scene_get_bvh(const wos_scene *scene, int *sorted_prim, uint32_t *sorted_codes, int *children,
                      void *stream);
/* Near-boundary candidate grid of a 2D scene: an acceleration structure with
 * no counterpart in the reference, answering the same closest-primitive
 * queries as its LBVH (include/scene.cuh:130-146) with identical results.
 * Cells of side cell_factor x the mean segment length that lie within
 * band_cells cells of the boundary list every segment that can be closest to
 * a point of the cell (lists longer than max_cand are dropped); queries in
 * other cells use the tree.  wos_scene_create builds it with (1.0, 16, 48);
 * this call rebuilds it, or removes it when cell_factor <= 0.  Synchronises
 * the device. */
int wos_scene_set_grid(wos_scene *scene, float cell_factor, int band_cells, int max_cand, void *stream);
/* Grid shape (nx = ny = 0 when there is none), cell side and list entries. */
int wos_scene_get_grid(const wos_scene *scene, int *nx, int *ny, float *cell, unsigned int *n_cand);
/* Start-node grid (no counterpart in the reference; results of every query
 * are unchanged): each square (2D) or cubic (3D) cell over the bounding box
 * records the deepest LBVH node below which every primitive that can be
 * closest to a point of the cell lies, and traversals from that cell start
 * there instead of at the root (in 2D a cell may record two nodes).
 * wos_scene_create builds it in 2D with two cells per mean segment length
 * along the longer side (at most 2048), and none in 3D (where it measured no
 * faster).  resolution: cells along the longer side, 0 removes the grid, -1
 * chooses as wos_scene_create does in 2D and 96 in 3D.  Synchronises the
 * device. */
int wos_scene_set_start_grid(wos_scene *scene, int resolution, void *stream);
int wos_scene_get_start_grid(const wos_scene *scene, int *nx, int *ny, int *nz);

/* Scene.closest_point_preliminary(p) -> Int (include/scene.cuh:497-537).
 * Ties: equal squared distance keeps the smallest primitive index.
 * out_d2 (nullable): squared distance. */
int wos_closest_point_index(const wos_scene *scene, const float *p, int n, int *out, float *out_d2, void *stream);
/* SceneDevice::sdf (include/scene.cuh:148-178) over a batch. */
int wos_sdf(const wos_scene *scene, const float *p, int n, float *out, void *stream);
/* Scene.closest_point(p) -> ClosestPointRecord (include/scene.cuh:539-657) */
int wos_closest_point(const wos_scene *scene, const float *p, int n, const wos_record *out, void *stream);
/* Scene.largest_inscribed_ball(its, epsilon=1e-3) (include/scene.cuh:659-704);
 * reads its.valid (nullable = all valid), its.p, its.n. */
int wos_largest_inscribed_ball(const wos_scene *scene, const wos_record *its, int n, float epsilon, float *out,
                               void *stream);
/* Scene.dirichlet(cRec) (include/scene.cuh:455-468); reads valid (nullable),
 * prim_id, t (2D) / uv (3D). */
int wos_dirichlet(const wos_scene *scene, const wos_record *rec, int n, float *out, void *stream);

/* WoS.single_walk(p, scene, sampler) -> ClosestPointRecord
 * (src/python/wos.cu:14, include/wos.cuh:374-533).  state/inc: the PCG32
 * sampler of each walk, read and written back (include/wos.cuh:438-439). */
int wos_single_walk(const wos_scene *scene, const wos_walk_params *params, const float *p, int n,
                    uint64_t *state, uint64_t *inc, const wos_record *out, void *stream);

/* Solver.walk_reduce(p, scene, seed) (wos/solver.py:49-65) over the CUDA walk
 * value (wos/wos.py:46-54, wos/wos_with_source.py:257-265): W = nwalks
 * walks per point, walk k = i*W + w keyed by TEA(seed, i) and k.  One
 * persistent launch; outputs (all nullable except as noted):
 *   mean     n_points floats       (needs walk_value)
 *   mean64   n_points doubles      per-point FP64 mean  (needs walk_value)
 *   var64    n_points doubles      per-point FP64 sample variance (needs walk_value)
 *   walk_value  n_points*W floats  per-walk value
 *   rec_prim    n_points*W ints    exit primitive (-1 when the walk is invalid)
 *   rec_t       n_points*W floats (2D t) or 2*n_points*W (3D uv)
 *   basis       K*n_points doubles (coefficient-major, basis[k*n_points + i]):
 *               sum over the W walks of point i of d value / d coeff_k of a
 *               Fourier source (0 for invalid walks).  Accumulated in 64-bit
 *               fixed point (units of 2^-e, e from nwalks and nsteps), so the
 *               sums are bit-reproducible whatever order the walks end in.
 * chunk: walks a wave claims per queue visit (<= 0: default). */
int wos_walk_reduce(const wos_scene *scene, const wos_walk_params *params, const float *points, int n_points,
                    int seed, int chunk, float *mean, double *mean64, double *var64, float *walk_value,
                    int *rec_prim, float *rec_t, double *basis, void *stream);

/* wos_walk_reduce for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, checks and per-walk
 * results as wos_walk_reduce; the queue head, start records and traversal
 * spill live in one static 64 MiB device array per GPU (held until the
 * process ends) instead of library allocations, so the call allocates
 * nothing.  Limit: 8 bytes per point plus one workgroup's traversal spill
 * must fit the array (about 8 million points); larger calls fail with an
 * error.  The points are visited in input order.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race; calls on one stream run in order. */
int wos_walk_reduce_noalloc(const wos_scene *scene, const wos_walk_params *params, const float *points,
                            int n_points, int seed, int chunk, float *mean, double *mean64, double *var64,
                            float *walk_value, int *rec_prim, float *rec_t, double *basis, void *stream);

/* Instrumented walk_reduce for measurement: same launch as wos_walk_reduce
 * (with the Fourier basis outputs when with_basis), and adds into counts
 * (8 device uint64, zeroed by the caller): AABB tests, primitive tests,
 * closest-primitive queries, walk steps, then the lane activity of the
 * persistent kernel: lanes taking a step summed over the waves' step
 * phases, the number of those phases, querying lanes summed over the waves'
 * node-visit iterations, the number of those iterations.  Slower than the
 * plain launch. */
int wos_walk_reduce_counts(const wos_scene *scene, const wos_walk_params *params, const float *points, int n_points,
                           int seed, int with_basis, float *walk_value, unsigned long long *counts, void *stream);

/* Reverse mode of walk_reduce with respect to the vertex values and the
 * Fourier coefficients: the gradient Dr.Jit's AD takes through
 * Scene.dirichlet's gathers (include/scene.cuh:455-468) and through the
 * contrib sum (include/wos.cuh:136-138), computed from the exit records of a
 * walk_reduce call with the same settings (basis: its per-point sums).
 * grad_out: dL/d mean (n_points).  Adds into grad_values (n_vertices,
 * nullable) and grad_coeffs (K, nullable); the coefficient gradients are
 * reduced in a fixed order and so are bit-reproducible. */
int wos_adjoint_values(const wos_scene *scene, int n_points, int nwalks, const int *rec_prim, const float *rec_t,
                       const double *basis, const float *grad_out, float *grad_values, float *grad_coeffs,
                       void *stream);

/* Reverse mode of walk_reduce with respect to every source parameter -- what
 * Dr.Jit's AD takes through Scene.source_function (include/scene.cuh:278-301,
 * wos/scene.py:359-381) for the Gaussian (5), sinusoid (5) and Fourier (4K:
 * ax, ay, phase, coeff per term) sources.  Each walk is re-run from its key
 * (the streams of a walk_reduce call with the same seed and settings), so
 * nothing is stored.  grad_out: dL/d mean (n_points).  Adds into grad_source
 * (n_source_params floats); a scene without a source adds nothing.  The
 * walks' terms are summed over fixed slices of walks in a fixed order, so the
 * result is bit-reproducible; allocates 4 bytes per parameter per slice
 * (at most 2^20 slices) on the stream. */
int wos_adjoint_source(const wos_scene *scene, const wos_walk_params *params, const float *points, int n_points,
                       int seed, const float *grad_out, float *grad_source, void *stream);

/* Shape gradient: CustomSolver backward (wos/solver.py:109-114) with
 * bwd_solver = OursCUDA (wos/wos_grad.py:130-151) in 2D or Ours3D
 * (wos/wos_grad_3d.py:117-137) in 3D, walks recomputed from (seed, point,
 * walk).  Adds dL/dvertices into grad_vertices (n_vertices*dim, nullable),
 * summed in 64-bit fixed point (bit-reproducible; allocates a 16*(dim-1)-byte
 * staging record per sample on the stream).
 * sample_grad_u (nullable): per-sample spatial gradient estimate,
 * n_points*nwalks*dim floats (0 for samples that do not contribute). */
int wos_shape_gradient(const wos_scene *scene, const wos_grad_params *params, const float *points, int n_points,
                       int seed, const float *grad_out, float *grad_vertices, float *sample_grad_u, void *stream);

/* Instrumented wos_shape_gradient for measurement (no gradient output): adds
 * into counts (4 device uint64, zeroed by the caller) the bound tests,
 * primitive tests, closest-primitive queries and walk steps of the samples
 * (primary walks, LIB queries, secondary walks).  Slower than the plain
 * launch; no counterpart in the reference. */
int wos_shape_gradient_counts(const wos_scene *scene, const wos_grad_params *params, const float *points,
                              int n_points, int seed, unsigned long long *counts, void *stream);

/* Forward mode of the same integrand: the derivative of each point's
 * OursCUDA / Ours3D estimate along a deformation that moves vertex v with
 * velocity[v] (n_vertices*dim) -- what diff_solve/runner.py:75-109 (renderD)
 * obtains with dr.forward_to on the image.  Per sample -grad u . V(exit),
 * V interpolated like the exit point; per point the mean over the walks
 * (out_mean f32, out_mean64 f64, out_var64 f64 variance of one sample; each
 * nullable).  The samples are those of wos_shape_gradient with the same
 * seed, so sum_i g_i out_mean[i] equals grad_vertices . velocity for
 * grad_out = g up to float rounding. */
int wos_shape_derivative(const wos_scene *scene, const wos_grad_params *params, const float *points, int n_points,
                         int seed, const float *velocity, float *out_mean, double *out_mean64, double *out_var64,
                         void *stream);

/* ---- batches of scenes ----------------------------------------------------
 * No counterpart in the reference, which solves one scene per call: a batch
 * runs Solver.walk_reduce (wos/solver.py:49-65) for many (scene, settings,
 * points, seed) items in one persistent launch per dimension.  Each item's
 * walks use exactly the streams of a single wos_walk_reduce call with the
 * same seed.  wos_batch_create uploads the item table and allocates every
 * buffer a run needs (synchronous); wos_batch_run and wos_batch_adjoint
 * allocate nothing and can be captured into one graph; runs of one batch on
 * concurrent streams race.  The item table holds the scenes' device buffers:
 * after wos_scene_set_source, wos_scene_set_grid, wos_scene_set_start_grid
 * or a change of wos_scene_set_use_bvh on an item's scene, run and adjoint
 * fail with an error until the batch is created again (new vertex values,
 * wos_scene_set_values, are read by the next run). */
typedef struct wos_batch wos_batch;
typedef struct {
    const wos_scene *scene;
    wos_walk_params  params;
    const float     *points;
    int              n_points;
    int              seed;
    float           *mean;        /* nullable */
    double          *mean64;      /* nullable */
    double          *var64;       /* nullable */
    float           *walk_value;  /* required: n_points * nwalks */
    int             *rec_prim;    /* nullable; required by the reverse pass */
    float           *rec_t;       /* nullable; required by the reverse pass */
    double          *basis;       /* nullable: Fourier sources only; K * n_points per-point sums */
    /* reverse pass (wos_batch_adjoint), all nullable: dL/d mean per point,
     * read when the pass runs; dL/d vertex values (n_vertices) and dL/d
     * Fourier coefficients (K), added into (the caller zeroes them) */
    const float     *grad_out;
    float           *grad_values;
    float           *grad_coeffs;
    /* shape gradient (wos_batch_shape_gradient), nullable: dL/d vertex
     * positions (n_vertices * dim), added into, by the OursCUDA / Ours3D
     * samples of grad_params (the bwd_solver of CustomSolver) with this
     * item's seed and grad_out; needs the scene's BVH */
    wos_grad_params  grad_params;
    float           *grad_vertices;
} wos_batch_item;
int wos_batch_create(const wos_batch_item *items, int n_items, wos_batch **out);
/* Primal pass: walk_reduce of every item (wos/solver.py:49-65). */
int wos_batch_run(wos_batch *batch, void *stream);
/* Reverse pass of every item with grad_values or grad_coeffs, from the exit
 * records of the last run: wos_adjoint_values per item (the transpose of
 * include/scene.cuh:251-264 and of the contrib sum, include/wos.cuh:136-138). */
int wos_batch_adjoint(wos_batch *batch, void *stream);
/* Shape-gradient pass of every item with grad_vertices: wos_shape_gradient
 * per item (CustomSolver backward, wos/solver.py:109-114, with OursCUDA /
 * Ours3D, wos/wos_grad.py:130-151, wos/wos_grad_3d.py:117-137), all of a
 * dimension's items in one persistent launch.  Reads each item's grad_out;
 * the samples are those of a single-scene wos_shape_gradient with the same
 * seed and settings, and both sum in the same fixed point, so each item's
 * grad_vertices equals the single-scene result bit for bit. */
int wos_batch_shape_gradient(wos_batch *batch, void *stream);
int wos_batch_destroy(wos_batch *batch);

/* ---- testing hook -----------------------------------------------------------
 * No counterpart in the reference: the next n requests for a launch's
 * traversal spill fail (without allocating), so tests can check that every
 * entry point returns an error instead of launching. */
int wos_debug_fail_workspace(int n);

#ifdef __cplusplus
}
#endif
#endif
