/*
 * wos_oracle.h -- CPU restatement of the zihay/diff-wos walk-on-spheres hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product (diff-wos_amd/) never links, imports or calls it.
 *
 * Every function names the reference file:line it restates.  Floating-point
 * work follows the same canonical operation order as the device code
 * (explicit fmaf, no other contraction, IEEE division and sqrt, the
 * polynomial sin/cos/log/exp below), so most per-walk results agree with the
 * GPU bit for bit; the statistical bar (3 sigma) is what the tests require.
 */
#ifndef WOS_ORACLE_H
#define WOS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers ---------------------------------------------------- */
typedef struct { uint64_t state, inc; } or_pcg32;
/* include/sampler2.cuh:20-28 */
void     or_pcg32_seed(or_pcg32 *r, uint64_t index, uint64_t initstate, uint64_t initseq);
/* include/sampler2.cuh:30-37 */
uint32_t or_pcg32_next_u32(or_pcg32 *r);
/* include/sampler2.cuh:39-44 */
float    or_pcg32_next_1d(or_pcg32 *r);
/* Dr.Jit PCG32.next_float64 (used by the wos Python modules): two 32-bit draws, low word first */
double   or_pcg32_next_f64(or_pcg32 *r);
/* wos/utils.py:250-256 (int32 arithmetic, arithmetic right shift) */
void     or_tea(int32_t v0, int32_t v1, int rounds, int32_t *o0, int32_t *o1);
/* batch helpers for tests */
void     or_pcg32_stream(uint64_t index, uint64_t initstate, uint64_t initseq, int n, uint32_t *out);

/* ---- elementary functions (own polynomials; pinned against libm in tests) */
void  or_sincos_turns(float u, float *s, float *c);   /* sin/cos(2*pi*u) */
float or_sinf(float x);
float or_logf(float x);
float or_expf(float x);

/* ---- scenes ------------------------------------------------------------- */
typedef struct {
    int dim;                 /* 2 or 3 */
    int n_vertices, n_prims;
    const float *vertices;   /* n_vertices * dim */
    const int   *indices;    /* n_prims * dim */
    const float *values;     /* n_vertices */
    float       *normals;    /* n_vertices * dim, filled by or_scene_init (2D only) */
    float aabb_min[3], aabb_max[3];
    int source_type;         /* 0 none, 1 gaussian, 2 sinusoid, 3 fourier */
    const float *source_params;
    int n_source_params;
    /* optional LBVH (or_bvh_create); when set, every closest-primitive query
     * uses it, with the brute-force scan's result (same tie rule) */
    const struct or_bvh *bvh;
} or_scene;

typedef struct {
    int   valid;
    float p[3], n[3];
    float t;        /* 2D segment parameter */
    float uv[2];    /* 3D barycentrics of (b, c) */
    float d;
    int   prim_id;
    float contrib;
} or_record;

typedef struct {
    int   nwalks, nsteps;
    float epsilon;
    int   double_sided, prevent_fd_artifacts;
    int   use_IS_for_greens;   /* interior samples from the tabulated Green's-function CDF */
} or_walk_params;

/* Tabulated CDF of r log(1/r) on [0, 1] (include/tabulated_G_cdf.cuh:5-7,
 * as its closed form): *n entries. */
const float *or_green_cdf(int *n);
/* include/wos.cuh:49-75: radius in [0, R] and its radial pdf; bin (nullable) */
float or_sample_green_radius(float u, float R, float *pdf, int *bin);
/* include/wos.cuh:115-139 (uniform) / 77-88 (is): offset, G and weight of one
 * interior sample of the disk of radius d; 0 when the reference adds nothing */
int   or_interior_sample(float u0, float u1, float d, int is, float *dx, float *dy, float *G, float *w);

/* include/scene.cuh:368-380 (2D vertex normals) and 417-435 (AABB) */
void  or_scene_init(or_scene *s);
/* include/scene.cuh:102-128 brute force; ties -> smallest index (squared-distance key) */
int   or_closest_index(const or_scene *s, const float *p, float *out_d2);
/* include/scene.cuh:148-178 */
float or_sdf(const or_scene *s, const float *p);
/* include/scene.cuh:180-223 */
void  or_closest_point(const or_scene *s, const float *p, or_record *rec);
/* include/scene.cuh:225-239 */
float or_largest_inscribed_ball(const or_scene *s, const or_record *rec, float epsilon);
/* include/scene.cuh:251-264 */
float or_dirichlet(const or_scene *s, const or_record *rec);
/* include/scene.cuh:278-301 plus the Fourier type (sum_k c_k sin(ax x + ay y + ph)) */
float or_source(const or_scene *s, float x, float y);

/* include/wos.cuh:97-145.  basis (nullable) accumulates d contrib / d c_k for
 * the Fourier source, length n_source_params/4. */
void  or_walk(const or_scene *s, const or_walk_params *w, const float *p,
              or_pcg32 *rng, or_record *rec, float *basis);
/* wos/wos.py:46-54, wos/wos_with_source.py:257-265: per-walk value */
float or_walk_value(const or_scene *s, const or_record *rec);

/* wos/solver.py:49-65 with FP64 accumulation.  Per point: mean, variance of
 * the walk values.  walk_val (nullable): per-walk values, P*W.  rec_prim /
 * rec_t (nullable): per-walk exit primitive (-1 invalid) and t (2D) / uv (3D,
 * 2 floats per walk).  basis (nullable): per point mean of d u / d c_k (P*K).
 * nthreads <= 0: one thread. */
/* Fixed-point exponent of the per-point Fourier basis sums (see the C file). */
int   or_basis_exponent(int nwalks, int nsteps);
/* basis (nullable): n_points x K, per point the sum over its walks of
 * d value / d coeff_k, summed as 64-bit integers in units of
 * 2^-or_basis_exponent(nwalks, nsteps). */
void  or_walk_reduce(const or_scene *s, const or_walk_params *w, const float *points,
                     int n_points, int64_t seed, int nthreads,
                     double *mean, double *var, float *walk_val,
                     int *rec_prim, float *rec_t, double *basis);

/* Adjoint of the walk value with respect to vertex values and Fourier source
 * coefficients (Dr.Jit reverse mode through include/scene.cuh:251-264 and the
 * contrib sum at include/wos.cuh:136-138).  grad_out: per point dL/du. */
void  or_adjoint_values(const or_scene *s, int n_points, int nwalks,
                        const int *rec_prim, const float *rec_t,
                        const double *grad_out, double *grad_values);

typedef struct {
    int   nwalks, nsteps;
    float epsilon, epsilon2, clamping;
    int   double_sided, control_variates;
} or_grad_params;

/* wos/wos_grad.py:130-151 (OursCUDA) and wos/wos_grad_3d.py:117-137 (Ours3D),
 * run through Solver.walk_reduce (wos/solver.py:49-65) and transposed:
 * grad_vertices[v] += grad_out[i]/W * d res / d v  (FP64). */
void  or_shape_gradient(const or_scene *s, const or_grad_params *g, const float *points,
                        int n_points, int64_t seed, const double *grad_out,
                        int nthreads, double *grad_vertices, double *sample_grad_u);


/* ---- batch helpers (ctypes convenience for the tests) -------------------- */
/* fn: 0 sin(2 pi u), 1 cos(2 pi u), 2 sin, 3 log, 4 exp */
void  or_batch_elementary(int fn, const float *in, int n, float *out);
void  or_batch_closest(const or_scene *s, const float *p, int n, int *idx, float *d2);
void  or_batch_sdf(const or_scene *s, const float *p, int n, float *out);
/* rec arrays: valid n, p n*dim, nrm n*dim, t n (2D) / 2n (3D), prim n */
void  or_batch_closest_point(const or_scene *s, const float *p, int n, int *valid, float *rp, float *rn,
                             float *rt, int *prim);
/* Scene.dirichlet over records (valid nullable; t: n (2D) or 2n (3D));
 * 0 for invalid records and primitives out of range */
void  or_batch_dirichlet(const or_scene *s, const int *valid, const int *prim, const float *t, int n, float *out);
void  or_batch_lib(const or_scene *s, const int *valid, const float *p, const float *nrm, int n, float eps,
                   float *out);
/* single walks from explicit PCG32 states (WoS::singleWalk semantics):
 * state/inc advanced in place; outputs as or_batch_closest_point plus contrib */
void  or_batch_single_walk(const or_scene *s, const or_walk_params *w, const float *p, int n, uint64_t *state,
                           uint64_t *inc, int *valid, float *rp, float *rn, float *rt, int *prim, float *contrib);

/* ---- LBVH restatement (used to pin the device build) -------------------- */
/* Morton codes of primitive centroids, stable sort by (code, index), Karras
 * 2012 internal nodes.  Outputs: sorted primitive ids (n), internal node
 * children (2*(n-1); leaf j encoded as ~j), parent of internal nodes (n-1) and
 * leaves (n), max depth.  Returns 0 on success. */
int   or_lbvh_build(const or_scene *s, uint32_t *codes_sorted, int *prim_sorted,
                    int *children, int *parent_internal, int *parent_leaf, int *max_depth);
/* Morton code of a point in [0,1]^dim (2D: 16 bits per axis, 3D: 10 bits) */
uint32_t or_morton(int dim, const float *unit);

/* CPU LBVH over the tree of or_lbvh_build with per-node boxes, for the
 * closest-primitive query of include/scene.cuh:130-146: nearest-first
 * traversal that prunes a box only when its distance exceeds the best
 * distance so far plus a scale-relative margin, so the result is the
 * brute-force scan's (smallest squared distance, then smallest index).
 * Assign the result to or_scene.bvh; NULL on failure. */
typedef struct or_bvh or_bvh;
or_bvh *or_bvh_create(const or_scene *s);
void    or_bvh_destroy(or_bvh *b);
/* queries through the tree (whatever s->bvh holds) */
int     or_bvh_closest(const or_scene *s, const or_bvh *b, const float *p, float *out_d2);

#ifdef __cplusplus
}
#endif
#endif
