// wos_build.h -- internal interface between the C ABI (wos_capi.hip), the
// scene builder (wos_build.hip) and the solver kernels (wos_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <cmath>

#include "wos_device.h"

namespace wos {

// Records a formatted message for wos_last_error() (thread-local).
void set_error(const char *fmt, ...);

struct BuildInputs {
    int dim, n_vertices, n_prims;
    const float *vtx;  // device, owned by the scene
    const int *idx;    // device, owned by the scene
};

struct BuildOutputs {
    float4 *nodes;          // (n_prims - 1) * 4 float4 (2D: capsule bounds with leaf segments inline)
    float4 *leaves;         // n_prims * (dim == 2 ? 4 : 6) float4: Morton order, then primitive order
    float *normals;         // n_vertices * dim
    int *sorted_prim;       // n_prims
    uint32_t *sorted_codes; // n_prims
    int *children;          // 2 * (n_prims - 1), nullable
    float amin[3], amax[3];
    int max_depth;
    int top_counts[17];     // internal nodes with depth < k, k = 0..16 (depth-major order)
    uint32_t error_bits;    // 1: degenerate primitive, 2: index out of range
};

int build_scene_gpu(BuildInputs &in, hipStream_t stream, BuildOutputs &out);

// Near-boundary candidate grid of a 2D scene (see grid_closest).
struct GridSettings {
    float cell_factor = 1.0f;   // cell side / mean segment length; <= 0 disables the grid
    int band_cells = 16;        // cells within this many cell sides of the boundary get lists
    int max_cand = 48;          // longer lists are dropped (those cells ask the tree)
    long max_cells = 1L << 23;  // the cell side grows to respect this
};
struct GridInputs {
    int n_prims, max_depth;
    const float *vtx;
    const int *idx;
    const float4 *nodes, *leaves;
    const int *sorted_prim;
    float amin[2], amax[2];
    float tol;
};
struct GridOutputs {
    uint32_t *start = nullptr;   // nx * ny + 1 (hipMalloc, owned by the caller afterwards)
    int *cand = nullptr;         // n_cand leaf indices (hipMalloc)
    uint32_t n_cand = 0;
    float ox = 0, oy = 0, inv_cell = 0, cell = 0;
    int nx = 0, ny = 0;
};
int build_grid_2d(const GridInputs &in, const GridSettings &gs, hipStream_t stream, GridOutputs &out);

// Start-node grid (see start_nodes_of): `resolution` cells along the longest
// side of the bounding box; 0 builds none; kStartGridAuto chooses it: in 2D
// two cells per mean segment length (clamped to [64, kStartGridAutoMax2]:
// finer cells start queries deeper, measured down to half a segment on the
// clover), in 3D kStartGridAuto3.
constexpr int kStartGridAuto = -1;
constexpr int kStartGridAutoMax2 = 2048;
constexpr int kStartGridAuto3 = 96;
constexpr int kStartGridResolution2 = kStartGridAuto;
constexpr int kStartGridResolution3 = 0;   // measured no faster on the displaced spheres; opt-in
struct StartGridOutputs {
    // per cell {first node, second or -1, key bits} (hipMalloc, owned by the
    // caller): packed = 1, 8 bytes {first | second << 16 (0: none), key}
    // when every internal node id fits 16 bits (2D), else int4 {first, second,
    // key, 0}
    int *start = nullptr;
    float ox = 0, oy = 0, oz = 0, inv_cell = 0;
    int nx = 0, ny = 0, nz = 0;
    int packed = 0;
};
int build_start_grid_2d(const GridInputs &in, int resolution, hipStream_t stream, StartGridOutputs &out);
struct GridInputs3 {
    int n_prims, max_depth;
    const float4 *nodes, *leaves;
    float amin[3], amax[3];
    float tol;
};
int build_start_grid_3d(const GridInputs3 &in, int resolution, hipStream_t stream, StartGridOutputs &out);

// Permutation of n points into Morton order within [amin, amax] (device).
int sort_points_morton(const float *points, int n, int D, const float *amin, const float *amax, int *order,
                       hipStream_t stream);

// ---- solver launches (wos_kernels.hip) -------------------------------------
struct Workspace;   // per-device scratch owned by the C ABI layer

struct WalkSettings {
    int nwalks, nsteps;
    float epsilon;
    int double_sided, prevent_fd_artifacts;
    int use_is;   // interior samples of a source from the tabulated Green's-function CDF
};

// Host copy of the tabulated Green's-function CDF (kGreenCdfSize floats,
// entry i = x^2 (1 - 2 ln x) at x = (i + 1) / kGreenCdfSize, evaluated in
// double and rounded to float).
const float *green_cdf_host();

struct WalkOutputs {
    float *walk_value;   // P*W, nullable
    int *rec_prim;       // P*W, nullable
    float *rec_t;        // P*W*(dim-1), nullable
    double *basis;       // K*P (coefficient-major) per-point sums over the walks, nullable
    float *mean;         // P, nullable
    double *mean64;      // P, nullable
    double *var64;       // P, nullable
};

// Per-point Fourier basis sums are accumulated as 64-bit integers in units
// of 2^-e, so their value does not depend on the order in which walks end.
// One interior sample adds b G w with |b| <= 1 and |G w| <= R^2 (1 + ln(R /
// 1e-4)) for a ball of radius R <= sqrt(1e5 / pi) (no sample is taken for a
// larger ball: the reference's pdf > 1e-5 guard); a walk takes at most
// nsteps samples, so the sum over nwalks walks stays below 2^61 / 2^e.
// The same exponent is restated in the oracle.
inline int basis_exponent(int nwalks, int nsteps) {
    const double r = 178.5;                  // > sqrt(1e5 / pi) = 178.41
    const double per_sample = r * r * 16.0;  // 1 + ln(178.5 / 1e-4) = 15.4 < 16
    const double bound = (double)(nwalks > 1 ? nwalks : 1) * (double)(nsteps > 1 ? nsteps : 1) * per_sample;
    int e = 0;
    (void)frexp(bound, &e);   // bound < 2^e
    return 61 - e;
}

// Launch tuning shared by all walk launches (set through wos_set_tuning).
struct Tuning {
    int chunk = 128;        // walks a wave claims per queue visit
    int wgs_per_cu = 8;     // persistent workgroups per CU
    int top_levels = 6;     // BVH levels kept in LDS
    int window = 0;         // node visits per lane between refills; 0: 8 in 2D, 16 in 3D
    int spatial_order = 1;  // visit query points in Morton order (0: input order)
};
extern Tuning g_tuning;

// 3D queries are longer (about 50 node visits against 12 in 2D), so a wider
// window pays for its idle lanes by running the step phase fewer times.
inline int window_for(int dim) { return g_tuning.window > 0 ? g_tuning.window : (dim == 2 ? 8 : 16); }

// Traversal spill of `stream` (grown, never shrunk) for `lanes` lanes: 0 with
// *out = null when the tree fits the LDS stack, -1 (error set) when it cannot
// be allocated.  Launches on one stream share it; other streams get their own.
int spill_buffer(const SceneView &s, long lanes, long *stride, int2 **out, hipStream_t stream);
// Queue counter of `stream` (null, error set, when it cannot be allocated).
unsigned long long *launch_counter(hipStream_t stream);

int launch_closest_index(const SceneView &s, const float *p, int n, int *out_idx, float *out_d2, hipStream_t st);
int launch_sdf(const SceneView &s, const float *p, int n, float *out, hipStream_t st);
int launch_closest_point(const SceneView &s, const float *p, int n, uint8_t *valid, float *op, float *on,
                         float *ot, int *oprim, float *ocontrib, hipStream_t st);
int launch_lib(const SceneView &s, const uint8_t *valid, const float *p, const float *nrm, int n, float eps,
               float *out, hipStream_t st);
int launch_dirichlet(const SceneView &s, const uint8_t *valid, const int *prim, const float *t, int n,
                     float *out, hipStream_t st);
int launch_single_walk(const SceneView &s, const WalkSettings &w, const float *p, int n, uint64_t *state,
                       uint64_t *inc, uint8_t *valid, float *op, float *on, float *ot, int *oprim,
                       float *ocontrib, hipStream_t st);
int launch_walk_reduce(const SceneView &s, const WalkSettings &w, const float *points, int n_points,
                       int32_t seed, int chunk, const WalkOutputs &o, hipStream_t st);
int launch_walk_reduce_noalloc(const SceneView &s, const WalkSettings &w, const float *points, int n_points,
                               int32_t seed, int chunk, const WalkOutputs &o, hipStream_t st);
int launch_adjoint_values(const SceneView &s, int n_points, int nwalks, const int *rec_prim, const float *rec_t,
                          const double *basis, const float *grad_out, float *grad_values, float *grad_coeffs,
                          hipStream_t st);

int launch_source_adjoint(const SceneView &s, const WalkSettings &w, const float *points, int n_points,
                          int32_t seed, const float *grad_out, float *grad_src, hipStream_t st);

struct GradSettings {
    int nwalks, nsteps;
    float epsilon, epsilon2, clamping;
    int double_sided, control_variates;
};
int launch_shape_gradient(const SceneView &s, const GradSettings &g, const float *points, int n_points,
                          int32_t seed, const float *grad_out, float *grad_vertices, float *sample_grad_u,
                          hipStream_t st, const float *velocity = nullptr, float *sample_jvp = nullptr);
int launch_shape_derivative(const SceneView &s, const GradSettings &g, const float *points, int n_points,
                            int32_t seed, const float *velocity, float *mean, double *mean64, double *var64,
                            hipStream_t st);

}  // namespace wos
