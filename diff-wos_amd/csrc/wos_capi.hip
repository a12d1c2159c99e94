// wos_capi.hip -- the C ABI declared in include/wos_amd.h: scene ownership,
// argument checks, workspace, and dispatch to the kernels.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>

#include <algorithm>
#include <map>
#include <mutex>

#include "../../include/wos_amd.h"
#include "wos_build.h"

struct wos_scene {
    int dim = 0, n_vertices = 0, n_prims = 0, use_bvh = 0;
    int source_type = 0, n_src = 0;
    int device = 0;
    float *vtx = nullptr, *val = nullptr, *nrm = nullptr, *src = nullptr;
    int *idx = nullptr, *sorted_prim = nullptr, *children = nullptr;
    uint32_t *sorted_codes = nullptr;
    float4 *nodes = nullptr, *leaves = nullptr;
    float amin[3] = {0, 0, 0}, amax[3] = {0, 0, 0};
    float tol = 0.0f;
    int max_depth = 0;
    int top_counts[17] = {0};
    wos::GridOutputs grid;        // 2D near-boundary candidate grid (start == null: none)
    wos::StartGridOutputs sgrid;  // 2D start-node grid (start == null: none)
    // bumped by every setter that frees a buffer or changes what a walk
    // reads besides the vertex values (batches check it before each run)
    unsigned long long generation = 0;
};

namespace wos {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// ---- workspace: traversal spill and queue counter, one per (device, stream).
// Launches on one stream run in order, so they can share one spill buffer and
// one counter; launches on different streams get different ones and never
// overwrite each other's traversal stacks.
namespace {
struct StreamWorkspace {
    int2 *spill = nullptr;
    size_t spill_bytes = 0;
    unsigned long long *counter = nullptr;
};
std::map<std::pair<int, hipStream_t>, StreamWorkspace> g_ws;
std::mutex g_ws_lock;

StreamWorkspace &workspace_of(hipStream_t stream) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    return g_ws[std::make_pair(dev, stream)];
}
}  // namespace

// Failure injection for the workspace error paths (set through
// wos_debug_fail_workspace): the next n spill requests fail without
// allocating, so tests can drive every launcher's error return.
static int g_fail_workspace_allocs = 0;

int spill_buffer(const SceneView &s, long lanes, long *stride, int2 **out, hipStream_t stream) {
    *stride = lanes;
    *out = nullptr;
    {
        std::lock_guard<std::mutex> g(g_ws_lock);
        if (g_fail_workspace_allocs > 0) {
            --g_fail_workspace_allocs;
            set_error("traversal spill allocation failed (injected)");
            return -1;
        }
    }
    long deep = (long)s.max_depth + 2 - kLdsStack;
    if (!s.use_bvh || deep <= 0) return 0;   // the LDS part holds every entry
    size_t need = (size_t)deep * (size_t)lanes * sizeof(int2);
    std::lock_guard<std::mutex> g(g_ws_lock);
    StreamWorkspace &w = workspace_of(stream);
    if (w.spill_bytes < need) {
        // the old buffer may still be read by kernels queued on this stream
        if (w.spill) {
            if (hipStreamSynchronize(stream) != hipSuccess) {
                set_error("cannot synchronise the stream before growing the traversal spill");
                return -1;
            }
            (void)hipFree(w.spill);
        }
        w.spill = nullptr;
        w.spill_bytes = 0;
        if (hipMalloc((void **)&w.spill, need) != hipSuccess) {
            (void)hipGetLastError();
            set_error("cannot allocate %zu bytes of traversal spill", need);
            return -1;
        }
        w.spill_bytes = need;
    }
    *out = w.spill;
    return 0;
}

unsigned long long *launch_counter(hipStream_t stream) {
    std::lock_guard<std::mutex> g(g_ws_lock);
    StreamWorkspace &w = workspace_of(stream);
    if (!w.counter) {
        if (hipMalloc((void **)&w.counter, 4 * sizeof(unsigned long long)) != hipSuccess) {
            (void)hipGetLastError();
            set_error("cannot allocate the queue counter");
            return nullptr;
        }
    }
    return w.counter;
}


static SceneView view_of(const wos_scene *s) {
    SceneView v;
    v.dim = s->dim;
    v.n_vertices = s->n_vertices;
    v.n_prims = s->n_prims;
    v.use_bvh = s->use_bvh;
    v.vtx = s->vtx;
    v.idx = s->idx;
    v.val = s->val;
    v.nrm = s->nrm;
    for (int k = 0; k < 3; ++k) { v.amin[k] = s->amin[k]; v.amax[k] = s->amax[k]; }
    v.source_type = s->source_type;
    v.n_src = s->n_src;
    v.src = s->src;
    v.nodes = s->nodes;
    v.leaves = s->leaves;
    v.pleaves = s->leaves + (s->dim == 2 ? 2 : 3) * (size_t)s->n_prims;
    v.tol = s->tol;
    v.max_depth = s->max_depth;
    v.stats = nullptr;
    int lv = std::min(std::max(g_tuning.top_levels, 0), 16);
    int nt = s->n_prims > 1 ? s->top_counts[lv] : 0;
    const int node_bytes = node_words(s->dim) * (int)sizeof(float4);
    while (nt > 0 && (size_t)nt * node_bytes > 32768) nt = s->top_counts[--lv];
    v.n_top = nt;
    v.grid_start = s->grid.start;
    v.grid_cand = s->grid.cand;
    v.grid_ox = s->grid.ox;
    v.grid_oy = s->grid.oy;
    v.grid_inv_cell = s->grid.inv_cell;
    v.grid_nx = s->grid.nx;
    v.grid_ny = s->grid.ny;
    v.start_node = s->sgrid.start;
    v.st_ox = s->sgrid.ox;
    v.st_oy = s->sgrid.oy;
    v.st_oz = s->sgrid.oz;
    v.st_inv_cell = s->sgrid.inv_cell;
    v.st_nx = s->sgrid.nx;
    v.st_ny = s->sgrid.ny;
    v.st_nz = s->sgrid.nz;
    v.st_packed = s->sgrid.packed;
    return v;
}

const float *green_cdf_host() {
    static float table[kGreenCdfSize];
    static std::once_flag once;
    std::call_once(once, [] {
        for (int i = 0; i < kGreenCdfSize; ++i) {
            const double x = (double)(i + 1) / (double)kGreenCdfSize;
            table[i] = (float)(x * x * (1.0 - 2.0 * log(x)));
        }
    });
    return table;
}

SceneView scene_view(const wos_scene *s) { return view_of(s); }
unsigned long long scene_generation(const wos_scene *s) { return s->generation; }

}  // namespace wos

using namespace wos;

#define CHK(x)                                                       \
    do {                                                             \
        hipError_t e_ = (x);                                         \
        if (e_ != hipSuccess) {                                      \
            set_error("%s: %s", #x, hipGetErrorString(e_));          \
            return -1;                                               \
        }                                                            \
    } while (0)

#define REQUIRE(cond, ...)            \
    do {                              \
        if (!(cond)) {                \
            set_error(__VA_ARGS__);   \
            return -1;                \
        }                             \
    } while (0)

static hipStream_t S(void *stream) { return reinterpret_cast<hipStream_t>(stream); }

static void free_scene(wos_scene *s) {
    if (!s) return;
    (void)hipFree(s->vtx);
    (void)hipFree(s->val);
    (void)hipFree(s->nrm);
    (void)hipFree(s->src);
    (void)hipFree(s->idx);
    (void)hipFree(s->sorted_prim);
    (void)hipFree(s->children);
    (void)hipFree(s->sorted_codes);
    (void)hipFree(s->nodes);
    (void)hipFree(s->leaves);
    (void)hipFree(s->grid.start);
    (void)hipFree(s->grid.cand);
    (void)hipFree(s->sgrid.start);
    delete s;
}

static GridInputs grid_inputs(const wos_scene *s) {
    GridInputs gi;
    gi.n_prims = s->n_prims;
    gi.max_depth = s->max_depth;
    gi.vtx = s->vtx;
    gi.idx = s->idx;
    gi.nodes = s->nodes;
    gi.leaves = s->leaves;
    gi.sorted_prim = s->sorted_prim;
    for (int k = 0; k < 2; ++k) { gi.amin[k] = s->amin[k]; gi.amax[k] = s->amax[k]; }
    gi.tol = s->tol;
    return gi;
}

static int build_grid(wos_scene *s, const GridSettings &gs, hipStream_t st) {
    return build_grid_2d(grid_inputs(s), gs, st, s->grid);
}

static int build_start_grid(wos_scene *s, int resolution, hipStream_t st) {
    if (s->dim == 2) return build_start_grid_2d(grid_inputs(s), resolution, st, s->sgrid);
    if (resolution == kStartGridAuto) resolution = kStartGridAuto3;
    GridInputs3 gi;
    gi.n_prims = s->n_prims;
    gi.max_depth = s->max_depth;
    gi.nodes = s->nodes;
    gi.leaves = s->leaves;
    for (int k = 0; k < 3; ++k) { gi.amin[k] = s->amin[k]; gi.amax[k] = s->amax[k]; }
    gi.tol = s->tol;
    return build_start_grid_3d(gi, resolution, st, s->sgrid);
}

static int check_source(int source_type, const float *params, int n) {
    REQUIRE(source_type >= 0 && source_type <= 3, "source_type must be 0..3 (got %d)", source_type);
    if (source_type == 1 || source_type == 2)
        REQUIRE(params && n >= 5, "source_type %d needs 5 parameters (got %d)", source_type, n);
    if (source_type == 3)
        REQUIRE(params && n >= 4 && n % 4 == 0, "Fourier source needs 4 parameters per term (got %d)", n);
    return 0;
}

extern "C" {

const char *wos_last_error(void) { return g_err; }

int wos_debug_fail_workspace(int n) {
    REQUIRE(n >= 0, "n must be >= 0 (got %d)", n);
    std::lock_guard<std::mutex> g(g_ws_lock);
    g_fail_workspace_allocs = n;
    return 0;
}

int wos_scene_create(int dim, const float *vertices, int n_vertices, const int *indices, int n_prims,
                     const float *values, int use_bvh, int source_type, const float *source_params,
                     int n_source_params, void *stream, wos_scene **out) {
    g_err[0] = 0;
    REQUIRE(out, "out is NULL");
    *out = nullptr;
    REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3 (got %d)", dim);
    REQUIRE(n_vertices > 0 && n_prims > 0, "empty scene (%d vertices, %d primitives)", n_vertices, n_prims);
    // node and leaf byte offsets are 32-bit in the traversal (64 B per 2D
    // node, 96 B per 3D node)
    REQUIRE(n_prims <= (dim == 2 ? (1 << 26) : (1 << 25)), "at most 2^%d primitives per %dD scene (got %d)",
            dim == 2 ? 26 : 25, dim, n_prims);
    REQUIRE(vertices && indices && values, "vertices, indices and values are required");
    if (check_source(source_type, source_params, n_source_params)) return -1;
    hipStream_t st = S(stream);
    wos_scene *s = new wos_scene();
    s->dim = dim;
    s->n_vertices = n_vertices;
    s->n_prims = n_prims;
    s->use_bvh = use_bvh ? 1 : 0;
    s->source_type = source_type;
    s->n_src = source_type ? n_source_params : 0;
    (void)hipGetDevice(&s->device);
    int nn = n_prims > 1 ? n_prims - 1 : 1;
    size_t vb = sizeof(float) * (size_t)n_vertices * dim;
#define ACHK(x)                \
    if ((x) != hipSuccess) {   \
        set_error("%s failed", #x); \
        free_scene(s);         \
        return -1;             \
    }
    ACHK(hipMalloc((void **)&s->vtx, vb));
    ACHK(hipMalloc((void **)&s->nrm, vb));
    ACHK(hipMalloc((void **)&s->val, sizeof(float) * (size_t)n_vertices));
    ACHK(hipMalloc((void **)&s->idx, sizeof(int) * (size_t)n_prims * dim));
    ACHK(hipMalloc((void **)&s->sorted_prim, sizeof(int) * (size_t)n_prims));
    ACHK(hipMalloc((void **)&s->sorted_codes, sizeof(uint32_t) * (size_t)n_prims));
    ACHK(hipMalloc((void **)&s->children, sizeof(int) * 2 * (size_t)nn));
    ACHK(hipMalloc((void **)&s->nodes, sizeof(float4) * node_words(dim) * (size_t)nn));
    ACHK(hipMalloc((void **)&s->leaves, sizeof(float4) * (dim == 2 ? 4 : 6) * (size_t)n_prims));
    ACHK(hipMalloc((void **)&s->src, sizeof(float) * (size_t)std::max(1, s->n_src)));
    ACHK(hipMemcpyAsync(s->vtx, vertices, vb, hipMemcpyDefault, st));
    ACHK(hipMemcpyAsync(s->val, values, sizeof(float) * (size_t)n_vertices, hipMemcpyDefault, st));
    ACHK(hipMemcpyAsync(s->idx, indices, sizeof(int) * (size_t)n_prims * dim, hipMemcpyDefault, st));
    if (s->n_src) ACHK(hipMemcpyAsync(s->src, source_params, sizeof(float) * (size_t)s->n_src, hipMemcpyDefault, st));
#undef ACHK
    BuildInputs in{dim, n_vertices, n_prims, s->vtx, s->idx};
    BuildOutputs bo;
    bo.nodes = s->nodes;
    bo.leaves = s->leaves;
    bo.normals = s->nrm;
    bo.sorted_prim = s->sorted_prim;
    bo.sorted_codes = s->sorted_codes;
    bo.children = s->children;
    if (build_scene_gpu(in, st, bo) != 0) {
        free_scene(s);
        return -1;
    }
    if (bo.error_bits & 2u) {
        free_scene(s);
        set_error("primitive index out of range [0, %d)", n_vertices);
        return -1;
    }
    if (bo.error_bits & 1u) {
        free_scene(s);
        set_error("degenerate primitive (zero length or zero area)");
        return -1;
    }
    float m = 1.0f;
    for (int k = 0; k < 3; ++k) {
        s->amin[k] = bo.amin[k];
        s->amax[k] = bo.amax[k];
        m = std::max(m, std::max(fabsf(bo.amin[k]), fabsf(bo.amax[k])));
    }
    s->tol = 1e-6f * m;
    s->max_depth = bo.max_depth;
    for (int k = 0; k < 17; ++k) s->top_counts[k] = bo.top_counts[k];
    if (build_start_grid(s, dim == 2 ? kStartGridResolution2 : kStartGridResolution3, st) != 0 ||
        (dim == 2 && build_grid(s, GridSettings{}, st) != 0)) {
        free_scene(s);
        return -1;
    }
    *out = s;
    return 0;
}

int wos_scene_set_grid(wos_scene *s, float cell_factor, int band_cells, int max_cand, void *stream) {
    REQUIRE(s, "scene is NULL");
    REQUIRE(s->dim == 2, "the candidate grid is built for 2D scenes only");
    REQUIRE(band_cells >= 0 && max_cand >= 0, "band_cells and max_cand must be non-negative");
    CHK(hipDeviceSynchronize());   // kernels in flight may still read the old grid
    ++s->generation;
    (void)hipFree(s->grid.start);
    (void)hipFree(s->grid.cand);
    s->grid = GridOutputs{};
    GridSettings gs;
    gs.cell_factor = cell_factor;
    gs.band_cells = band_cells;
    gs.max_cand = max_cand;
    return build_grid(s, gs, S(stream));
}

int wos_scene_set_start_grid(wos_scene *s, int resolution, void *stream) {
    REQUIRE(s, "scene is NULL");
    const int cap = s->dim == 2 ? 8192 : 256;
    REQUIRE(resolution >= kStartGridAuto && resolution <= cap,
            "resolution must be in [0, %d] or %d (automatic) (got %d)", cap, kStartGridAuto, resolution);
    CHK(hipDeviceSynchronize());   // kernels in flight may still read the old grid
    ++s->generation;
    (void)hipFree(s->sgrid.start);
    s->sgrid = StartGridOutputs{};
    return build_start_grid(s, resolution, S(stream));
}

int wos_scene_get_start_grid(const wos_scene *s, int *nx, int *ny, int *nz) {
    REQUIRE(s, "scene is NULL");
    if (nx) *nx = s->sgrid.nx;
    if (ny) *ny = s->sgrid.ny;
    if (nz) *nz = s->sgrid.nz;
    return 0;
}

int wos_scene_get_grid(const wos_scene *s, int *nx, int *ny, float *cell, unsigned int *n_cand) {
    REQUIRE(s, "scene is NULL");
    if (nx) *nx = s->grid.nx;
    if (ny) *ny = s->grid.ny;
    if (cell) *cell = s->grid.cell;
    if (n_cand) *n_cand = s->grid.n_cand;
    return 0;
}

int wos_set_tuning(int chunk, int wgs_per_cu, int top_levels, int window) {
    REQUIRE(chunk >= 64 && chunk <= (1 << 20), "chunk must be in [64, 2^20] (got %d)", chunk);
    REQUIRE(wgs_per_cu >= 1 && wgs_per_cu <= 32, "wgs_per_cu must be in [1, 32] (got %d)", wgs_per_cu);
    REQUIRE(top_levels >= 0 && top_levels <= 16, "top_levels must be in [0, 16] (got %d)", top_levels);
    REQUIRE(window >= 0 && window <= 4096, "window must be in [0, 4096] (0: per dimension; got %d)", window);
    g_tuning.window = window;
    g_tuning.chunk = chunk;
    g_tuning.wgs_per_cu = wgs_per_cu;
    g_tuning.top_levels = top_levels;
    return 0;
}

int wos_set_spatial_order(int enable) {
    g_tuning.spatial_order = enable ? 1 : 0;
    return 0;
}

int wos_get_tuning(int *chunk, int *wgs_per_cu, int *top_levels, int *window) {
    if (window) *window = g_tuning.window;
    if (chunk) *chunk = g_tuning.chunk;
    if (wgs_per_cu) *wgs_per_cu = g_tuning.wgs_per_cu;
    if (top_levels) *top_levels = g_tuning.top_levels;
    return 0;
}

int wos_scene_destroy(wos_scene *scene) {
    if (!scene) return 0;
    (void)hipDeviceSynchronize();
    free_scene(scene);
    return 0;
}

int wos_scene_set_values(wos_scene *s, const float *values, void *stream) {
    REQUIRE(s && values, "scene and values are required");
    CHK(hipMemcpyAsync(s->val, values, sizeof(float) * (size_t)s->n_vertices, hipMemcpyDefault, S(stream)));
    return 0;
}

int wos_scene_set_source(wos_scene *s, int source_type, const float *params, int n, void *stream) {
    REQUIRE(s, "scene is NULL");
    if (check_source(source_type, params, n)) return -1;
    int need = source_type ? n : 0;
    ++s->generation;
    if (need > s->n_src) {
        CHK(hipStreamSynchronize(S(stream)));
        (void)hipFree(s->src);
        s->src = nullptr;
        CHK(hipMalloc((void **)&s->src, sizeof(float) * (size_t)need));
    }
    if (need) CHK(hipMemcpyAsync(s->src, params, sizeof(float) * (size_t)need, hipMemcpyDefault, S(stream)));
    s->source_type = source_type;
    s->n_src = need;
    return 0;
}

int wos_scene_set_use_bvh(wos_scene *s, int use_bvh) {
    REQUIRE(s, "scene is NULL");
    if ((use_bvh ? 1 : 0) != s->use_bvh) ++s->generation;
    s->use_bvh = use_bvh ? 1 : 0;
    return 0;
}

int wos_scene_get_normals(const wos_scene *s, float *out, void *stream) {
    REQUIRE(s && out, "scene and out are required");
    CHK(hipMemcpyAsync(out, s->nrm, sizeof(float) * (size_t)s->n_vertices * s->dim, hipMemcpyDefault, S(stream)));
    return 0;
}

int wos_scene_get_info(const wos_scene *s, float *amin, float *amax, int *dim, int *nv, int *np, int *depth) {
    REQUIRE(s, "scene is NULL");
    for (int k = 0; k < 3; ++k) {
        if (amin) amin[k] = s->amin[k];
        if (amax) amax[k] = s->amax[k];
    }
    if (dim) *dim = s->dim;
    if (nv) *nv = s->n_vertices;
    if (np) *np = s->n_prims;
    if (depth) *depth = s->max_depth;
    return 0;
}

int wos_scene_get_bvh(const wos_scene *s, int *sorted_prim, uint32_t *codes, int *children, void *stream) {
    REQUIRE(s, "scene is NULL");
    hipStream_t st = S(stream);
    if (sorted_prim) CHK(hipMemcpyAsync(sorted_prim, s->sorted_prim, sizeof(int) * (size_t)s->n_prims, hipMemcpyDefault, st));
    if (codes) CHK(hipMemcpyAsync(codes, s->sorted_codes, sizeof(uint32_t) * (size_t)s->n_prims, hipMemcpyDefault, st));
    if (children && s->n_prims > 1)
        CHK(hipMemcpyAsync(children, s->children, sizeof(int) * 2 * (size_t)(s->n_prims - 1), hipMemcpyDefault, st));
    return 0;
}

int wos_closest_point_index(const wos_scene *s, const float *p, int n, int *out, float *out_d2, void *stream) {
    REQUIRE(s && (n == 0 || (p && out)), "scene, p and out are required");
    REQUIRE(n >= 0, "negative count");
    return launch_closest_index(view_of(s), p, n, out, out_d2, S(stream));
}

int wos_sdf(const wos_scene *s, const float *p, int n, float *out, void *stream) {
    REQUIRE(s && (n == 0 || (p && out)), "scene, p and out are required");
    REQUIRE(n >= 0, "negative count");
    return launch_sdf(view_of(s), p, n, out, S(stream));
}

int wos_closest_point(const wos_scene *s, const float *p, int n, const wos_record *o, void *stream) {
    REQUIRE(s && o && (n == 0 || p), "scene, p and out are required");
    REQUIRE(n >= 0, "negative count");
    return launch_closest_point(view_of(s), p, n, o->valid, o->p, o->n, o->t, o->prim_id, o->contrib, S(stream));
}

int wos_largest_inscribed_ball(const wos_scene *s, const wos_record *its, int n, float epsilon, float *out,
                               void *stream) {
    REQUIRE(s && its && (n == 0 || (its->p && its->n && out)), "scene, its.p, its.n and out are required");
    REQUIRE(n >= 0, "negative count");
    return launch_lib(view_of(s), its->valid, its->p, its->n, n, epsilon, out, S(stream));
}

int wos_dirichlet(const wos_scene *s, const wos_record *rec, int n, float *out, void *stream) {
    REQUIRE(s && rec && (n == 0 || (rec->prim_id && rec->t && out)), "scene, rec.prim_id, rec.t and out are required");
    REQUIRE(n >= 0, "negative count");
    return launch_dirichlet(view_of(s), rec->valid, rec->prim_id, rec->t, n, out, S(stream));
}

static int check_walk(const wos_walk_params *w) {
    REQUIRE(w, "walk params are NULL");
    REQUIRE(w->nwalks >= 1, "nwalks must be >= 1 (got %d)", w->nwalks);
    REQUIRE(w->nsteps >= 0, "nsteps must be >= 0 (got %d)", w->nsteps);
    REQUIRE(w->epsilon >= 0.0f, "epsilon must be >= 0");
    return 0;
}

int wos_single_walk(const wos_scene *s, const wos_walk_params *w, const float *p, int n, uint64_t *state,
                    uint64_t *inc, const wos_record *o, void *stream) {
    REQUIRE(s && o && (n == 0 || (p && state && inc)), "scene, p, sampler state/inc and out are required");
    REQUIRE(n >= 0, "negative count");
    if (check_walk(w)) return -1;
    WalkSettings ws{w->nwalks, w->nsteps, w->epsilon, w->double_sided, w->prevent_fd_artifacts,
                    w->use_IS_for_greens ? 1 : 0};
    return launch_single_walk(view_of(s), ws, p, n, state, inc, o->valid, o->p, o->n, o->t, o->prim_id, o->contrib,
                              S(stream));
}

int wos_walk_reduce(const wos_scene *s, const wos_walk_params *w, const float *points, int n_points, int seed,
                    int chunk, float *mean, double *mean64, double *var64, float *walk_value, int *rec_prim,
                    float *rec_t, double *basis, void *stream) {
    REQUIRE(s && (n_points == 0 || points), "scene and points are required");
    REQUIRE(n_points >= 0, "negative count");
    if (check_walk(w)) return -1;
    REQUIRE((long long)n_points * w->nwalks < (1LL << 40), "too many walks");
    REQUIRE(!(rec_prim && !rec_t) && !(rec_t && !rec_prim), "rec_prim and rec_t go together");
    if (basis) REQUIRE(s->source_type == 3, "basis output needs a Fourier source (source_type 3)");
    WalkSettings ws{w->nwalks, w->nsteps, w->epsilon, w->double_sided, w->prevent_fd_artifacts,
                    w->use_IS_for_greens ? 1 : 0};
    WalkOutputs o{walk_value, rec_prim, rec_t, basis, mean, mean64, var64};
    return launch_walk_reduce(view_of(s), ws, points, n_points, (int32_t)seed, chunk, o, S(stream));
}

int wos_walk_reduce_noalloc(const wos_scene *s, const wos_walk_params *w, const float *points, int n_points,
                            int seed, int chunk, float *mean, double *mean64, double *var64, float *walk_value,
                            int *rec_prim, float *rec_t, double *basis, void *stream) {
    REQUIRE(s && (n_points == 0 || points), "scene and points are required");
    REQUIRE(n_points >= 0, "negative count");
    if (check_walk(w)) return -1;
    REQUIRE(!(rec_prim && !rec_t) && !(rec_t && !rec_prim), "rec_prim and rec_t go together");
    if (basis) REQUIRE(s->source_type == 3, "basis output needs a Fourier source (source_type 3)");
    WalkSettings ws{w->nwalks, w->nsteps, w->epsilon, w->double_sided, w->prevent_fd_artifacts,
                    w->use_IS_for_greens ? 1 : 0};
    WalkOutputs o{walk_value, rec_prim, rec_t, basis, mean, mean64, var64};
    return launch_walk_reduce_noalloc(view_of(s), ws, points, n_points, (int32_t)seed, chunk, o, S(stream));
}

int wos_walk_reduce_counts(const wos_scene *s, const wos_walk_params *w, const float *points, int n_points,
                           int seed, int with_basis, float *walk_value, unsigned long long *counts, void *stream) {
    REQUIRE(s && counts && (n_points == 0 || (points && walk_value)), "scene, points, walk_value and counts are required");
    if (check_walk(w)) return -1;
    if (n_points == 0) return 0;
    SceneView v = view_of(s);
    v.stats = counts;
    WalkSettings ws{w->nwalks, w->nsteps, w->epsilon, w->double_sided, w->prevent_fd_artifacts,
                    w->use_IS_for_greens ? 1 : 0};
    WalkOutputs o{walk_value, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double *basis = nullptr;
    if (with_basis && s->source_type == 3) {
        size_t nb = (size_t)(s->n_src / 4) * (size_t)n_points * sizeof(double);
        CHK(hipMallocAsync((void **)&basis, nb, S(stream)));
        o.basis = basis;
    }
    int rc = launch_walk_reduce(v, ws, points, n_points, (int32_t)seed, 0, o, S(stream));
    if (basis) CHK(hipFreeAsync(basis, S(stream)));
    return rc;
}

int wos_adjoint_values(const wos_scene *s, int n_points, int nwalks, const int *rec_prim, const float *rec_t,
                       const double *basis, const float *grad_out, float *grad_values, float *grad_coeffs,
                       void *stream) {
    REQUIRE(s && (n_points == 0 || (rec_prim && rec_t && grad_out)), "scene, records and grad_out are required");
    REQUIRE(n_points >= 0 && nwalks >= 1, "bad sizes");
    if (basis) REQUIRE(s->source_type == 3, "basis input needs a Fourier source (source_type 3)");
    return launch_adjoint_values(view_of(s), n_points, nwalks, rec_prim, rec_t, basis, grad_out, grad_values,
                                 grad_coeffs, S(stream));
}

int wos_adjoint_source(const wos_scene *s, const wos_walk_params *w, const float *points, int n_points, int seed,
                       const float *grad_out, float *grad_source, void *stream) {
    REQUIRE(s && (n_points == 0 || (points && grad_out && grad_source)),
            "scene, points, grad_out and grad_source are required");
    REQUIRE(n_points >= 0, "negative count");
    if (check_walk(w)) return -1;
    WalkSettings ws{w->nwalks, w->nsteps, w->epsilon, w->double_sided, w->prevent_fd_artifacts,
                    w->use_IS_for_greens ? 1 : 0};
    return launch_source_adjoint(view_of(s), ws, points, n_points, (int32_t)seed, grad_out, grad_source, S(stream));
}

int wos_shape_gradient(const wos_scene *s, const wos_grad_params *g, const float *points, int n_points, int seed,
                       const float *grad_out, float *grad_vertices, float *sample_grad_u, void *stream) {
    REQUIRE(s && g && (n_points == 0 || points), "scene, params and points are required");
    REQUIRE(n_points >= 0 && g->nwalks >= 1 && g->nsteps >= 0, "bad sizes");
    REQUIRE(n_points == 0 || !grad_vertices || grad_out, "grad_vertices needs grad_out");
    GradSettings gs{g->nwalks, g->nsteps, g->epsilon, g->epsilon2, g->clamping, g->double_sided, g->control_variates};
    return launch_shape_gradient(view_of(s), gs, points, n_points, (int32_t)seed, grad_out, grad_vertices,
                                 sample_grad_u, S(stream));
}

int wos_shape_gradient_counts(const wos_scene *s, const wos_grad_params *g, const float *points, int n_points,
                              int seed, unsigned long long *counts, void *stream) {
    REQUIRE(s && g && counts && (n_points == 0 || points), "scene, params, points and counts are required");
    REQUIRE(n_points >= 0 && g->nwalks >= 1 && g->nsteps >= 0, "bad sizes");
    SceneView v = view_of(s);
    v.stats = counts;
    GradSettings gs{g->nwalks, g->nsteps, g->epsilon, g->epsilon2, g->clamping, g->double_sided, g->control_variates};
    return launch_shape_gradient(v, gs, points, n_points, (int32_t)seed, nullptr, nullptr, nullptr, S(stream));
}

int wos_shape_derivative(const wos_scene *s, const wos_grad_params *g, const float *points, int n_points, int seed,
                         const float *velocity, float *out_mean, double *out_mean64, double *out_var64,
                         void *stream) {
    REQUIRE(s && g && (n_points == 0 || (points && velocity)), "scene, params, points and velocity are required");
    REQUIRE(n_points >= 0 && g->nwalks >= 1 && g->nsteps >= 0, "bad sizes");
    GradSettings gs{g->nwalks, g->nsteps, g->epsilon, g->epsilon2, g->clamping, g->double_sided, g->control_variates};
    return launch_shape_derivative(view_of(s), gs, points, n_points, (int32_t)seed, velocity, out_mean, out_mean64,
                                   out_var64, S(stream));
}

}  // extern "C"
