// wos_kernels.hip -- walk-on-spheres kernels for gfx950.
//
//   k_closest_index / k_sdf / k_closest_point / k_lib / k_dirichlet
//       Scene<N>::closestPointPreliminary / sdf / closestPoint /
//       largestInscribedBall / dirichlet (include/scene.cuh:102-264, 497-704)
//   k_single_walk      WoS<N>::singleWalk (include/wos.cuh:374-533): one lane
//                      per walk, sampler state read and written back
//   k_walk_persistent  Solver.walk_reduce (wos/solver.py:49-65) over
//                      WoSDevice::walk (include/wos.cuh:97-145): persistent
//                      waves that refill finished lanes from a chunked queue
//   k_reduce           per-point FP64 mean / variance of the walk values
//   k_adjoint_values   reverse mode of the walk value w.r.t. vertex values and
//                      Fourier source coefficients, from stored exit records
//   k_shape_gradient   OursCUDA / Ours3D (wos/wos_grad.py:130-151,
//                      wos/wos_grad_3d.py:117-137) transposed onto vertices
#include <algorithm>
#include <mutex>

#include "wos_build.h"
#include "wos_grad.h"

namespace wos {

constexpr int kMaxBasis = 16;   // Fourier source terms carried per lane
// Basis template argument of the walk kernel: KB > 0 carries exactly KB
// Fourier terms (no per-term guards), kAnyBasis up to kMaxBasis terms guarded
// by the run-time count, 0 none.
constexpr int kAnyBasis = -1;
__host__ __device__ constexpr int basis_cap(int KB) { return KB < 0 ? kMaxBasis : KB; }
// Walks that end wait (outputs held in registers) until this many lanes of
// the wave are finished or idle, then write their outputs and are refilled
// together: the write / refill path then runs with several lanes active
// instead of once per round for one or two.
constexpr int kFinishBatch = 8;
constexpr int kGridCap = 4096;  // workgroups of a grid-stride launch

// This file's device copy of the tabulated Green's-function CDF, filled once
// per device before the first launch that may sample it.
static __device__ float g_green_cdf[kGreenCdfSize];

static int ensure_green_cdf() {
    static std::mutex lock;
    static bool ready[64] = {false};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        set_error("cannot query the current device");
        return -1;
    }
    std::lock_guard<std::mutex> g(lock);
    if (ready[dev]) return 0;
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_green_cdf), green_cdf_host(), sizeof(float) * kGreenCdfSize) != hipSuccess) {
        (void)hipGetLastError();
        set_error("cannot upload the Green's-function CDF table");
        return -1;
    }
    ready[dev] = true;
    return 0;
}

#define KCHK()                                                        \
    do {                                                              \
        hipError_t e_ = hipGetLastError();                            \
        if (e_ != hipSuccess) {                                       \
            set_error("kernel launch: %s", hipGetErrorString(e_));    \
            return -1;                                                \
        }                                                             \
    } while (0)

static int grid_of(long n, int cap = kGridCap) {
    long g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// Per-workgroup traversal state: the LDS stack and, when s.n_top > 0, an LDS
// copy of the first s.n_top nodes (dynamic shared memory).
#define DECLARE_STACK(spill_, stride_)                                          \
    load_top_nodes(s);                                                          \
    TravStack st{(spill_), (stride_)};                                          \
    const int slot = threadIdx.x;                                               \
    const long lane_g = (long)blockIdx.x * kBlock + threadIdx.x;

// ---------------------------------------------------------------------------
// point queries
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kBlock) void k_closest_index(SceneView s, const float *p, int n, int *out,
                                                          float *out_d2, int2 *spill, long stride) {
    DECLARE_STACK(spill, stride);
    for (long i = lane_g; i < n; i += (long)gridDim.x * kBlock) {
        float q[3];
        load_point<D>(p, i, q);
        float d2;
        int id = closest_prim<D>(s, q, __builtin_inff(), d2, st, slot, lane_g);
        out[i] = id;
        if (out_d2) out_d2[i] = d2;
    }
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_sdf(SceneView s, const float *p, int n, float *out, int2 *spill,
                                                long stride) {
    DECLARE_STACK(spill, stride);
    for (long i = lane_g; i < n; i += (long)gridDim.x * kBlock) {
        float q[3];
        load_point<D>(p, i, q);
        float d2;
        int id = closest_prim<D>(s, q, __builtin_inff(), d2, st, slot, lane_g);
        out[i] = id < 0 ? __builtin_inff() : signed_dist<D>(s, id, q);
    }
}

template <int D>
__device__ __forceinline__ void store_record(const Rec &r, long i, uint8_t *valid, float *op, float *on, float *ot,
                                             int *oprim, float *ocontrib, float contrib) {
    if (valid) valid[i] = (uint8_t)r.valid;
    if (op)
        for (int k = 0; k < D; ++k) op[D * i + k] = r.valid ? r.p[k] : 0.0f;
    if (on)
        for (int k = 0; k < D; ++k) on[D * i + k] = r.valid ? r.n[k] : 0.0f;
    if (ot) {
        if (D == 2) ot[i] = r.valid ? r.t : 0.0f;
        else { ot[2 * i] = r.valid ? r.uv[0] : 0.0f; ot[2 * i + 1] = r.valid ? r.uv[1] : 0.0f; }
    }
    if (oprim) oprim[i] = r.valid ? r.prim : -1;
    if (ocontrib) ocontrib[i] = contrib;
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_closest_point(SceneView s, const float *p, int n, uint8_t *valid,
                                                          float *op, float *on, float *ot, int *oprim,
                                                          float *ocontrib, int2 *spill, long stride) {
    DECLARE_STACK(spill, stride);
    for (long i = lane_g; i < n; i += (long)gridDim.x * kBlock) {
        float q[3];
        load_point<D>(p, i, q);
        float d2;
        int id = closest_prim<D>(s, q, __builtin_inff(), d2, st, slot, lane_g);
        Rec r;
        r.valid = 0;
        if (id >= 0) closest_record<D>(s, id, q, r);
        store_record<D>(r, i, valid, op, on, ot, oprim, ocontrib, 0.0f);
    }
}

// include/scene.cuh:225-239; the sdf at p + n m lies within m (+ rounding) of
// the boundary point p, which bounds the search.
template <int D>
__device__ float lib_one(const SceneView &s, const float *p, const float *nrm, float eps, const TravStack &st,
                         int slot, long lane_g) {
    float b = fminf(fabsf(s.amax[0] - s.amin[0]), fabsf(s.amax[1] - s.amin[1]));
    if (D == 3) b = fminf(b, fabsf(s.amax[2] - s.amin[2]));
    b = b / 2.0f;
    float a = eps;
    for (int it = 0; it < 10; ++it) {
        float m = (a + b) / 2.0f;
        float q[3] = {fmaf(nrm[0], m, p[0]), fmaf(nrm[1], m, p[1]), D == 3 ? fmaf(nrm[2], m, p[2]) : 0.0f};
        float d2;
        int id = closest_prim<D>(s, q, bound_d2(s, m), d2, st, slot, lane_g);
        float d = fabsf(signed_dist<D>(s, id, q));
        if (fabsf(d - m) < 1e-5f) a = m; else b = m;
    }
    return a;
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_lib(SceneView s, const uint8_t *valid, const float *p, const float *nrm,
                                                int n, float eps, float *out, int2 *spill, long stride) {
    DECLARE_STACK(spill, stride);
    for (long i = lane_g; i < n; i += (long)gridDim.x * kBlock) {
        if (valid && !valid[i]) { out[i] = 0.0f; continue; }
        float q[3], nn[3];
        load_point<D>(p, i, q);
        load_point<D>(nrm, i, nn);
        out[i] = lib_one<D>(s, q, nn, eps, st, slot, lane_g);
    }
}

// lib_one for many 3D boundary points on persistent lanes: each lane runs
// its point's ten bisection queries as resumable queries, advanced `window`
// node visits per round, each seeded with the previous iteration's closest
// primitive (any seed leaves the answer exact: a bounded search that ends
// beyond its bound continues unbounded, as closest_prim falls back), and
// takes the next point of its wave's chunk when the tenth ends.  Needs the
// BVH (use_bvh, two or more triangles).
__global__ __launch_bounds__(kBlock) void k_lib_persistent3(SceneView s, const uint8_t *valid, const float *p,
                                                            const float *nrm, int n, float eps,
                                                            unsigned long long *counter, int chunk, int window,
                                                            float *out, int2 *spill, long stride) {
    DECLARE_STACK(spill, stride);
    const int lane = __lane_id();
    float upper = fminf(fabsf(s.amax[0] - s.amin[0]), fabsf(s.amax[1] - s.amin[1]));
    upper = fminf(upper, fabsf(s.amax[2] - s.amin[2])) / 2.0f;
    long long i = -1;
    bool busy = false, widened = false;
    int it = 0, prim = -1;
    float a = 0.0f, b = 0.0f, m = 0.0f, bound = 0.0f;
    float pp[3] = {0.0f, 0.0f, 0.0f}, nn[3] = {0.0f, 0.0f, 0.0f}, q3[3] = {0.0f, 0.0f, 0.0f};
    Query qq;
    qq.node = qq.sp = 0;
    qq.best = qq.thr = 0.0f;
    qq.best_id = -1;
    qq.nbox = qq.nprim = 0;
    // the query of iteration it (m from the current a, b), seeded with prim
    auto begin = [&]() {
        m = (a + b) / 2.0f;
        for (int k = 0; k < 3; ++k) q3[k] = fmaf(nn[k], m, pp[k]);
        bound = bound_d2(s, m);
        if (prim >= 0) query_begin_at<3>(qq, s, q3, bound, prim, prim_d2<3>(s, prim, q3), st, slot, lane_g);
        else query_begin_at<3>(qq, s, q3, bound, 0x7fffffff, __builtin_inff(), st, slot, lane_g);
        widened = false;
    };
    long long c_next = 0, c_end = 0;
    bool exhausted = false;
    for (;;) {
        const uint64_t idle = __ballot(!busy);
        if (idle != 0 && !exhausted) {
            const int need = __popcll(idle);
            const long long avail = c_end - c_next;
            long long base2 = 0, end2 = 0;
            bool fresh = false;
            if (avail < need) {
                unsigned long long bq = 0;
                if (lane == 0) bq = atomicAdd(counter, (unsigned long long)chunk);
                const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)bq);
                const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(bq >> 32));
                const long long bb = (long long)(((uint64_t)hi << 32) | lo);
                if (bb >= n) {
                    exhausted = true;
                } else {
                    fresh = true;
                    base2 = bb;
                    end2 = bb + chunk < n ? bb + chunk : n;
                }
            }
            if (!busy) {
                const int r = rank_in_mask(idle);
                long long j = -1;
                if (r < avail) j = c_next + r;
                else if (fresh && (r - avail) < end2 - base2) j = base2 + (r - avail);
                if (j >= 0) {
                    i = j;
                    if (valid && !valid[i]) {
                        out[i] = 0.0f;
                    } else {
                        load_point<3>(p, i, pp);
                        load_point<3>(nrm, i, nn);
                        a = eps;
                        b = upper;
                        it = 0;
                        prim = -1;
                        begin();
                        busy = true;
                    }
                }
            }
            if (avail >= need) {
                c_next += need;
            } else if (fresh) {
                const long long used2 = need - avail;
                c_next = base2 + used2 < end2 ? base2 + used2 : end2;
                c_end = end2;
            } else {
                c_next = c_end;
            }
        }
        if (exhausted && __ballot(busy) == 0) break;
        for (int k = 0; k < window; ++k) {
            if (busy && query_step<3, false>(s, q3, qq, st, slot, lane_g, s.tol)) {
                if (!widened && (qq.best_id == 0x7fffffff || qq.best > bound)) {
                    // nothing within the bound: the unbounded search
                    query_begin_at<3>(qq, s, q3, __builtin_inff(), qq.best_id, qq.best, st, slot, lane_g);
                    widened = true;
                } else {
                    prim = qq.best_id;
                    const float d = fabsf(signed_dist<3>(s, prim, q3));
                    if (fabsf(d - m) < 1e-5f) a = m;
                    else b = m;
                    if (++it == 10) {
                        out[i] = a;
                        busy = false;
                    } else {
                        begin();
                    }
                }
            }
            if (__ballot(busy) == 0) break;
        }
    }
}

template <int D>
__global__ void k_dirichlet(SceneView s, const uint8_t *valid, const int *prim, const float *t, int n, float *out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int f = prim[i];
        bool ok = (!valid || valid[i]) && f >= 0 && f < s.n_prims;
        float v = 0.0f;
        if (ok) v = (D == 2) ? dirichlet_at<2>(s, f, t[i], 0.0f, 0.0f) : dirichlet_at<3>(s, f, 0.0f, t[2 * i], t[2 * i + 1]);
        out[i] = v;
    }
}

// ---------------------------------------------------------------------------
// one walk on one lane (include/wos.cuh:97-145)
// ---------------------------------------------------------------------------
constexpr int kMaxSourceParams = 4 * kMaxBasis;   // Fourier: (ax, ay, phase, coeff) per term

// Adds d f / d theta (x, y) * Gw to acc[theta] for every parameter theta of
// the scene's source (include/scene.cuh:278-301: Gaussian (mu_x, mu_y,
// sigma_x, sigma_y, A), sinusoid (A, B, C, amplitude, offset); Fourier
// sum_k c_k sin(ax_k x + ay_k y + ph_k) as (ax, ay, ph, c) per term).
// KMAX: Fourier terms the unrolled loop covers (the scene's K <= KMAX).
template <int KMAX = kMaxBasis>
__device__ __forceinline__ void add_source_param_grad(const SceneView &s, float x, float y, float Gw, float *acc) {
    const float *q = s.src;
    if (s.source_type == 1) {
        const float ex = x - q[0], ey = y - q[1];
        const float isx2 = 1.0f / (q[2] * q[2]), isy2 = 1.0f / (q[3] * q[3]);
        const float e = exp_poly(-(ex * ex) * (0.5f * isx2) - (ey * ey) * (0.5f * isy2));
        const float f = q[4] * e;
        acc[0] = fmaf(f * ex * isx2, Gw, acc[0]);
        acc[1] = fmaf(f * ey * isy2, Gw, acc[1]);
        acc[2] = fmaf(f * (ex * ex) * isx2 / q[2], Gw, acc[2]);
        acc[3] = fmaf(f * (ey * ey) * isy2 / q[3], Gw, acc[3]);
        acc[4] = fmaf(e, Gw, acc[4]);
    } else if (s.source_type == 2) {
        const float phi = fmaf(q[0], x, fmaf(q[1], y, q[2]));
        const float ac = q[3] * sin_cw(phi + 1.57079632679489661923f);
        acc[0] = fmaf(ac * x, Gw, acc[0]);
        acc[1] = fmaf(ac * y, Gw, acc[1]);
        acc[2] = fmaf(ac, Gw, acc[2]);
        acc[3] = fmaf(sin_cw(phi), Gw, acc[3]);
        acc[4] += Gw;
    } else if (s.source_type == 3) {
        const int K = s.n_src / 4;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            if (k < K) {
                const float phi = fmaf(q[4 * k], x, fmaf(q[4 * k + 1], y, q[4 * k + 2]));
                const float ac = q[4 * k + 3] * sin_cw(phi + 1.57079632679489661923f);
                acc[4 * k] = fmaf(ac * x, Gw, acc[4 * k]);
                acc[4 * k + 1] = fmaf(ac * y, Gw, acc[4 * k + 1]);
                acc[4 * k + 2] = fmaf(ac, Gw, acc[4 * k + 2]);
                acc[4 * k + 3] = fmaf(sin_cw(phi), Gw, acc[4 * k + 3]);
            }
        }
    }
}

// init_prim >= 0: the closest primitive of p0 is already known (init_d2).
// PGRAD: also accumulate d contrib / d theta for every source parameter
// into pacc (kMaxSourceParams floats, zeroed by the caller).
template <int D, bool PGRAD = false>
__device__ void walk_one(const SceneView &s, const WalkSettings &w, const float *p0, Pcg32 &rng, int init_prim,
                         float init_d2, const TravStack &st, int slot, long lane_g, Rec &rec, float &contrib,
                         float *pacc = nullptr) {
    float p[3] = {p0[0], p0[1], D == 3 ? p0[2] : 0.0f};
    rec.valid = 0;
    rec.prim = -1;
    contrib = 0.0f;
    float d2;
    int prim = init_prim;
    if (prim < 0) prim = closest_prim<D>(s, p, __builtin_inff(), d2, st, slot, lane_g);
    else d2 = init_d2;
    if (prim < 0) return;
    float sd = signed_dist<D>(s, prim, p);
    if (sd >= 0.0f && !w.double_sided) return;
    float d = fabsf(sd);
    if (w.prevent_fd_artifacts) d *= 0.5f;
    const bool src = (D == 2) && s.source_type != 0;
    for (int i = 0; i < w.nsteps; ++i) {
        if (d < w.epsilon) {
            closest_record<D>(s, prim, p, rec, sgn(sd));   // normal toward the walk's side
            return;
        }
        if constexpr (PGRAD) {
            if (src) {
                const float u0 = rng.next_1d(), u1 = rng.next_1d();
                float dyx, dyy, G, wt;
                if (interior_sample(u0, u1, d, w.use_is != 0, g_green_cdf, dyx, dyy, G, wt)) {
                    const float yx = p[0] + dyx, yy = p[1] + dyy;
                    contrib = fmaf(source_at(s, yx, yy) * G, wt, contrib);
                    add_source_param_grad(s, yx, yy, G * wt, pacc);
                }
            }
        } else {
            if (src) source_step<D>(s, p, d, w.use_is != 0, g_green_cdf, rng, contrib);
        }
        move_point<D>(p, d, rng);
        if (s.stats) atomicAdd(&s.stats[3], 1ull);
        prim = closest_prim<D>(s, p, bound_d2(s, 2.0f * d), d2, st, slot, lane_g);
        d = sqrtf(d2);
    }
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_single_walk(SceneView s, WalkSettings w, const float *p, int n,
                                                        uint64_t *state, uint64_t *inc, uint8_t *valid, float *op,
                                                        float *on, float *ot, int *oprim, float *ocontrib,
                                                        int2 *spill, long stride) {
    DECLARE_STACK(spill, stride);
    for (long i = lane_g; i < n; i += (long)gridDim.x * kBlock) {
        float q[3];
        load_point<D>(p, i, q);
        Pcg32 rng;
        rng.state = state[i];
        rng.inc = inc[i];
        Rec r;
        float contrib;
        walk_one<D>(s, w, q, rng, -1, 0.0f, st, slot, lane_g, r, contrib);
        store_record<D>(r, i, valid, op, on, ot, oprim, ocontrib, r.valid ? contrib : 0.0f);
        state[i] = rng.state;
        inc[i] = rng.inc;
    }
}

// Reverse mode of walk_reduce w.r.t. every source parameter: each walk k = i
// W + w is re-run from its key (TEA(seed, i), sequence k -- the streams of
// the forward launch, so the same paths), accumulating d contrib / d theta
// along the way (walk_one's steps, one per round); a walk that ends on the
// boundary contributes grad_out[i] / W times it (an invalid walk's value, and
// so its derivative, is 0).  Lanes are persistent: a lane whose walk ends
// writes the walk's terms to stage[(k - k0) * np + q] and takes the next walk
// of [k0, k1) from its wave's chunk, so a wave is not held by its longest
// walk.  k_source_slices then reduces fixed slices of 64 * batches walks.
// NP: parameters carried per lane (8 for the Gaussian and sinusoid, 4 K_max
// for the Fourier source), so the accumulators stay in registers.
// init_prim / init_sd: each point's start record (closest primitive and
// signed distance, as walk_one's start computes them).  Steps whose query is
// pending advance it `window` node visits per round (the walk kernel's
// resumable queries, seeded with the previous primitive).
template <int NP>
__global__ __launch_bounds__(kBlock) void k_source_walks(SceneView s, WalkSettings w, const float *points,
                                                         int32_t seed, const float *grad_out, const int *init_prim,
                                                         const float *init_sd, long long k0, long long k1,
                                                         unsigned long long *counter, int chunk, int window,
                                                         float *stage, int2 *spill, long stride) {
    DECLARE_STACK(spill, stride);
    const int np = s.n_src;
    const int W = w.nwalks;
    const int lane = __lane_id();
    const long long n = k1 - k0;
    long long j = -1;             // the lane's walk, relative to k0
    uint32_t i = 0;               // its point
    bool busy = false, querying = false;
    float p[3] = {0.0f, 0.0f, 0.0f};
    float d = 0.0f;
    int prim = -1, step = 0;
    Pcg32 rng;
    rng.state = rng.inc = 0;
    Query qq;
    qq.node = qq.sp = 0;
    qq.best = qq.thr = 0.0f;
    qq.best_id = -1;
    qq.nbox = qq.nprim = 0;
    const bool tree = s.use_bvh && s.n_prims > 1;
    float acc[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[q] = 0.0f;
    long long c_next = 0, c_end = 0;
    bool exhausted = false;
    for (;;) {
        const uint64_t idle = __ballot(!busy);
        if (idle != 0 && !exhausted) {
            const int need = __popcll(idle);
            const long long avail = c_end - c_next;
            long long base2 = 0, end2 = 0;
            bool fresh = false;
            if (avail < need) {
                unsigned long long b = 0;
                if (lane == 0) b = atomicAdd(counter, (unsigned long long)chunk);
                const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
                const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
                const long long bb = (long long)(((uint64_t)hi << 32) | lo);
                if (bb >= n) {
                    exhausted = true;
                } else {
                    fresh = true;
                    base2 = bb;
                    end2 = bb + chunk < n ? bb + chunk : n;
                }
            }
            if (!busy) {
                const int r = rank_in_mask(idle);
                long long jj = -1;
                if (r < avail) jj = c_next + r;
                else if (fresh && (r - avail) < end2 - base2) jj = base2 + (r - avail);
                if (jj >= 0) {
                    // walk_one's start: the closest primitive and the side
                    j = jj;
                    const long long k = k0 + j;
                    i = point_of_walk(k, W);
                    load_point<2>(points, i, p);
                    rng = walk_stream(seed, i, k);
#pragma unroll
                    for (int q = 0; q < NP; ++q) acc[q] = 0.0f;
                    step = 0;
                    prim = init_prim[i];
                    const float sd = init_sd[i];
                    bool start = prim >= 0;
                    if (start) {
                        start = !(sd >= 0.0f && !w.double_sided);
                        d = fabsf(sd);
                        if (w.prevent_fd_artifacts) d *= 0.5f;
                    }
                    if (start) {
                        busy = true;
                        querying = false;
                    } else {
#pragma unroll
                        for (int q = 0; q < NP; ++q)
                            if (q < np) stage[j * np + q] = acc[q] * 0.0f;
                    }
                }
            }
            if (avail >= need) {
                c_next += need;
            } else if (fresh) {
                const long long used2 = need - avail;
                c_next = base2 + used2 < end2 ? base2 + used2 : end2;
                c_end = end2;
            } else {
                c_next = c_end;
            }
        }
        if (exhausted && __ballot(busy) == 0) break;
        if (busy && !querying) {
            // one iteration of walk_one's loop: stop on the boundary, give up
            // after nsteps, or take a step
            const bool valid = step < w.nsteps && d < w.epsilon;
            if (valid || step >= w.nsteps) {
                const float gw = valid ? grad_out[i] / (float)W : 0.0f;
#pragma unroll
                for (int q = 0; q < NP; ++q)
                    if (q < np) stage[j * np + q] = acc[q] * gw;
                busy = false;
            } else {
                if (s.source_type != 0) {
                    const float u0 = rng.next_1d(), u1 = rng.next_1d();
                    float dyx, dyy, G, wt;
                    if (interior_sample(u0, u1, d, w.use_is != 0, g_green_cdf, dyx, dyy, G, wt))
                        add_source_param_grad<NP / 4>(s, p[0] + dyx, p[1] + dyy, G * wt, acc);
                }
                move_point<2>(p, d, rng);
                ++step;
                if (tree) {
                    // the walk kernel's query: from the start-grid nodes,
                    // seeded with the previous primitive (within 2d of the
                    // new point); the same exact closest primitive
                    query_begin_from<2>(qq, start_nodes_of<2>(s, p), bound_d2(s, 2.0f * d), prim,
                                        prim_d2<2>(s, prim, p), s.tol, st, slot, lane_g);
                    querying = true;
                } else {
                    float d2;
                    prim = closest_prim<2>(s, p, bound_d2(s, 2.0f * d), d2, st, slot, lane_g);
                    d = sqrtf(d2);
                }
            }
        }
        for (int it = 0; it < window; ++it) {
            if (busy && querying && query_step<2, false>(s, p, qq, st, slot, lane_g, s.tol)) {
                prim = qq.best_id;
                d = sqrtf(qq.best);
                querying = false;
            }
            if (__ballot(busy && querying) == 0) break;
        }
    }
}

// Slices sl of 64 * batches staged walks (walk k0 + j at stage[j * np]):
// each 64-walk batch is reduced across a wave by shuffles and lane q adds
// parameter q's batch sums in order, into part[sl * np + q].  A function of
// the staged values alone (not of which wave took a slice).
__global__ __launch_bounds__(kBlock) void k_source_slices(const float *stage, long long n, int np, int batches,
                                                          long long sl0, long long n_slices, float *part) {
    const int lane = __lane_id();
    const long long wave = ((long long)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const long long n_waves = ((long long)gridDim.x * kBlock) >> 6;
    for (long long sl = wave; sl < n_slices; sl += n_waves) {
        float mine = 0.0f;
        for (int b = 0; b < batches; ++b) {
            const long long jw = (sl * batches + b) * 64 + lane;
            for (int q = 0; q < np; ++q) {
                float v = jw < n ? stage[jw * np + q] : 0.0f;
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
                if (lane == q) mine += v;
            }
        }
        if (lane < np) part[(sl0 + sl) * np + lane] = mine;
    }
}

// grad_src[q] += the slices' sums of parameter q (blockIdx.x), added in FP64
// in a fixed order: thread t takes slices t, t + 256, ..., then a tree.
__global__ __launch_bounds__(kBlock) void k_source_adjoint_sum(const float *part, long long n_slices, int np,
                                                               float *grad_src) {
    __shared__ double red[kBlock];
    const int q = blockIdx.x;
    double acc = 0.0;
    for (long long sl = threadIdx.x; sl < n_slices; sl += kBlock) acc += (double)part[sl * np + q];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad_src[q] += (float)red[0];
}

// ---------------------------------------------------------------------------
// persistent walk kernel
// ---------------------------------------------------------------------------
struct PersistArgs {
    const float *points;
    int n_points;
    WalkSettings w;
    int32_t seed;
    long long total;                 // n_points * nwalks
    unsigned long long *counter;     // zero before launch
    int chunk;                       // walks claimed per queue visit
    int window;                      // node visits per lane between refills
    const int *init_prim;            // per point: closest primitive of the start point
    const float *init_sd;            // per point: signed distance of the start point
    const int *order;                // visiting order of the points (nullable: input order)
    float *walk_value;
    int *rec_prim;
    float *rec_t;
    unsigned long long *basis;       // K x n_points fixed-point sums (units of 1 / basis_scale), zeroed
    double basis_scale;              // 2^basis_exponent
    // staged outputs (nullable): per walk one 16-byte record {value, prim,
    // t or u, v} and, with a basis, kq float4 of basis sums, written as whole
    // vectors and split / reduced per point by k_reduce_staged
    float4 *stage;
    float4 *stage_basis;
    int kq;
    int K;
    int2 *spill;
    long spill_stride;
};

// All walks of a point start at the same place, so the (unbounded) start
// query runs once per point here instead of once per walk.
template <int D>
__global__ __launch_bounds__(kBlock) void k_point_init(SceneView s, const float *p, int n, const int *order,
                                                       int *prim, float *sd, int2 *spill, long stride) {
    DECLARE_STACK(spill, stride);
    for (long j = lane_g; j < n; j += (long)gridDim.x * kBlock) {
        // neighbouring lanes take neighbouring points when an order is given
        // (2D launches: 0.21 -> 0.16 ms for the clover's 126,820 points; the
        // 3D start queries measured slower in Morton order, 7 -> 17 ms on
        // 262,144 points, and keep the input order)
        const long i = order ? order[j] : j;
        float q[3];
        load_point<D>(p, i, q);
        float d2;
        int id = closest_prim<D>(s, q, __builtin_inff(), d2, st, slot, lane_g);
        prim[i] = id;
        sd[i] = id < 0 ? __builtin_inff() : signed_dist<D>(s, id, q);
    }
}


// What k_points_persistent3 writes per point.
enum : int {
    kPqIndex = 0,    // closest index (id) and its squared distance (f, nullable)
    kPqInit = 1,     // start records of a walk launch: closest primitive (id), signed distance (f)
    kPqSdf = 2,      // signed distance (f)
    kPqRecord = 3    // closest-point record (store_record's outputs)
};
struct PointOut {
    int *id;
    float *f;
    uint8_t *valid;
    float *op, *on, *ot;
    int *oprim;
    float *ocontrib;
};

// Unbounded 3D closest-primitive queries of many points on the walk kernel's
// persistent machinery: each lane advances its query `window` node visits
// per round and takes the next point of its wave's chunk as soon as its
// query ends, so a wave is not held by its slowest query (the start queries
// of interior points vary severalfold in length).  MODE: what is written
// (kPq*).  Needs the BVH (use_bvh, two or more triangles).
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_points_persistent3(SceneView s, const float *p, int n,
                                                               unsigned long long *counter, int chunk, int window,
                                                               PointOut o, int2 *spill, long stride) {
    DECLARE_STACK(spill, stride);
    const int lane = __lane_id();
    long long i = -1;
    bool busy = false;
    float q3[3] = {0.0f, 0.0f, 0.0f};
    Query q;
    q.node = q.sp = 0;
    q.best = q.thr = 0.0f;
    q.best_id = -1;
    q.nbox = q.nprim = 0;
    long long c_next = 0, c_end = 0;
    bool exhausted = false;
    for (;;) {
        const uint64_t idle = __ballot(!busy);
        if (idle != 0 && !exhausted) {
            const int need = __popcll(idle);
            const long long avail = c_end - c_next;
            long long base2 = 0, end2 = 0;
            bool fresh = false;
            if (avail < need) {
                unsigned long long b = 0;
                if (lane == 0) b = atomicAdd(counter, (unsigned long long)chunk);
                const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
                const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
                const long long bb = (long long)(((uint64_t)hi << 32) | lo);
                if (bb >= n) {
                    exhausted = true;
                } else {
                    fresh = true;
                    base2 = bb;
                    end2 = bb + chunk < n ? bb + chunk : n;
                }
            }
            if (!busy) {
                const int r = rank_in_mask(idle);
                long long j = -1;
                if (r < avail) j = c_next + r;
                else if (fresh && (r - avail) < end2 - base2) j = base2 + (r - avail);
                if (j >= 0) {
                    i = j;
                    load_point<3>(p, i, q3);
                    query_begin_at<3>(q, s, q3, __builtin_inff(), 0x7fffffff, __builtin_inff(), st, slot, lane_g);
                    busy = true;
                }
            }
            if (avail >= need) {
                c_next += need;
            } else if (fresh) {
                const long long used2 = need - avail;
                c_next = base2 + used2 < end2 ? base2 + used2 : end2;
                c_end = end2;
            } else {
                c_next = c_end;
            }
        }
        if (exhausted && __ballot(busy) == 0) break;
        for (int it = 0; it < window; ++it) {
            if (busy && query_step<3, true>(s, q3, q, st, slot, lane_g, s.tol)) {
                query_flush_stats<true>(s, q);
                if (s.stats) atomicAdd(&s.stats[2], 1ull);
                const int id = q.best_id == 0x7fffffff ? -1 : q.best_id;
                if constexpr (MODE == kPqIndex) {
                    o.id[i] = id;
                    if (o.f) o.f[i] = q.best;
                } else if constexpr (MODE == kPqInit || MODE == kPqSdf) {
                    if (MODE == kPqInit) o.id[i] = id;
                    o.f[i] = id < 0 ? __builtin_inff() : signed_dist<3>(s, id, q3);
                } else {
                    Rec r;
                    r.valid = 0;
                    if (id >= 0) closest_record<3>(s, id, q3, r);
                    store_record<3>(r, i, o.valid, o.op, o.on, o.ot, o.oprim, o.ocontrib, 0.0f);
                }
                busy = false;
            }
            if (__ballot(busy) == 0) break;
        }
    }
}


// Fixed-point basis accumulators -> doubles in place (x / scale).
__global__ void k_basis_finish(unsigned long long *acc, long n, double inv_scale) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        const double v = (double)(long long)acc[j] * inv_scale;
        acc[j] = (unsigned long long)__double_as_longlong(v);
    }
}

template <int D, int KB>
__device__ __forceinline__ void write_walk(const SceneView &s, const PersistArgs &a, long long k, bool valid,
                                           int prim, const float *p, float contrib, const float *bas) {
    float v = 0.0f, t0 = 0.0f, t1 = 0.0f;
    int pr = -1;
    if (valid) {
        Rec r;
        closest_record<D>(s, prim, p, r);
        v = dirichlet_at<D>(s, prim, r.t, r.uv[0], r.uv[1]) + contrib;
        pr = prim;
        t0 = (D == 2) ? r.t : r.uv[0];
        t1 = (D == 2) ? 0.0f : r.uv[1];
    }
    if (a.stage) {
        // walks end out of order: a few whole-vector stores per walk instead
        // of one 4-byte store into each of up to eleven arrays
        a.stage[k] = make_float4(v, __int_as_float(pr), t0, t1);
        if constexpr (KB != 0) {
            float4 *bp = a.stage_basis + k * a.kq;
#pragma unroll
            for (int j = 0; j < (basis_cap(KB) + 3) / 4; ++j) {
                if (KB > 0 || j < a.kq) {
                    float x[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) x[u] = (valid && 4 * j + u < basis_cap(KB)) ? bas[4 * j + u] : 0.0f;
                    bp[j] = make_float4(x[0], x[1], x[2], x[3]);
                }
            }
        }
        return;
    }
    if (a.walk_value) a.walk_value[k] = v;
    if (a.rec_prim) a.rec_prim[k] = pr;
    if (a.rec_t) {
        if (D == 2) a.rec_t[k] = t0;
        else { a.rec_t[2 * k] = t0; a.rec_t[2 * k + 1] = t1; }
    }
    if constexpr (KB != 0) {
        if (valid) add_basis<basis_cap(KB)>(a.basis + point_of_walk(k, a.w.nwalks), a.n_points, a.K, a.basis_scale, bas);
    }
}

// Each wave keeps 64 walks in flight.  A lane's life cycle: kStart (sample
// the source term, jump to the sphere, begin the closest-primitive query) ->
// kQuery (the query advances a few node visits per round) -> kReady (the
// closest primitive of the new point is known: stop on the boundary or take
// the next step) -> kDone when the walk ends.  Finished walks write their
// outputs in batches of kFinishBatch lanes; the lanes then take the next walk
// indices of the wave's chunk (ballot + mbcnt) and the wave claims a new chunk from the
// global counter when its chunk runs dry, so neither walk lengths nor query
// lengths idle the SIMD for long.  Every walk re-derives its random stream
// from (seed, point, walk index), so results do not depend on the schedule.
// KB: Fourier basis sums carried per lane (see kAnyBasis).  The 2D kernels
// are held to 64 VGPRs so that 8 waves share each SIMD (measured 14 % faster
// on the clover than 6 waves at 65-78 VGPRs); the 3D kernel measured best
// without the limit.
template <int D, bool SRC, int KB, bool STATS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(D == 2 ? 8 : 1))) void k_walk_persistent(SceneView s, PersistArgs a) {
    DECLARE_STACK(a.spill, a.spill_stride);
    const int lane = __lane_id();
    long long k = -1;
    int mode = kIdle;
    float p[3] = {0.0f, 0.0f, 0.0f};
    float d = 0.0f, contrib = 0.0f;
    int step = 0, prim = -1;
    Pcg32 rng;
    rng.state = rng.inc = 0;
    Query q;
    q.node = q.sp = 0;
    q.best = q.thr = 0.0f;
    q.best_id = -1;
    q.nbox = q.nprim = 0;
    float bas[KB != 0 ? basis_cap(KB) : 1];
#pragma unroll
    for (int i = 0; i < (KB != 0 ? basis_cap(KB) : 1); ++i) bas[i] = 0.0f;
    long long c_next = 0, c_end = 0;
    bool exhausted = false, done_valid = false;
    const int W = a.w.nwalks;
    // instrumented launches (STATS): the lane's work tallies (bound tests,
    // primitive tests, queries, steps) and the wave's lane-activity tallies,
    // added to the launch counters once, when the wave ends
    unsigned long long work[4] = {0ull, 0ull, 0ull, 0ull};
    unsigned long long act[4] = {0ull, 0ull, 0ull, 0ull};

    for (;;) {
        // 1. completed queries: stop on the boundary (include/wos.cuh:108-113),
        //    run out of steps (144), or take the next step
        if (mode == kReady) {
            prim = q.best_id;
            d = sqrtf(q.best);
            done_valid = d < a.w.epsilon && step < a.w.nsteps;
            const bool done_invalid = !done_valid && step >= a.w.nsteps;
            mode = (done_valid || done_invalid) ? kDone : kStart;
        }
        // 2. write finished walks once kFinishBatch lanes are done or idle
        //    (or nothing else is left to run), then refill idle lanes from
        //    the wave's chunk
        if (__popcll(__ballot(mode == kDone || mode == kIdle)) >= kFinishBatch || exhausted ||
            __ballot(mode == kStart || mode == kQuery) == 0) {
            if (mode == kDone) {
                write_walk<D, KB>(s, a, k, done_valid, prim, p, contrib, bas);
                mode = kIdle;
                k = -1;
            }
        }
        uint64_t idle = __ballot(mode == kIdle);
        if (idle != 0 && !exhausted) {
            int need = __popcll(idle);
            long long avail = c_end - c_next;
            long long base2 = 0, end2 = 0;
            bool fresh = false;
            if (avail < need) {
                unsigned long long b = 0;
                if (lane == 0) b = atomicAdd(a.counter, (unsigned long long)a.chunk);
                uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
                uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
                long long bb = (long long)(((uint64_t)hi << 32) | lo);
                if (bb >= a.total) {
                    exhausted = true;
                } else {
                    fresh = true;
                    base2 = bb;
                    end2 = bb + a.chunk < a.total ? bb + a.chunk : a.total;
                }
            }
            if (mode == kIdle) {
                int r = rank_in_mask(idle);
                long long wslot = -1;   // walk slot (slot is the lane's LDS stack slot)
                if (r < avail) wslot = c_next + r;
                else if (fresh && (r - avail) < end2 - base2) wslot = base2 + (r - avail);
                if (wslot >= 0) {
                    // wslot -> (point, walk): points in the visiting order
                    const uint32_t pi = point_of_walk(wslot, W);
                    const long long i = a.order ? (long long)a.order[pi] : (long long)pi;
                    const long long kk = i * W + ((uint32_t)wslot - pi * (uint32_t)W);
                    int ip = a.init_prim[i];
                    float sd = a.init_sd[i];
                    load_point<D>(a.points, i, p);
                    contrib = 0.0f;
#pragma unroll
                    for (int j = 0; j < (KB != 0 ? basis_cap(KB) : 1); ++j) bas[j] = 0.0f;
                    // include/wos.cuh:100-106, then the first loop test
                    bool outside = ip < 0 || (sd >= 0.0f && !a.w.double_sided);
                    float d0 = fabsf(sd);
                    if (a.w.prevent_fd_artifacts) d0 *= 0.5f;
                    if (outside || a.w.nsteps <= 0) {
                        write_walk<D, KB>(s, a, kk, false, ip, p, 0.0f, bas);
                    } else if (d0 < a.w.epsilon) {
                        write_walk<D, KB>(s, a, kk, true, ip, p, 0.0f, bas);
                    } else {
                        k = kk;
                        rng = walk_stream(a.seed, i, kk);
                        prim = ip;
                        d = d0;
                        step = 0;
                        mode = kStart;
                    }
                }
            }
            if (avail >= need) {
                c_next += need;
            } else if (fresh) {
                long long used2 = need - avail;
                c_next = base2 + used2 < end2 ? base2 + used2 : end2;
                c_end = end2;
            } else {
                c_next = c_end;
            }
        }
        if (exhausted && __ballot(mode != kIdle) == 0) break;

        // 3. a new step: source sample, jump to the sphere, begin the query
        //    within 2d of the new point, seeded with the previous primitive.
        //    The random numbers are drawn in stream order (source sample,
        //    then direction); the loads for the new point are issued before
        //    the source arithmetic at the old point so that they overlap it.
        if constexpr (STATS) {
            // lane activity: lanes taking a step, per wave step phase
            const uint64_t m = __ballot(mode == kStart);
            if (m != 0) {
                act[0] += (unsigned long long)__popcll(m);
                act[1] += 1ull;
            }
        }
        if (mode == kStart) {
            float su0 = 0.0f, su1 = 0.0f;
            if (SRC) {
                su0 = rng.next_1d();
                su1 = rng.next_1d();
            }
            const float px0 = p[0], py0 = p[1];
            move_point<D>(p, d, rng);
            ++step;
            if constexpr (STATS) work[3] += 1ull;
            int4 start = make_int4(0, -1, 0, 0);
            float4 sg = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float sinv = 0.0f;
            if (s.use_bvh) {
                start = start_nodes_of<D>(s, p);
                if constexpr (D == 2) {
                    const float4 *lf = at_byte_offset(s.pleaves, (uint32_t)prim << 5);
                    sg = lf[0];
                    sinv = lf[1].x;
                }
            }
            if (SRC) {
                // include/wos.cuh:115-139: one interior sample, f G w
                float dyx, dyy, G, w;
                if (interior_sample(su0, su1, d, a.w.use_is != 0, g_green_cdf, dyx, dyy, G, w)) {
                    float yx = px0 + dyx, yy = py0 + dyy;
                    if constexpr (KB != 0) {
                        float f = 0.0f;
#pragma unroll
                        for (int j = 0; j < basis_cap(KB); ++j) {
                            if (KB > 0 || j < a.K) {
                                const const_f32 *c = as_constant(s.src) + 4 * j;
                                float b = sin_cw(fmaf(c[0], yx, fmaf(c[1], yy, c[2])));
                                f = fmaf(c[3], b, f);
                                bas[j] = fmaf(b * G, w, bas[j]);
                                // one term at a time: interleaving the terms'
                                // arithmetic costs registers (occupancy); the
                                // coefficients are scalar loads (as_constant)
                                __builtin_amdgcn_sched_barrier(0);
                            }
                        }
                        contrib = fmaf(f * G, w, contrib);
                    } else {
                        contrib = fmaf(source_at(s, yx, yy) * G, w, contrib);
                    }
                }
            }
            if (!s.use_bvh) {
                float d2;
                int id = closest_brute<D>(s, p, d2);
                q.best = d2;
                q.best_id = id;
                mode = kReady;
            } else {
                if constexpr (STATS) {
                    work[2] += 1ull;
                    work[1] += 1ull;   // the seed primitive
                }
                float seed_d2;
                if constexpr (D == 2) {
                    float h, dx, dy;
                    seed_d2 = seg_d2(p[0], p[1], sg.x, sg.y, sg.z, sg.w, sinv, h, dx, dy);
                } else {
                    seed_d2 = prim_d2<D>(s, prim, p);
                }
                if (s.n_prims == 1) {
                    q.best = seed_d2;
                    q.best_id = prim;
                    mode = kReady;
                } else {
                    query_begin_from<D>(q, start, bound_d2(s, 2.0f * d), prim, seed_d2, s.tol, st, slot, lane_g);
                    mode = kQuery;
                }
            }
        }
        // 4. advance every pending query by up to `window` node visits
        for (int it = 0; it < a.window; ++it) {
            if constexpr (STATS) {
                // lane activity: querying lanes, per wave node-visit iteration
                const uint64_t m = __ballot(mode == kQuery);
                if (m != 0) {
                    act[2] += (unsigned long long)__popcll(m);
                    act[3] += 1ull;
                }
            }
            if (mode == kQuery) {
                if (query_step<D, STATS>(s, p, q, st, slot, lane_g, s.tol)) {
                    mode = kReady;
                    if constexpr (STATS) {
                        work[0] += q.nbox;
                        work[1] += q.nprim;
                    }
                }
            }
            if (__ballot(mode == kQuery) == 0) break;
        }
    }
    if constexpr (STATS) {
        for (int c = 0; c < 4; ++c) {
            unsigned long long v = work[c];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            work[c] = v;
        }
        if (lane == 0)
            for (int c = 0; c < 4; ++c) {
                atomicAdd(&s.stats[c], work[c]);
                atomicAdd(&s.stats[4 + c], act[c]);
            }
    }
}

// One wave per point: FP64 sum and sum of squares of its walk values.
__global__ __launch_bounds__(kBlock) void k_reduce(const float *vals, int n_points, int W, float *mean,
                                                   double *mean64, double *var64) {
    const int lane = __lane_id();
    long wave = ((long)blockIdx.x * kBlock + threadIdx.x) >> 6;
    long nwaves = ((long)gridDim.x * kBlock) >> 6;
    for (long i = wave; i < n_points; i += nwaves) {
        double sm = 0.0, s2 = 0.0;
        for (int w = lane; w < W; w += 64) {
            double v = vals[i * W + w];
            sm += v;
            s2 += v * v;
        }
        for (int off = 32; off > 0; off >>= 1) {
            sm += __shfl_xor(sm, off);
            s2 += __shfl_xor(s2, off);
        }
        if (lane == 0) {
            double m = sm / W;
            if (mean) mean[i] = (float)m;
            if (mean64) mean64[i] = m;
            if (var64) var64[i] = W > 1 ? (s2 - sm * m) / (W - 1) : 0.0;
        }
    }
}

// One wave per point over staged walk outputs: splits the records into the
// per-walk outputs (each nullable; whole lines, since a point's walks are
// contiguous), forms the FP64 mean / variance like k_reduce (same order, so
// identical results), and sums the basis (K <= kMaxBasis terms, kq float4
// per walk; invalid walks carry zeros) in 64-bit fixed point, units of
// 1 / scale, written as doubles basis[q * n_points + i].
template <int D>
__global__ __launch_bounds__(kBlock) void k_reduce_staged(const float4 *stage, const float4 *stage_basis, int kq,
                                                          int K, int n_points, int W, double scale,
                                                          float *walk_value, int *rec_prim, float *rec_t,
                                                          float *mean, double *mean64, double *var64,
                                                          double *basis) {
    const int lane = __lane_id();
    long wave = ((long)blockIdx.x * kBlock + threadIdx.x) >> 6;
    long nwaves = ((long)gridDim.x * kBlock) >> 6;
    for (long i = wave; i < n_points; i += nwaves) {
        double sm = 0.0, s2 = 0.0;
        long long bs[kMaxBasis];
#pragma unroll
        for (int q = 0; q < kMaxBasis; ++q) bs[q] = 0;
        for (int w = lane; w < W; w += 64) {
            const long long k = i * W + w;
            const float4 r = stage[k];
            if (walk_value) walk_value[k] = r.x;
            if (rec_prim) rec_prim[k] = __float_as_int(r.y);
            if (rec_t) {
                if (D == 2) rec_t[k] = r.z;
                else { rec_t[2 * k] = r.z; rec_t[2 * k + 1] = r.w; }
            }
            const double v = r.x;
            sm += v;
            s2 += v * v;
            if (basis) {
#pragma unroll
                for (int j = 0; j < kMaxBasis / 4; ++j) {
                    if (j < kq) {
                        const float4 b = stage_basis[k * kq + j];
                        const float x[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
                        for (int u = 0; u < 4; ++u) bs[4 * j + u] += __double2ll_rn((double)x[u] * scale);
                    }
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            sm += __shfl_xor(sm, off);
            s2 += __shfl_xor(s2, off);
        }
        if (basis) {
#pragma unroll
            for (int q = 0; q < kMaxBasis; ++q) {
                if (q < K) {
                    long long x = bs[q];
                    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
                    bs[q] = x;
                }
            }
        }
        if (lane == 0) {
            double m = sm / W;
            if (mean) mean[i] = (float)m;
            if (mean64) mean64[i] = m;
            if (var64) var64[i] = W > 1 ? (s2 - sm * m) / (W - 1) : 0.0;
            if (basis)
                for (int q = 0; q < K; ++q) basis[(long)q * n_points + i] = (double)bs[q] / scale;
        }
    }
}

// ---------------------------------------------------------------------------
// adjoint of the walk value w.r.t. vertex values and Fourier coefficients
// ---------------------------------------------------------------------------
constexpr int kLdsValues = 6144;   // vertex-value histogram (64-bit entries) kept in LDS up to this size

// Deterministic FP64 sums over the points: kSumSlices fixed slices per
// row, each reduced by one workgroup in a fixed order (thread strides, then
// a tree in LDS), then the slices of a row added in order.
constexpr int kSumSlices = 64;
__device__ __forceinline__ double block_sum(double acc) {
    __shared__ double part[kBlock];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    return part[0];
}

// part[blockIdx.x] = sum of |grad_out| over slice blockIdx.x.
__global__ __launch_bounds__(kBlock) void k_abs_slices(const float *grad_out, int n, double *part) {
    const long per = ((long)n + kSumSlices - 1) / kSumSlices;
    const long b = blockIdx.x * per, e = min((long)n, b + per);
    double acc = 0.0;
    for (long i = b + threadIdx.x; i < e; i += kBlock) acc += fabs((double)grad_out[i]);
    const double t = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// scale[0] = the value adjoint's fixed-point scale (value_scale_block) from
// the slice sums of |grad_out|.
__global__ void k_value_scale_final(const double *part, double *scale) {
    if (threadIdx.x != 0) return;
    double sum = 0.0;
    for (int b = 0; b < kSumSlices; ++b) sum += part[b];
    int e = 0;
    (void)frexp(sum, &e);
    scale[0] = sum > 0.0 ? ldexp(1.0, 61 - e) : 1.0;
}

// Fixed-point value sums -> grad_values (added into).
__global__ void k_value_finish(const unsigned long long *acc, int n, const double *scale, float *grad_values) {
    const double inv = 1.0 / scale[0];
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        if (acc[v] != 0ull) grad_values[v] += (float)((double)(long long)acc[v] * inv);
}

// Value part of the reverse pass: every walk's g_i / W times its exit
// primitive's interpolation weights, summed per vertex in 64-bit fixed
// point (units of 1 / scale[0]) into acc, through an LDS histogram sized to
// the mesh (small meshes leave room for more resident workgroups).
template <int D, bool LDSHIST>
__global__ __launch_bounds__(kBlock) void k_adjoint_values(SceneView s, int n_points, int W, const int *rec_prim,
                                                           const float *rec_t, const float *grad_out,
                                                           const double *scale, unsigned long long *acc) {
    extern __shared__ unsigned long long h[];
    if (LDSHIST)
        for (int v = threadIdx.x; v < s.n_vertices; v += kBlock) h[v] = 0ull;
    __syncthreads();
    const double sc = scale[0];
    long long total = (long long)n_points * W;
    for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < total; k += (long long)gridDim.x * kBlock) {
        int f = rec_prim[k];
        if (f < 0) continue;
        float g = grad_out[point_of_walk(k, W)] / (float)W;
        {
            if constexpr (D == 2) {
                float t = rec_t[k];
                int i0 = s.idx[2 * f], i1 = s.idx[2 * f + 1];
                const unsigned long long g0 = value_term(g * (1.0f - t), sc), g1 = value_term(g * t, sc);
                unsigned long long *dst = LDSHIST ? h : acc;
                atomicAdd(&dst[i0], g0);
                atomicAdd(&dst[i1], g1);
            } else {
                float u = rec_t[2 * k], v = rec_t[2 * k + 1];
                int i0 = s.idx[3 * f], i1 = s.idx[3 * f + 1], i2 = s.idx[3 * f + 2];
                const unsigned long long g0 = value_term(g * (1.0f - u - v), sc), g1 = value_term(g * u, sc),
                                         g2 = value_term(g * v, sc);
                unsigned long long *dst = LDSHIST ? h : acc;
                atomicAdd(&dst[i0], g0);
                atomicAdd(&dst[i1], g1);
                atomicAdd(&dst[i2], g2);
            }
        }
    }
    __syncthreads();
    if (LDSHIST)
        for (int v = threadIdx.x; v < s.n_vertices; v += kBlock)
            if (h[v] != 0ull) atomicAdd(&acc[v], h[v]);
}

// dL/d coeff_q = sum_i grad_out[i] / W * basis[q][i] (basis: per-point sums
// over the walks) as deterministic FP64 slice sums: part[q * kSumSlices +
// slice] for q = blockIdx.y, slice = blockIdx.x; k_adjoint_coeffs_final adds
// each coefficient's slices in order into grad_coeffs[q].
__global__ __launch_bounds__(kBlock) void k_adjoint_coeffs(int n_points, const double *basis,
                                                           const float *grad_out, double *part) {
    const int q = blockIdx.y;
    const double *bq = basis + (long)q * n_points;
    const long per = ((long)n_points + kSumSlices - 1) / kSumSlices;
    const long b = blockIdx.x * per, e = min((long)n_points, b + per);
    double acc = 0.0;
    for (long i = b + threadIdx.x; i < e; i += kBlock) acc = fma((double)grad_out[i], bq[i], acc);
    const double t = block_sum(acc);
    if (threadIdx.x == 0) part[q * kSumSlices + blockIdx.x] = t;
}

__global__ void k_adjoint_coeffs_final(const double *part, int K, int W, float *grad_coeffs) {
    const int q = threadIdx.x;
    if (q >= K) return;
    double sum = 0.0;
    for (int b = 0; b < kSumSlices; ++b) sum += part[q * kSumSlices + b];
    grad_coeffs[q] += (float)(sum / W);
}

// ---------------------------------------------------------------------------
// shape gradient (OursCUDA / Ours3D), one lane per (point, walk) sample
// ---------------------------------------------------------------------------
struct GradArgs {
    const float *points;
    int n_points;
    GradSettings g;
    int32_t seed;
    long long total;                 // n_points * nwalks samples
    unsigned long long *counter;     // zero before launch
    int chunk, window;
    const int *init_prim;            // per point: closest primitive of the start point
    const float *init_sd;            // per point: signed distance of the start point
    const int *order;                // visiting order of the points (nullable: input order)
    const float *grad_out;
    float *grad_vertices;
    float *sample_grad_u;
    const float *velocity;           // forward mode: per-vertex velocity (n_vertices * D), nullable
    float *sample_jvp;               // forward mode: per-sample -grad u . V(exit) (n_points * nwalks)
    float4 *vstage;                  // reverse mode: staged vertex-gradient terms (D - 1 per sample), nullable
    int2 *spill;
    long spill_stride;
};

// Workgroups of the fixed-point vertex scatter of a small mesh (each keeps an
// LDS histogram of the sums, written whole and summed in a second pass).
constexpr int kVgradBlocks = 256;

// Sample source of the single-scene shape-gradient launch (grad_samples in
// wos_grad.h): one scene and the launch arguments, uniform across the grid.
struct GradOne {
    const SceneView &s;
    const GradArgs &a;
    __device__ __forceinline__ long long total() const { return a.total; }
    __device__ __forceinline__ unsigned long long *counter() const { return a.counter; }
    __device__ __forceinline__ int chunk() const { return a.chunk; }
    __device__ __forceinline__ int window() const { return a.window; }
    __device__ __forceinline__ const SceneView &scene(int) const { return s; }
    __device__ __forceinline__ const GradSettings &settings(int) const { return a.g; }
    __device__ __forceinline__ const SceneView &trav(int, const SceneView &) const { return s; }
    __device__ __forceinline__ void bind(int, SceneView &) const {}
    // slot -> (point, sample): points in the visiting order
    __device__ __forceinline__ int take(long long wslot, long long &i, long long &kk) const {
        const int W = a.g.nwalks;
        const uint32_t pi = point_of_walk(wslot, W);
        i = a.order ? (long long)a.order[pi] : (long long)pi;
        kk = i * W + ((uint32_t)wslot - pi * (uint32_t)W);
        return 0;
    }
    __device__ __forceinline__ int start_prim(int, long long i) const { return a.init_prim[i]; }
    __device__ __forceinline__ float start_sd(int, long long i) const { return a.init_sd[i]; }
    template <int D>
    __device__ __forceinline__ void load(int, long long i, float *p) const { load_point<D>(a.points, i, p); }
    __device__ __forceinline__ int32_t seed(int) const { return a.seed; }
    template <int D>
    __device__ __forceinline__ void skip(int, long long kk) const {
        if (a.sample_grad_u)
            for (int c = 0; c < D; ++c) a.sample_grad_u[kk * D + c] = 0.0f;
        if (a.sample_jvp) a.sample_jvp[kk] = 0.0f;
        if (a.vstage) {
            const float zero[3] = {0.0f, 0.0f, 0.0f};
            stage_vertex_grad<D>(a.vstage, kk, false, 0, 0.0f, 0.0f, 0.0f, zero);
        }
    }
    // A finished sample: its spatial gradient, its forward-mode derivative
    // and the scatter -g_i / W w grad u onto the exit primitive's vertices.
    template <int D>
    __device__ __forceinline__ void finish(int, long long k, bool ok, int rprim, float rt0, float rt1,
                                           const float *gu) const {
        const int W = a.g.nwalks;
        if (a.sample_grad_u)
            for (int c = 0; c < D; ++c) a.sample_grad_u[k * D + c] = gu[c];
        if (a.sample_jvp) {
            // forward mode of the same integrand: the exit point moves with
            // the interpolated vertex velocity
            float jv = 0.0f;
            if (ok) {
                if constexpr (D == 2) {
                    const int i0 = s.idx[2 * rprim], i1 = s.idx[2 * rprim + 1];
                    for (int c = 0; c < 2; ++c)
                        jv = fmaf(-gu[c], fmaf(rt0, a.velocity[2 * i1 + c], (1.0f - rt0) * a.velocity[2 * i0 + c]), jv);
                } else {
                    const int i0 = s.idx[3 * rprim], i1 = s.idx[3 * rprim + 1], i2 = s.idx[3 * rprim + 2];
                    const float wa = 1.0f - rt0 - rt1;
                    for (int c = 0; c < 3; ++c) {
                        const float vc = fmaf(rt1, a.velocity[3 * i2 + c],
                                              fmaf(rt0, a.velocity[3 * i1 + c], wa * a.velocity[3 * i0 + c]));
                        jv = fmaf(-gu[c], vc, jv);
                    }
                }
            }
            a.sample_jvp[k] = jv;
        }
        if (a.vstage)
            stage_vertex_grad<D>(a.vstage, k, ok, rprim, rt0, rt1, a.grad_out[point_of_walk(k, W)] / (float)W, gu);
    }
};

// Shape-gradient samples (OursCUDA / Ours3D) of one scene: grad_samples
// (wos_grad.h) over GradOne.  3D is held to 80 VGPRs (6 waves per SIMD): the
// sample state needs ~106 (4 waves); at 80, with a few registers spilled,
// the 200k-triangle shape gradient measured 7 % faster on one MI355X and the
// same on another, and 3 % / 6 % faster than at 96 (5 waves) / 72 (7 waves).
// 2D is held to 64 (8 waves, a few spills): 1.5 % faster than at 80 on the
// clover.  Finished samples write in batches as
// in the walk kernel (2D: 4 % faster).  STATS: count bound tests, primitive
// tests, queries and walk steps into s.stats (the instrumented launch of
// the measurements).
template <int D, bool STATS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(D == 2 ? 8 : 6))) void k_grad_persistent(
    SceneView s, GradArgs a) {
    DECLARE_STACK(a.spill, a.spill_stride);
    const GradOne src{s, a};
    grad_samples<D, STATS>(src, st, slot, lane_g);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
#define DISPATCH_D(s, ...)                      \
    if ((s).dim == 2) { __VA_ARGS__(2) }        \
    else { __VA_ARGS__(3) }

Tuning g_tuning;

static size_t top_bytes(const SceneView &s) {
    return s.use_bvh ? (size_t)s.n_top * sizeof(float4) * node_words(s.dim) : 0;
}

static int cu_count() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (cus <= 0) cus = 256;
    }
    return cus;
}

static int occupancy_grid() { return cu_count() * (g_tuning.wgs_per_cu > 0 ? g_tuning.wgs_per_cu : 8); }

// Visiting order of the query points for the walk kernels: Morton order
// within the scene's box for point sets large enough to pay for the sort
// (null: input order).
constexpr int kOrderMinPoints = 4096;
static int spatial_order(const SceneView &s, const float *points, int n_points, int **order, hipStream_t st) {
    *order = nullptr;
    if (n_points < kOrderMinPoints || g_tuning.spatial_order == 0) return 0;
    if (hipMallocAsync((void **)order, sizeof(int) * (size_t)n_points, st) != hipSuccess) {
        set_error("cannot allocate the point order (%d points)", n_points);
        return -1;
    }
    return sort_points_morton(points, n_points, s.dim, s.amin, s.amax, *order, st);
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }
__global__ void k_zero_u64_n(unsigned long long *p, long n) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) p[j] = 0ull;
}

// Points per queue claim of k_points_persistent3.
constexpr int kPointChunk = 64;

// 3D scenes with a BVH take the persistent point-query kernel.
static bool persistent_points(const SceneView &s) { return s.dim == 3 && s.use_bvh && s.n_prims > 1; }

int launch_closest_index(const SceneView &s, const float *p, int n, int *out, float *out_d2, hipStream_t st) {
    if (n <= 0) return 0;
    int g = grid_of(n, occupancy_grid());
    long stride;
    int2 *spill = nullptr;
    if (spill_buffer(s, (long)g * kBlock, &stride, &spill, st) != 0) return -1;
    if (persistent_points(s)) {
        unsigned long long *counter = launch_counter(st);
        if (!counter) return -1;
        k_zero_u64<<<1, 1, 0, st>>>(counter);
        PointOut o{};
        o.id = out;
        o.f = out_d2;
        k_points_persistent3<kPqIndex><<<g, kBlock, top_bytes(s), st>>>(s, p, n, counter, kPointChunk, window_for(3),
                                                                         o, spill, stride);
    } else if (s.dim == 2) {
        k_closest_index<2><<<g, kBlock, top_bytes(s), st>>>(s, p, n, out, out_d2, spill, stride);
    } else {
        k_closest_index<3><<<g, kBlock, top_bytes(s), st>>>(s, p, n, out, out_d2, spill, stride);
    }
    KCHK();
    return 0;
}

int launch_sdf(const SceneView &s, const float *p, int n, float *out, hipStream_t st) {
    if (n <= 0) return 0;
    int g = grid_of(n, occupancy_grid());
    long stride;
    int2 *spill = nullptr;
    if (spill_buffer(s, (long)g * kBlock, &stride, &spill, st) != 0) return -1;
    if (persistent_points(s)) {
        unsigned long long *counter = launch_counter(st);
        if (!counter) return -1;
        PointOut o{};
        o.f = out;
        k_zero_u64<<<1, 1, 0, st>>>(counter);
        k_points_persistent3<kPqSdf><<<g, kBlock, top_bytes(s), st>>>(s, p, n, counter, kPointChunk, window_for(3),
                                                                       o, spill, stride);
    } else if (s.dim == 2) {
        k_sdf<2><<<g, kBlock, top_bytes(s), st>>>(s, p, n, out, spill, stride);
    } else {
        k_sdf<3><<<g, kBlock, top_bytes(s), st>>>(s, p, n, out, spill, stride);
    }
    KCHK();
    return 0;
}

int launch_closest_point(const SceneView &s, const float *p, int n, uint8_t *valid, float *op, float *on, float *ot,
                         int *oprim, float *ocontrib, hipStream_t st) {
    if (n <= 0) return 0;
    int g = grid_of(n, occupancy_grid());
    long stride;
    int2 *spill = nullptr;
    if (spill_buffer(s, (long)g * kBlock, &stride, &spill, st) != 0) return -1;
    if (persistent_points(s)) {
        unsigned long long *counter = launch_counter(st);
        if (!counter) return -1;
        const PointOut o{nullptr, nullptr, valid, op, on, ot, oprim, ocontrib};
        k_zero_u64<<<1, 1, 0, st>>>(counter);
        k_points_persistent3<kPqRecord><<<g, kBlock, top_bytes(s), st>>>(s, p, n, counter, kPointChunk,
                                                                          window_for(3), o, spill, stride);
    } else if (s.dim == 2) {
        k_closest_point<2><<<g, kBlock, top_bytes(s), st>>>(s, p, n, valid, op, on, ot, oprim, ocontrib, spill, stride);
    } else {
        k_closest_point<3><<<g, kBlock, top_bytes(s), st>>>(s, p, n, valid, op, on, ot, oprim, ocontrib, spill, stride);
    }
    KCHK();
    return 0;
}

int launch_lib(const SceneView &s, const uint8_t *valid, const float *p, const float *nrm, int n, float eps,
               float *out, hipStream_t st) {
    if (n <= 0) return 0;
    int g = grid_of(n, occupancy_grid());
    long stride;
    int2 *spill = nullptr;
    if (spill_buffer(s, (long)g * kBlock, &stride, &spill, st) != 0) return -1;
    if (persistent_points(s)) {
        unsigned long long *counter = launch_counter(st);
        if (!counter) return -1;
        k_zero_u64<<<1, 1, 0, st>>>(counter);
        k_lib_persistent3<<<g, kBlock, top_bytes(s), st>>>(s, valid, p, nrm, n, eps, counter, kPointChunk,
                                                            window_for(3), out, spill, stride);
    } else if (s.dim == 2) {
        k_lib<2><<<g, kBlock, top_bytes(s), st>>>(s, valid, p, nrm, n, eps, out, spill, stride);
    } else {
        k_lib<3><<<g, kBlock, top_bytes(s), st>>>(s, valid, p, nrm, n, eps, out, spill, stride);
    }
    KCHK();
    return 0;
}

int launch_dirichlet(const SceneView &s, const uint8_t *valid, const int *prim, const float *t, int n, float *out,
                     hipStream_t st) {
    if (n <= 0) return 0;
    if (s.dim == 2) k_dirichlet<2><<<grid_of(n), kBlock, 0, st>>>(s, valid, prim, t, n, out);
    else k_dirichlet<3><<<grid_of(n), kBlock, 0, st>>>(s, valid, prim, t, n, out);
    KCHK();
    return 0;
}

int launch_single_walk(const SceneView &s, const WalkSettings &w, const float *p, int n, uint64_t *state,
                       uint64_t *inc, uint8_t *valid, float *op, float *on, float *ot, int *oprim, float *ocontrib,
                       hipStream_t st) {
    if (n <= 0) return 0;
    if (w.use_is && s.source_type != 0 && ensure_green_cdf() != 0) return -1;
    int g = grid_of(n, occupancy_grid());
    long stride;
    int2 *spill = nullptr;
    if (spill_buffer(s, (long)g * kBlock, &stride, &spill, st) != 0) return -1;
    if (s.dim == 2)
        k_single_walk<2><<<g, kBlock, top_bytes(s), st>>>(s, w, p, n, state, inc, valid, op, on, ot, oprim, ocontrib, spill,
                                               stride);
    else
        k_single_walk<3><<<g, kBlock, top_bytes(s), st>>>(s, w, p, n, state, inc, valid, op, on, ot, oprim, ocontrib, spill,
                                               stride);
    KCHK();
    return 0;
}

// Per-launch start records of the persistent kernels (closest primitive and
// signed distance of each query point) and the points' visiting order.  The
// buffers are stream-ordered allocations released on the launch stream when
// the guard leaves scope -- after the kernels that read them are queued, or
// at once on an error return.
struct StartRecords {
    hipStream_t st;
    int *prim = nullptr;
    float *sd = nullptr;
    int *order = nullptr;
    explicit StartRecords(hipStream_t s) : st(s) {}
    StartRecords(const StartRecords &) = delete;
    StartRecords &operator=(const StartRecords &) = delete;
    ~StartRecords() {
        if (prim) (void)hipFreeAsync(prim, st);
        if (sd) (void)hipFreeAsync(sd, st);
        if (order) (void)hipFreeAsync(order, st);
    }
    int alloc(const SceneView &s, const float *points, int n_points) {
        if (hipMallocAsync((void **)&prim, sizeof(int) * (size_t)n_points, st) != hipSuccess ||
            hipMallocAsync((void **)&sd, sizeof(float) * (size_t)n_points, st) != hipSuccess) {
            (void)hipGetLastError();
            set_error("cannot allocate %d start records", n_points);
            return -1;
        }
        return spatial_order(s, points, n_points, &order, st);
    }
};

// Walk indices are split into (point, walk) in 32-bit arithmetic on the device.
static bool walk_count_fits(long long total) {
    if (total < (1LL << 32)) return true;
    set_error("at most 2^32 - 1 walks per launch (points x walks = %lld)", total);
    return false;
}

// Buffers of one walk_reduce launch beyond the caller's inputs and outputs.
struct WalkScratch {
    unsigned long long *counter = nullptr;   // the work queue's head, zeroed by the launch
    int *init_prim = nullptr;                // n_points start records
    float *init_sd = nullptr;
    const int *order = nullptr;              // visiting order of the points (nullable: input order)
    int2 *spill = nullptr;                   // traversal stack beyond the LDS part
    long spill_stride = 0;
    int grid = 1, grid_init = 1;             // persistent walk grid, start-query grid
    // staged outputs (nullable: the walk kernel then writes the caller's
    // per-walk arrays directly and adds the basis with atomics)
    float4 *stage = nullptr, *stage_basis = nullptr;
};

// Checks shared by both walk_reduce launchers; *K: Fourier terms.
static int walk_reduce_check(const SceneView &s, const WalkSettings &w, int n_points, const WalkOutputs &o, int *K) {
    if (!walk_count_fits((long long)n_points * w.nwalks)) return -1;
    *K = (s.source_type == 3) ? s.n_src / 4 : 0;
    if (o.basis && *K > kMaxBasis) {
        set_error("Fourier source has %d terms; at most %d are supported", *K, kMaxBasis);
        return -1;
    }
    if ((o.mean || o.mean64 || o.var64) && !o.walk_value) {
        set_error("per-point statistics need the per-walk value buffer");
        return -1;
    }
    if (w.use_is && s.source_type != 0 && ensure_green_cdf() != 0) return -1;
    return 0;
}

// Grids of the full-occupancy launch: one wave per 64 walks up to eight
// workgroups per CU for the walks, one lane per point for the start queries.
static void walk_reduce_grids(int n_points, int nwalks, int *g, int *gi) {
    const long long total = (long long)n_points * nwalks;
    long waves_needed = (long)((total + 63) / 64);
    int cap = occupancy_grid();
    long gg = (waves_needed + 3) / 4;
    *g = (int)std::max(1L, std::min(gg, (long)cap));
    *gi = grid_of(n_points, cap);
}

// Start records of a walk launch (closest primitive and signed distance per
// point) on grid gi; counter: the launch's queue head (3D scenes with a BVH
// use it here; the caller zeroes it again before the walks).
static void launch_point_init(const SceneView &s, const float *points, int n_points, const int *order, int *prim,
                              float *sd, int2 *spill, long stride, unsigned long long *counter, int gi,
                              hipStream_t st) {
    const size_t tb = top_bytes(s);
    if (persistent_points(s)) {
        k_zero_u64<<<1, 1, 0, st>>>(counter);
        PointOut o{};
        o.id = prim;
        o.f = sd;
        k_points_persistent3<kPqInit><<<gi, kBlock, tb, st>>>(s, points, n_points, counter, kPointChunk,
                                                               window_for(3), o, spill, stride);
    } else if (s.dim == 2) {
        k_point_init<2><<<gi, kBlock, tb, st>>>(s, points, n_points, order, prim, sd, spill, stride);
    } else {
        k_point_init<3><<<gi, kBlock, tb, st>>>(s, points, n_points, nullptr, prim, sd, spill, stride);
    }
}

static int walk_reduce_launch(const SceneView &s, const WalkSettings &w, const float *points, int n_points,
                              int32_t seed, int chunk, int K, const WalkOutputs &o, const WalkScratch &ws,
                              hipStream_t st) {
    PersistArgs a;
    a.points = points;
    a.n_points = n_points;
    a.w = w;
    a.seed = seed;
    a.total = (long long)n_points * w.nwalks;
    a.counter = ws.counter;
    a.chunk = chunk > 0 ? chunk : (g_tuning.chunk > 0 ? g_tuning.chunk : 256);
    a.window = window_for(s.dim);
    a.init_prim = ws.init_prim;
    a.init_sd = ws.init_sd;
    a.order = ws.order;
    a.walk_value = o.walk_value;
    a.rec_prim = o.rec_prim;
    a.rec_t = o.rec_t;
    a.basis = reinterpret_cast<unsigned long long *>(o.basis);
    a.basis_scale = ldexp(1.0, basis_exponent(w.nwalks, w.nsteps));
    a.stage = ws.stage;
    a.stage_basis = ws.stage_basis;
    a.kq = ws.stage_basis ? (K + 3) / 4 : 0;
    a.K = K;
    a.spill = ws.spill;
    a.spill_stride = ws.spill_stride;
    const int g = ws.grid, gi = ws.grid_init;
    size_t tb = top_bytes(s);
    bool src = s.dim == 2 && s.source_type != 0;
    bool bas = src && o.basis != nullptr && K > 0;
    const long nbas = (long)K * n_points;
    if (!ws.stage && o.basis && K > 0) k_zero_u64_n<<<grid_of(nbas), kBlock, 0, st>>>(a.basis, nbas);
    launch_point_init(s, points, n_points, ws.order, ws.init_prim, ws.init_sd, a.spill, a.spill_stride, a.counter, gi, st);
    k_zero_u64<<<1, 1, 0, st>>>(a.counter);
    if (s.dim == 2) {
        if (s.stats) {
            if (bas && a.K == 8) k_walk_persistent<2, true, 8, true><<<g, kBlock, tb, st>>>(s, a);
            else if (bas) k_walk_persistent<2, true, kAnyBasis, true><<<g, kBlock, tb, st>>>(s, a);
            else if (src) k_walk_persistent<2, true, 0, true><<<g, kBlock, tb, st>>>(s, a);
            else k_walk_persistent<2, false, 0, true><<<g, kBlock, tb, st>>>(s, a);
        } else {
            if (bas && a.K == 8) k_walk_persistent<2, true, 8, false><<<g, kBlock, tb, st>>>(s, a);
            else if (bas) k_walk_persistent<2, true, kAnyBasis, false><<<g, kBlock, tb, st>>>(s, a);
            else if (src) k_walk_persistent<2, true, 0, false><<<g, kBlock, tb, st>>>(s, a);
            else k_walk_persistent<2, false, 0, false><<<g, kBlock, tb, st>>>(s, a);
        }
    } else {
        if (s.stats) k_walk_persistent<3, false, 0, true><<<g, kBlock, tb, st>>>(s, a);
        else k_walk_persistent<3, false, 0, false><<<g, kBlock, tb, st>>>(s, a);
    }
    KCHK();
    if (ws.stage) {
        long waves = n_points;
        int gr = (int)((waves * 64 + kBlock - 1) / kBlock);
        if (gr > 8192) gr = 8192;
        double *bo = (o.basis && K > 0) ? o.basis : nullptr;
        if (s.dim == 2)
            k_reduce_staged<2><<<gr, kBlock, 0, st>>>(ws.stage, ws.stage_basis, a.kq, K, n_points, w.nwalks,
                                                     a.basis_scale, o.walk_value, o.rec_prim, o.rec_t, o.mean,
                                                     o.mean64, o.var64, bo);
        else
            k_reduce_staged<3><<<gr, kBlock, 0, st>>>(ws.stage, nullptr, 0, 0, n_points, w.nwalks, a.basis_scale,
                                                     o.walk_value, o.rec_prim, o.rec_t, o.mean, o.mean64, o.var64,
                                                     nullptr);
        KCHK();
        return 0;
    }
    if (o.basis && K > 0) {
        k_basis_finish<<<grid_of(nbas), kBlock, 0, st>>>(a.basis, nbas, 1.0 / a.basis_scale);
        KCHK();
    }
    if (o.mean || o.mean64 || o.var64) {
        long waves = n_points;
        int gr = (int)((waves * 64 + kBlock - 1) / kBlock);
        if (gr > 8192) gr = 8192;
        k_reduce<<<gr, kBlock, 0, st>>>(o.walk_value, n_points, w.nwalks, o.mean, o.mean64, o.var64);
        KCHK();
    }
    return 0;
}

int launch_walk_reduce(const SceneView &s, const WalkSettings &w, const float *points, int n_points, int32_t seed,
                       int chunk, const WalkOutputs &o, hipStream_t st) {
    if (n_points <= 0) return 0;
    int K = 0;
    if (walk_reduce_check(s, w, n_points, o, &K) != 0) return -1;
    WalkScratch ws;
    ws.counter = launch_counter(st);
    if (!ws.counter) return -1;
    walk_reduce_grids(n_points, w.nwalks, &ws.grid, &ws.grid_init);
    long lanes = (long)std::max(ws.grid, ws.grid_init) * kBlock;
    if (spill_buffer(s, lanes, &ws.spill_stride, &ws.spill, st) != 0) return -1;
    StartRecords sr(st);
    if (sr.alloc(s, points, n_points) != 0) return -1;
    ws.init_prim = sr.prim;
    ws.init_sd = sr.sd;
    ws.order = sr.order;
    // staged outputs: 16 bytes per walk, plus the basis rounded up to float4
    const size_t total = (size_t)n_points * (size_t)w.nwalks;
    const int kq = (o.basis && K > 0) ? (K + 3) / 4 : 0;
    if (hipMallocAsync((void **)&ws.stage, sizeof(float4) * total * (size_t)(1 + kq), st) != hipSuccess) {
        (void)hipGetLastError();
        set_error("cannot allocate the staged outputs of %zu walks", total);
        return -1;
    }
    ws.stage_basis = kq ? ws.stage + total : nullptr;
    const int rc = walk_reduce_launch(s, w, points, n_points, seed, chunk, K, o, ws, st);
    (void)hipFreeAsync(ws.stage, st);
    return rc;
}

// ---- allocation-free walk_reduce -----------------------------------------
// For callers that cannot allocate device memory: the queue head, the start
// records and the traversal spill that launch_walk_reduce allocates are cut
// from one file-scope device array instead.  The points keep their input
// order (the Morton sort needs buffers of its own); that changes which lanes
// walk which points, never a walk's value.  Every call on a device shares
// that device's array, so calls that overlap on one device (other streams,
// other host threads) race; calls on one stream run in order.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the current device, looked up on the first
// call there (each device holds its own copy of the array).
static int scratch_address(unsigned char **out) {
    static std::mutex lock;
    static unsigned char *addr[64] = {nullptr};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        set_error("cannot query the current device");
        return -1;
    }
    std::lock_guard<std::mutex> g(lock);
    if (!addr[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            set_error("cannot find the scratch array on device %d: %s", dev, hipGetErrorString(e));
            return -1;
        }
        addr[dev] = (unsigned char *)ptr;
    }
    *out = addr[dev];
    return 0;
}

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

int launch_walk_reduce_noalloc(const SceneView &s, const WalkSettings &w, const float *points, int n_points,
                               int32_t seed, int chunk, const WalkOutputs &o, hipStream_t st) {
    if (n_points <= 0) return 0;
    int K = 0;
    if (walk_reduce_check(s, w, n_points, o, &K) != 0) return -1;
    // layout: [queue head | start primitives | start distances | spill]
    const size_t rec = align256(sizeof(int) * (size_t)n_points);
    const size_t fixed = 256 + 2 * rec;
    const long deep = s.use_bvh ? std::max(0L, (long)s.max_depth + 2 - kLdsStack) : 0L;
    const size_t per_wg = (size_t)deep * sizeof(int2) * kBlock;   // spill of one workgroup
    if (fixed + per_wg > kScratchBytes) {
        long long most = (long long)((kScratchBytes - 256 - per_wg) / (2 * sizeof(int))) - 64;
        set_error("%d points need %zu bytes of scratch; the allocation-free walk_reduce has %zu "
                  "(at most about %lld points in this scene)", n_points, fixed + per_wg, kScratchBytes,
                  std::max(0LL, most));
        return -1;
    }
    unsigned char *base = nullptr;
    if (scratch_address(&base) != 0) return -1;
    WalkScratch ws;
    walk_reduce_grids(n_points, w.nwalks, &ws.grid, &ws.grid_init);
    if (per_wg > 0) {
        // fewer workgroups when the spill of a full grid does not fit (the
        // walk kernel is persistent and the start queries grid-stride)
        const long fit = (long)((kScratchBytes - fixed) / per_wg);
        ws.grid = (int)std::min((long)ws.grid, fit);
        ws.grid_init = (int)std::min((long)ws.grid_init, fit);
    }
    ws.counter = (unsigned long long *)base;
    ws.init_prim = (int *)(base + 256);
    ws.init_sd = (float *)(base + 256 + rec);
    ws.order = nullptr;
    ws.spill_stride = (long)std::max(ws.grid, ws.grid_init) * kBlock;
    ws.spill = per_wg > 0 ? (int2 *)(base + fixed) : nullptr;
    return walk_reduce_launch(s, w, points, n_points, seed, chunk, K, o, ws, st);
}

int launch_adjoint_values(const SceneView &s, int n_points, int nwalks, const int *rec_prim, const float *rec_t,
                          const double *basis, const float *grad_out, float *grad_values, float *grad_coeffs,
                          hipStream_t st) {
    if (n_points <= 0) return 0;
    int K = (s.source_type == 3) ? s.n_src / 4 : 0;
    if (basis && K > kMaxBasis) {
        set_error("Fourier source has %d terms; at most %d are supported", K, kMaxBasis);
        return -1;
    }
    long long total = (long long)n_points * nwalks;
    if (!walk_count_fits(total)) return -1;
    // up to eight workgroups per CU, each walking a strided share of the
    // records into its own LDS histogram (sized to the mesh: the 48 KB of a
    // fixed-size one allowed three per CU, 0.48 -> 0.36 ms on the clover)
    int g = grid_of((long)((total + 15) / 16), 8 * cu_count());
    bool lds = s.n_vertices <= kLdsValues;
    const size_t hb = lds ? sizeof(unsigned long long) * (size_t)s.n_vertices : 0;
    // scratch: fixed-point value sums (n_vertices), their scale, slice sums
    const bool want_c = basis && grad_coeffs && K > 0;
    if (!grad_values && !want_c) return 0;
    unsigned long long *acc = nullptr;
    const size_t nb = sizeof(unsigned long long) * ((size_t)s.n_vertices + 1 + (size_t)kSumSlices * (K + 1));
    if (hipMallocAsync((void **)&acc, nb, st) != hipSuccess) {
        (void)hipGetLastError();
        set_error("cannot allocate the adjoint scratch (%d vertices)", s.n_vertices);
        return -1;
    }
    double *scale = reinterpret_cast<double *>(acc + s.n_vertices);
    double *part = scale + 1;
    if (grad_values) {
        k_zero_u64_n<<<grid_of(s.n_vertices), kBlock, 0, st>>>(acc, s.n_vertices);
        k_abs_slices<<<kSumSlices, kBlock, 0, st>>>(grad_out, n_points, part);
        k_value_scale_final<<<1, 64, 0, st>>>(part, scale);
        if (s.dim == 2) {
            if (lds) k_adjoint_values<2, true><<<g, kBlock, hb, st>>>(s, n_points, nwalks, rec_prim, rec_t, grad_out, scale, acc);
            else k_adjoint_values<2, false><<<g, kBlock, 0, st>>>(s, n_points, nwalks, rec_prim, rec_t, grad_out, scale, acc);
        } else {
            if (lds) k_adjoint_values<3, true><<<g, kBlock, hb, st>>>(s, n_points, nwalks, rec_prim, rec_t, grad_out, scale, acc);
            else k_adjoint_values<3, false><<<g, kBlock, 0, st>>>(s, n_points, nwalks, rec_prim, rec_t, grad_out, scale, acc);
        }
        k_value_finish<<<grid_of(s.n_vertices), kBlock, 0, st>>>(acc, s.n_vertices, scale, grad_values);
    }
    if (want_c) {
        k_adjoint_coeffs<<<dim3(kSumSlices, K), kBlock, 0, st>>>(n_points, basis, grad_out, part + kSumSlices);
        k_adjoint_coeffs_final<<<1, 64, 0, st>>>(part + kSumSlices, K, nwalks, grad_coeffs);
    }
    (void)hipFreeAsync(acc, st);
    KCHK();
    return 0;
}

int launch_source_adjoint(const SceneView &s, const WalkSettings &w, const float *points, int n_points, int32_t seed,
                          const float *grad_out, float *grad_src, hipStream_t st) {
    if (n_points <= 0 || s.source_type == 0 || s.n_src <= 0) return 0;
    if (s.dim != 2) {
        set_error("source terms exist for 2D scenes only");
        return -1;
    }
    if (s.n_src > kMaxSourceParams) {
        set_error("source has %d parameters; at most %d are supported", s.n_src, kMaxSourceParams);
        return -1;
    }
    const long long total = (long long)n_points * w.nwalks;
    if (!walk_count_fits(total)) return -1;
    if (w.use_is && ensure_green_cdf() != 0) return -1;
    // slices of 64 * batches walks, batches the least power of two that keeps
    // the partial sums within 2^20 slices (a function of the walk count only)
    int batches = 1;
    while ((total + 64LL * batches - 1) / (64LL * batches) > (1LL << 20)) batches *= 2;
    const long long slice = 64LL * batches;
    const long long n_slices = (total + slice - 1) / slice;
    // walks staged per round: whole slices, about 2^22 walks
    const long long round = std::max<long long>(1, (1LL << 22) / slice) * slice;
    const long long per_round = std::min<long long>(round, total);
    const int g = grid_of((long)per_round, occupancy_grid());
    const int gi = grid_of(n_points, occupancy_grid());
    long stride;
    int2 *spill = nullptr;
    if (spill_buffer(s, (long)std::max(g, gi) * kBlock, &stride, &spill, st) != 0) return -1;
    unsigned long long *counter = launch_counter(st);
    if (!counter) return -1;
    StartRecords sr(st);
    if (sr.alloc(s, points, n_points) != 0) return -1;
    float *part = nullptr, *stage = nullptr;
    if (hipMallocAsync((void **)&part, sizeof(float) * (size_t)n_slices * s.n_src, st) != hipSuccess ||
        hipMallocAsync((void **)&stage, sizeof(float) * (size_t)per_round * s.n_src, st) != hipSuccess) {
        (void)hipGetLastError();
        if (part) (void)hipFreeAsync(part, st);
        set_error("cannot allocate the source-gradient terms of %lld walks", per_round);
        return -1;
    }
    const int gs = (int)std::max<long long>(1, std::min<long long>((per_round / slice + 3) / 4, occupancy_grid()));
    launch_point_init(s, points, n_points, sr.order, sr.prim, sr.sd, spill, stride, counter, gi, st);
    const int win = window_for(2);
    for (long long k0 = 0; k0 < total; k0 += round) {
        const long long k1 = std::min(total, k0 + round);
        k_zero_u64<<<1, 1, 0, st>>>(counter);
        if (s.n_src <= 8)
            k_source_walks<8><<<g, kBlock, top_bytes(s), st>>>(s, w, points, seed, grad_out, sr.prim, sr.sd, k0, k1,
                                                                counter, 128, win, stage, spill, stride);
        else if (s.n_src <= 32)
            k_source_walks<32><<<g, kBlock, top_bytes(s), st>>>(s, w, points, seed, grad_out, sr.prim, sr.sd, k0, k1,
                                                                 counter, 128, win, stage, spill, stride);
        else
            k_source_walks<kMaxSourceParams><<<g, kBlock, top_bytes(s), st>>>(s, w, points, seed, grad_out, sr.prim,
                                                                               sr.sd, k0, k1, counter, 128, win,
                                                                               stage, spill, stride);
        k_source_slices<<<gs, kBlock, 0, st>>>(stage, k1 - k0, s.n_src, batches, k0 / slice,
                                               (k1 - k0 + slice - 1) / slice, part);
    }
    k_source_adjoint_sum<<<s.n_src, kBlock, 0, st>>>(part, n_slices, s.n_src, grad_src);
    (void)hipFreeAsync(stage, st);
    (void)hipFreeAsync(part, st);
    KCHK();
    return 0;
}

int launch_shape_gradient(const SceneView &s, const GradSettings &g, const float *points, int n_points, int32_t seed,
                          const float *grad_out, float *grad_vertices, float *sample_grad_u, hipStream_t st,
                          const float *velocity, float *sample_jvp) {
    if (n_points <= 0) return 0;
    if (!s.use_bvh) {
        set_error("the shape gradient needs the scene's BVH (use_bvh)");
        return -1;
    }
    GradArgs a;
    a.points = points;
    a.n_points = n_points;
    a.g = g;
    a.seed = seed;
    a.total = (long long)n_points * g.nwalks;
    if (!walk_count_fits(a.total)) return -1;
    a.counter = launch_counter(st);
    if (!a.counter) return -1;
    a.chunk = g_tuning.chunk > 0 ? g_tuning.chunk : 128;
    a.window = window_for(s.dim);
    a.grad_out = grad_out;
    a.grad_vertices = grad_vertices;
    a.vstage = nullptr;
    a.sample_grad_u = sample_grad_u;
    a.velocity = velocity;
    a.sample_jvp = velocity ? sample_jvp : nullptr;
    long waves_needed = (long)((a.total + 63) / 64);
    int gr = (int)((waves_needed + 3) / 4);
    int cap = occupancy_grid();
    if (gr > cap) gr = cap;
    if (gr < 1) gr = 1;
    int gi = grid_of(n_points, cap);
    long lanes = (long)std::max(gr, gi) * kBlock;
    if (spill_buffer(s, lanes, &a.spill_stride, &a.spill, st) != 0) return -1;
    StartRecords sr(st);
    if (sr.alloc(s, points, n_points) != 0) return -1;
    a.init_prim = sr.prim;
    a.init_sd = sr.sd;
    a.order = sr.order;
    // reverse mode: staged terms (32 B per sample), the vertex sums, their range and slice sums
    unsigned char *vbuf = nullptr;
    VgradRange *range = nullptr;
    unsigned long long *vacc = nullptr, *vhist = nullptr;
    double *vpart = nullptr;
    const bool rev = grad_vertices && grad_out;
    if (rev) {
        const size_t stage_b = sizeof(float4) * (size_t)(s.dim - 1) * (size_t)a.total;
        const size_t acc_b = sizeof(unsigned long long) * (size_t)s.n_vertices * s.dim;
        // with LDS histograms: kVgradBlocks histograms of the sums, summed in a second pass
        const bool small = (long)s.n_vertices * s.dim <= (long)kVgradLds * kVgradWindows;
        const size_t part_b = small ? acc_b * kVgradBlocks : 0;
        if (hipMallocAsync((void **)&vbuf, stage_b + acc_b + 256 + sizeof(double) * kVgradSlices + part_b, st) !=
            hipSuccess) {
            (void)hipGetLastError();
            set_error("cannot allocate the staged gradient terms of %lld samples", a.total);
            return -1;
        }
        a.vstage = reinterpret_cast<float4 *>(vbuf);
        vacc = reinterpret_cast<unsigned long long *>(vbuf + stage_b);
        range = reinterpret_cast<VgradRange *>(vbuf + stage_b + acc_b);
        vpart = reinterpret_cast<double *>(vbuf + stage_b + acc_b + 256);
        vhist = small ? reinterpret_cast<unsigned long long *>(vbuf + stage_b + acc_b + 256 +
                                                               sizeof(double) * kVgradSlices)
                      : nullptr;
        k_vgrad_setup<<<1, 1, 0, st>>>(range, 0, a.total, s.n_vertices, s.idx, vacc, grad_vertices);
    }
    size_t tb = top_bytes(s);
    launch_point_init(s, points, n_points, sr.order, sr.prim, sr.sd, a.spill, a.spill_stride, a.counter, gi, st);
    k_zero_u64<<<1, 1, 0, st>>>(a.counter);
    if (s.stats) {
        if (s.dim == 2) k_grad_persistent<2, true><<<gr, kBlock, tb, st>>>(s, a);
        else k_grad_persistent<3, true><<<gr, kBlock, tb, st>>>(s, a);
    } else {
        if (s.dim == 2) k_grad_persistent<2, false><<<gr, kBlock, tb, st>>>(s, a);
        else k_grad_persistent<3, false><<<gr, kBlock, tb, st>>>(s, a);
    }
    if (rev) {
        // small meshes: kVgradBlocks histograms in LDS per window of the sums,
        // then summed; large ones: atomics into the sums in memory
        const long m = (long)s.n_vertices * s.dim;
        const bool lds = vhist != nullptr;
        const int gs = lds ? kVgradBlocks : grid_of((long)std::min<long long>(a.total, 1LL << 30), occupancy_grid());
        const dim3 gsz(gs, 1, lds ? vgrad_windows(m) : 1);
        const size_t hb = lds ? sizeof(unsigned long long) * (size_t)std::min<long>(m, kVgradLds) : 0;
        const int gv = grid_of(m);
        if (s.dim == 2) {
            k_vgrad_bound<2><<<dim3(kVgradSlices, 1), kBlock, 0, st>>>(a.vstage, range, vpart);
            k_vgrad_scale<<<1, 64, 0, st>>>(range, vpart);
            k_vgrad_scatter<2><<<gsz, kBlock, hb, st>>>(a.vstage, range, lds ? kVgradLds : 0, vhist);
            if (lds) k_vgrad_partials<2><<<gv, kBlock, 0, st>>>(range, vhist, kVgradBlocks);
            k_vgrad_finish<2><<<dim3(gv, 1), kBlock, 0, st>>>(range);
        } else {
            k_vgrad_bound<3><<<dim3(kVgradSlices, 1), kBlock, 0, st>>>(a.vstage, range, vpart);
            k_vgrad_scale<<<1, 64, 0, st>>>(range, vpart);
            k_vgrad_scatter<3><<<gsz, kBlock, hb, st>>>(a.vstage, range, lds ? kVgradLds : 0, vhist);
            if (lds) k_vgrad_partials<3><<<gv, kBlock, 0, st>>>(range, vhist, kVgradBlocks);
            k_vgrad_finish<3><<<dim3(gv, 1), kBlock, 0, st>>>(range);
        }
        (void)hipFreeAsync(vbuf, st);
    }
    KCHK();
    return 0;
}

// Forward mode of the shape-gradient integrand along a vertex velocity
// field: per-sample -grad u . V(exit) into a scratch buffer, then the FP64
// per-point mean / variance of k_reduce.
int launch_shape_derivative(const SceneView &s, const GradSettings &g, const float *points, int n_points,
                            int32_t seed, const float *velocity, float *mean, double *mean64, double *var64,
                            hipStream_t st) {
    if (n_points <= 0) return 0;
    float *jvp = nullptr;
    const size_t n = (size_t)n_points * (size_t)g.nwalks;
    if (hipMallocAsync((void **)&jvp, sizeof(float) * n, st) != hipSuccess) {
        set_error("cannot allocate %zu per-sample derivatives", n);
        return -1;
    }
    int rc = launch_shape_gradient(s, g, points, n_points, seed, nullptr, nullptr, nullptr, st, velocity, jvp);
    if (rc == 0 && (mean || mean64 || var64)) {
        int gr = (int)(((long)n_points * 64 + kBlock - 1) / kBlock);
        if (gr > 8192) gr = 8192;
        k_reduce<<<gr, kBlock, 0, st>>>(jvp, n_points, g.nwalks, mean, mean64, var64);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_error("kernel launch: %s", hipGetErrorString(e));
            rc = -1;
        }
    }
    (void)hipFreeAsync(jvp, st);
    return rc;
}

}  // namespace wos
