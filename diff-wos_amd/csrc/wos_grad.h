// wos_grad.h -- the shape-gradient sample loop (OursCUDA / Ours3D,
// wos/wos_grad.py:130-151, wos/wos_grad_3d.py:117-137) shared by the
// single-scene kernel (wos_kernels.hip) and the batched one (wos_batch.hip).
//
// A sample source policy Src supplies everything that differs between the
// two: the launch's sample count, queue and tuning; the mapping of a claimed
// sample slot to (item, point, sample index); per-item scene, settings,
// points, start records and seed; and what a finished sample writes.
#pragma once

#include "wos_build.h"

namespace wos {

constexpr int kGradFinishBatch = 8;   // finished samples written together (as in the walk kernel)

// Phases of a shape-gradient sample (wos/wos_grad.py:130-151 with
// wos/wos.py:59-129; wos/wos_grad_3d.py:117-137 with wos/wos3d.py:97-156).
enum : int { kPrimary = 0, kLib = 1, kSecQuery = 2, kSecWalk = 3 };

// Half the smallest bounding-box extent: the upper end of the LIB bisection
// (include/scene.cuh:225-249).
template <int D>
__device__ __forceinline__ float lib_upper(const SceneView &s) {
    float b = fminf(fabsf(s.amax[0] - s.amin[0]), fabsf(s.amax[1] - s.amin[1]));
    if (D == 3) b = fminf(b, fabsf(s.amax[2] - s.amin[2]));
    return b / 2.0f;
}

// Point y of secondary evaluation `side` on the sphere of radius R around
// c = x_b - n R (n = inward normal), from the sample's angle draws.
template <int D>
__device__ __forceinline__ void secondary_point(const float *xb, const float *n, float R, const float *dr, int side,
                                                float *y) {
    float c[3];
    for (int q = 0; q < 3; ++q) c[q] = fmaf(-n[q], R, xb[q]);
    if constexpr (D == 2) {
        const float sn = dr[0], cs = dr[1];
        const float pd[2] = {-n[1], n[0]};
        const float ss = side ? -sn : sn;
        y[0] = fmaf(R, fmaf(n[0], cs, pd[0] * ss), c[0]);
        y[1] = fmaf(R, fmaf(n[1], cs, pd[1] * ss), c[1]);
        y[2] = 0.0f;
    } else {
        const float ct = dr[0], sn = dr[1], cs = dr[2];
        float up[3] = {0.0f, 0.0f, 1.0f};
        if (!(fabsf(n[2]) < 0.9f)) { up[0] = 1.0f; up[2] = 0.0f; }
        float pd[3];
        cross3(n, up, pd);
        float l = sqrtf(dot3(pd, pd));
        pd[0] = pd[0] / l; pd[1] = pd[1] / l; pd[2] = pd[2] / l;
        float qd[3];
        cross3(n, pd, qd);
        const float r = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        const float ss = side ? -sn : sn, cc = side ? -cs : cs;
        for (int q = 0; q < 3; ++q) y[q] = fmaf(R, fmaf(ct, n[q], r * fmaf(cc, pd[q], ss * qd[q])), c[q]);
    }
}

// Tangential derivative of the boundary data on primitive f (wos/scene.py:
// 451-460, wos/wos3d.py:137-156), added to gu.
template <int D>
__device__ __forceinline__ void add_tangent_derivative(const SceneView &s, int f, float *gu) {
    if constexpr (D == 2) {
        int i0 = s.idx[2 * f], i1 = s.idx[2 * f + 1];
        float ex = s.vtx[2 * i1] - s.vtx[2 * i0], ey = s.vtx[2 * i1 + 1] - s.vtx[2 * i0 + 1];
        float len = sqrtf(dot2(ex, ey, ex, ey));
        float dv = (s.val[i1] - s.val[i0]) / len;
        gu[0] += dv * (ex / len);
        gu[1] += dv * (ey / len);
    } else {
        int i0 = s.idx[3 * f], i1 = s.idx[3 * f + 1], i2 = s.idx[3 * f + 2];
        float a[3], ab[3], ac[3];
        for (int q = 0; q < 3; ++q) {
            a[q] = s.vtx[3 * i0 + q];
            ab[q] = s.vtx[3 * i1 + q] - a[q];
            ac[q] = s.vtx[3 * i2 + q] - a[q];
        }
        float lab = sqrtf(dot3(ab, ab));
        float t[3] = {ab[0] / lab, ab[1] / lab, ab[2] / lab};
        float fn[3];
        cross3(ab, ac, fn);
        float lf = sqrtf(dot3(fn, fn));
        fn[0] = fn[0] / lf; fn[1] = fn[1] / lf; fn[2] = fn[2] / lf;
        float nn[3];
        cross3(fn, t, nn);
        float ln = sqrtf(dot3(nn, nn));
        nn[0] = nn[0] / ln; nn[1] = nn[1] / ln; nn[2] = nn[2] / ln;
        float va = s.val[i0], vb = s.val[i1], vc = s.val[i2];
        float gt = (vb - va) / lab;
        float gn = ((vc - va) - gt * dot3(ac, t)) / dot3(ac, nn);
        for (int q = 0; q < 3; ++q) gu[q] += fmaf(gt, t[q], gn * nn[q]);
    }
}


// ---------------------------------------------------------------------------
// vertex gradient in 64-bit fixed point
// ---------------------------------------------------------------------------
// A finished sample adds -w_j c to vertex j of its exit primitive, with w the
// exit point's interpolation weights (non-negative, summing to 1) and c = g
// grad u (g = dL/d mean of its point over W).  The samples stage their exit
// primitive (-1: the sample adds nothing), its t or (u, v) and c -- one
// float4 per sample in 2D, two in 3D.  The sums are then formed over each
// range of samples (a scene's launch, or one batch item) as 64-bit integers
// in units of 2^-e, e = 61 - ceil(log2 sum over the samples of max_k |c_k|)
// (no vertex component gathers more), so the gradient does not depend on
// the order of the additions.
struct VgradRange {
    long long begin, n;          // samples [begin, begin + n) of the staging buffer
    int n_vertices;
    const int *idx;              // the scene's primitive vertex indices
    unsigned long long *acc;     // n_vertices * D sums
    float *grad_vertices;        // added into
    double scale;                // set by k_vgrad_scale
};
constexpr int kVgradSlices = 64;

template <int D>
__device__ __forceinline__ void stage_vertex_grad(float4 *stage, long long slot, bool ok, int rprim, float rt0,
                                                  float rt1, float g, const float *gu) {
    const float pr = __int_as_float(ok ? rprim : -1);
    if constexpr (D == 2) {
        stage[slot] = make_float4(pr, rt0, g * gu[0], g * gu[1]);
    } else {
        stage[2 * slot] = make_float4(pr, rt0, rt1, g * gu[0]);
        stage[2 * slot + 1] = make_float4(g * gu[1], g * gu[2], 0.0f, 0.0f);
    }
}

// A staged sample: exit primitive (-1: none), weights of its vertices, c.
template <int D>
__device__ __forceinline__ int staged_sample(const float4 *stage, long long k, float *wt, float *c) {
    if constexpr (D == 2) {
        const float4 w0 = stage[k];
        wt[0] = 1.0f - w0.y;
        wt[1] = w0.y;
        c[0] = w0.z;
        c[1] = w0.w;
        return __float_as_int(w0.x);
    } else {
        const float4 w0 = stage[2 * k], w1 = stage[2 * k + 1];
        wt[0] = 1.0f - w0.y - w0.z;
        wt[1] = w0.y;
        wt[2] = w0.z;
        c[0] = w0.w;
        c[1] = w1.x;
        c[2] = w1.y;
        return __float_as_int(w0.x);
    }
}

__device__ __forceinline__ double vgrad_block_sum(double acc) {
    __shared__ double part[256];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    return part[0];
}

// One range written from the host's values (no host buffer to keep alive).
static __global__ void k_vgrad_setup(VgradRange *r, long long begin, long long n, int n_vertices, const int *idx,
                                     unsigned long long *acc, float *grad_vertices) {
    r->begin = begin;
    r->n = n;
    r->n_vertices = n_vertices;
    r->idx = idx;
    r->acc = acc;
    r->grad_vertices = grad_vertices;
    r->scale = 1.0;
}

// Range blockIdx.y, slice blockIdx.x: sum of max_k |c_k| over the slice's
// samples (part[range * kVgradSlices + slice]); also zeroes the range's sums.
template <int D>
__global__ __launch_bounds__(256) void k_vgrad_bound(const float4 *stage, VgradRange *ranges, double *part) {
    VgradRange &r = ranges[blockIdx.y];
    for (long j = (long)blockIdx.x * 256 + threadIdx.x; j < (long)r.n_vertices * D; j += (long)gridDim.x * 256)
        r.acc[j] = 0ull;
    const long long per = (r.n + kVgradSlices - 1) / kVgradSlices;
    const long long b = r.begin + blockIdx.x * per, e = min(r.begin + r.n, b + per);
    double acc = 0.0;
    for (long long k = b + threadIdx.x; k < e; k += 256) {
        float wt[3], c[3];
        if (staged_sample<D>(stage, k, wt, c) < 0) continue;
        float m = 0.0f;
        for (int q = 0; q < D; ++q) m = fmaxf(m, fabsf(c[q]));
        acc += (double)m;
    }
    const double t = vgrad_block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.y * kVgradSlices + blockIdx.x] = t;
}

// Per range (blockIdx.x): the slices added in order, then the scale.
static __global__ void k_vgrad_scale(VgradRange *ranges, const double *part) {
    if (threadIdx.x != 0) return;
    double sum = 0.0;
    for (int b = 0; b < kVgradSlices; ++b) sum += part[blockIdx.x * kVgradSlices + b];
    int e = 0;
    (void)frexp(sum, &e);
    ranges[blockIdx.x].scale = sum > 0.0 ? ldexp(1.0, 61 - e) : 1.0;
}

// Range blockIdx.y: every staged sample adds its rounded terms, through an
// LDS histogram of the sums [blockIdx.z * lds_cap, +lds_cap) when the range's
// n_vertices * D sums fit gridDim.z such windows (dynamic LDS; terms outside
// the window are left to the other blockIdx.z), else straight to the range's
// sums (blockIdx.z 0 only).  Each thread takes one contiguous span of samples
// and sums the terms of consecutive samples on one primitive in registers
// before adding them.  A histogram is then written whole to
// partial[blockIdx.x * m + window] when partial is given (summed by
// k_vgrad_partials), else added to the sums entry by entry.
constexpr int kVgradLds = 6144;
// Windows (blockIdx.z) of kVgradLds sums at most: larger meshes add their
// terms to the sums in memory.
constexpr int kVgradWindows = 4;
inline int vgrad_windows(long m) {
    const long w = (m + kVgradLds - 1) / kVgradLds;
    return w <= kVgradWindows ? (int)std::max(w, 1L) : 1;
}
template <int D>
__device__ __forceinline__ void vgrad_flush(unsigned long long *dst, long lo, long hi, const int *idx, int f,
                                            const unsigned long long *acc) {
    if (f < 0) return;
    for (int j = 0; j < D; ++j) {
        const long v = idx[(long)D * f + j];
        for (int q = 0; q < D; ++q) {
            const long e = v * D + q;
            if (e >= lo && e < hi && acc[j * D + q] != 0ull) atomicAdd(&dst[e - lo], acc[j * D + q]);
        }
    }
}
template <int D>
__global__ __launch_bounds__(256) void k_vgrad_scatter(const float4 *stage, const VgradRange *ranges, int lds_cap,
                                                       unsigned long long *partial) {
    extern __shared__ unsigned long long hv[];
    const VgradRange &r = ranges[blockIdx.y];
    const long m = (long)r.n_vertices * D;
    const bool lds = m <= (long)lds_cap * gridDim.z;
    const long lo = lds ? (long)blockIdx.z * lds_cap : 0;
    const long hi = lds ? (lo + lds_cap < m ? lo + lds_cap : m) : m;
    if (lo >= hi || (!lds && blockIdx.z > 0)) return;  // uniform over the workgroup
    if (lds)
        for (long j = threadIdx.x; j < hi - lo; j += 256) hv[j] = 0ull;
    __syncthreads();
    unsigned long long *dst = lds ? hv : r.acc;
    const double sc = r.scale;
    const long long nthr = (long long)gridDim.x * 256;
    const long long span = (r.n + nthr - 1) / nthr;
    const long long k0 = r.begin + ((long long)blockIdx.x * 256 + threadIdx.x) * span;
    const long long k1 = k0 + span < r.begin + r.n ? k0 + span : r.begin + r.n;
    int cur = -1;
    unsigned long long acc[D * D];
    for (int j = 0; j < D * D; ++j) acc[j] = 0ull;
    for (long long k = k0; k < k1; ++k) {
        float wt[3], c[3];
        const int f = staged_sample<D>(stage, k, wt, c);
        if (f < 0) continue;
        if (f != cur) {
            vgrad_flush<D>(dst, lo, hi, r.idx, cur, acc);
            cur = f;
            for (int j = 0; j < D * D; ++j) acc[j] = 0ull;
        }
        for (int j = 0; j < D; ++j)
            for (int q = 0; q < D; ++q)
                acc[j * D + q] += (unsigned long long)__double2ll_rn((double)(-(wt[j] * c[q])) * sc);
    }
    vgrad_flush<D>(dst, lo, hi, r.idx, cur, acc);
    __syncthreads();
    if (lds && partial)
        for (long j = threadIdx.x; j < hi - lo; j += 256) partial[(long)blockIdx.x * m + lo + j] = hv[j];
    else if (lds)
        for (long j = threadIdx.x; j < hi - lo; j += 256)
            if (hv[j] != 0ull) atomicAdd(&r.acc[lo + j], hv[j]);
}

// Range 0's sums from the nb histograms of k_vgrad_scatter (integer sums:
// any order gives the same bits).
template <int D>
__global__ void k_vgrad_partials(const VgradRange *ranges, const unsigned long long *partial, int nb) {
    const VgradRange &r = ranges[0];
    const long m = (long)r.n_vertices * D;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (long)gridDim.x * blockDim.x) {
        unsigned long long x = 0ull;
        for (int b = 0; b < nb; ++b) x += partial[(long)b * m + j];
        r.acc[j] = x;
    }
}

// Range blockIdx.y: sums -> grad_vertices (added into).
template <int D>
__global__ void k_vgrad_finish(const VgradRange *ranges) {
    const VgradRange &r = ranges[blockIdx.y];
    const double inv = 1.0 / r.scale;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < (long)r.n_vertices * D;
         j += (long)gridDim.x * blockDim.x)
        if (r.acc[j] != 0ull) r.grad_vertices[j] += (float)((double)(long long)r.acc[j] * inv);
}

// Shape-gradient samples, one per (point, walk), on the persistent
// machinery of the walk kernel: every closest-primitive query of a sample --
// the primary walk's steps, the ten LIB bisection queries, the two secondary
// starts and their walks' steps -- is a resumable query advanced `window`
// node visits per round, and lanes take new samples from the wave's chunk as
// theirs finish.  Queries are seeded with a primitive known to be near (the
// previous step's, or the primary exit primitive, which is within m of a LIB
// point and 2R of a secondary start), so they need no distance bound; the
// answer is the exact closest primitive either way, and the random draws
// follow the single-lane order (primary walk, angle draws, secondary walk 0,
// secondary walk 1).  STATS: count bound tests, primitive tests, queries and
// walk steps into the scene's stats (instrumented single-scene launches).
template <int D, bool STATS, class Src>
__device__ __forceinline__ void grad_samples(const Src &src, const TravStack &st, int slot, long lane_g) {
    const int lane = __lane_id();
    int item = 0;
    long long k = -1;
    int mode = kIdle, phase = kPrimary;
    float p[3] = {0.0f, 0.0f, 0.0f};
    float d = 0.0f;
    int step = 0, prim = -1;
    Pcg32 rng;
    rng.state = rng.inc = 0;
    Query q;
    q.node = q.sp = 0;
    q.best = q.thr = 0.0f;
    q.best_id = -1;
    q.nbox = q.nprim = 0;
    SceneView tv;   // the traversal's view of the lane's scene (batched launches)
    tv.nodes = nullptr;
    tv.leaves = nullptr;
    tv.n_top = 0;
    tv.dim = D;
    // primary exit record and the sample's derived quantities
    float xb[3] = {0.0f, 0.0f, 0.0f}, nb[3] = {0.0f, 0.0f, 0.0f};   // exit point, rec.n
    int rprim = -1;
    float rt0 = 0.0f, rt1 = 0.0f, u_ref = 0.0f;
    float la = 0.0f, lb = 0.0f, R = 0.0f, grad = 0.0f;
    float dr[3] = {0.0f, 0.0f, 0.0f};
    int lit = 0, side = 0;
    bool start_known = false, done_ok = false;
    float wside = -1.0f;   // sign of the sdf at the sample's start point
    long long c_next = 0, c_end = 0;
    bool exhausted = false;
    const long long total = src.total();

    for (;;) {
        // 1. completed queries
        if (mode == kReady) {
            const SceneView &s = src.scene(item);
            const GradSettings &g = src.settings(item);
            const int found = q.best_id;
            bool finish = false, ok = false;
            float u_side = 0.0f;
            bool have_u = false;
            if (phase == kPrimary || phase == kSecWalk) {
                if (!start_known) {
                    prim = found;
                    d = sqrtf(q.best);
                }
                start_known = false;
                const float eps = phase == kPrimary ? g.epsilon : g.epsilon2;
                if (d < eps && step < g.nsteps) {
                    Rec r;
                    // the primary exit's normal points to the side the walk
                    // started on (its sign test is arbitrary at depth ~0)
                    closest_record<D>(s, prim, p, r, wside);
                    if (phase == kPrimary) {
                        for (int c = 0; c < 3; ++c) { xb[c] = r.p[c]; nb[c] = r.n[c]; }
                        rprim = prim;
                        rt0 = r.t;
                        if (D == 3) { rt0 = r.uv[0]; rt1 = r.uv[1]; }
                        u_ref = dirichlet_at<D>(s, prim, r.t, r.uv[0], r.uv[1]);
                        la = 1e-3f;
                        lb = lib_upper<D>(s);
                        lit = 0;
                        phase = kLib;
                    } else {
                        u_side = dirichlet_at<D>(s, prim, r.t, r.uv[0], r.uv[1]);
                        have_u = true;
                    }
                } else if (step >= g.nsteps) {
                    if (phase == kPrimary) finish = true;
                    else { u_side = 0.0f; have_u = true; }
                } else {
                    mode = kStart;
                }
            } else if (phase == kLib) {
                const float m = (la + lb) / 2.0f;
                const float dd = fabsf(signed_dist<D>(s, found, p));
                if (fabsf(dd - m) < 1e-5f) la = m; else lb = m;
                if (++lit == 10) {
                    R = la;
                    if constexpr (D == 2) {
                        float t = (float)rng.next_f64();
                        const float lo = g.clamping * kInv2Pi;
                        t = fminf(fmaxf(t, lo), 1.0f - lo);
                        sincos_turns(t, dr[0], dr[1]);
                    } else {
                        const double z64 = 1.0 - 2.0 * rng.next_f64();
                        const double ph64 = rng.next_f64();
                        dr[0] = fminf(fmaxf((float)z64, -1.0f), 1.0f - g.clamping);
                        sincos_turns((float)ph64, dr[1], dr[2]);
                    }
                    side = 0;
                    grad = 0.0f;
                    phase = kSecQuery;
                }
            } else {   // kSecQuery: the start of secondary walk `side`
                const float sd = signed_dist<D>(s, found, p);
                if (fabsf(sd) < g.epsilon2) {
                    // wos/wos.py:96-98: the walk would stop at once
                    Rec r;
                    closest_record<D>(s, found, p, r);
                    u_side = dirichlet_at<D>(s, found, r.t, r.uv[0], r.uv[1]);
                    have_u = true;
                } else if ((sd >= 0.0f && !g.double_sided) || g.nsteps <= 0) {
                    u_side = 0.0f;   // invalid walk (include/wos.cuh:102-103, 144)
                    have_u = true;
                } else {
                    prim = found;
                    d = fabsf(sd);
                    step = 0;
                    phase = kSecWalk;
                    mode = kStart;
                }
            }
            if (have_u) {
                float P;
                if constexpr (D == 2) P = 1.0f / (dr[1] - 1.0f);
                else {
                    const float om = 1.0f - dr[0];
                    P = -1.0f / (1.41421356237309504880f * (om * sqrtf(om)));
                }
                const float num = (D == 3 || g.control_variates) ? (u_side - u_ref) : u_side;
                grad = grad + P * num / R;
                if (++side == 2) {
                    finish = true;
                    ok = true;
                } else {
                    phase = kSecQuery;
                }
            }
            if (finish) {
                done_ok = ok;
                mode = kDone;
            } else if (mode == kReady) {
                // a LIB or secondary-start query comes next: p and the seed
                if (phase == kLib) {
                    const float m = (la + lb) / 2.0f;
                    for (int c = 0; c < D; ++c) p[c] = fmaf(nb[c], m, xb[c]);
                } else {
                    const float n[3] = {-nb[0], -nb[1], D == 3 ? -nb[2] : 0.0f};
                    secondary_point<D>(xb, n, R, dr, side, p);
                }
                const float seed_d2 = prim_d2<D>(s, rprim, p);
                if constexpr (STATS) {
                    atomicAdd(&s.stats[1], 1ull);   // the seed primitive
                    atomicAdd(&s.stats[2], 1ull);
                }
                if (s.n_prims == 1) {
                    q.best = seed_d2;
                    q.best_id = rprim;
                } else {
                    query_begin_at<D>(q, s, p, __builtin_inff(), rprim, seed_d2, st, slot, lane_g);
                    mode = kQuery;
                }
            }
        }
        // finished samples write their outputs once kGradFinishBatch lanes of
        // the wave are done or idle (or nothing else is left to run)
        if (__popcll(__ballot(mode == kDone || mode == kIdle)) >= kGradFinishBatch || exhausted ||
            __ballot(mode != kDone && mode != kIdle) == 0) {
            if (mode == kDone) {
                float gu[3] = {0.0f, 0.0f, 0.0f};
                if (done_ok) {
                    const float n[3] = {-nb[0], -nb[1], D == 3 ? -nb[2] : 0.0f};
                    const float gr = grad / 2.0f;
                    for (int c = 0; c < D; ++c) gu[c] = gr * n[c];
                    add_tangent_derivative<D>(src.scene(item), rprim, gu);
                }
                src.template finish<D>(item, k, done_ok, rprim, rt0, rt1, gu);
                mode = kIdle;
                k = -1;
            }
        }
        // 2. refill idle lanes from the wave's chunk
        uint64_t idle = __ballot(mode == kIdle);
        if (idle != 0 && !exhausted) {
            int need = __popcll(idle);
            long long avail = c_end - c_next;
            long long base2 = 0, end2 = 0;
            bool fresh = false;
            if (avail < need) {
                unsigned long long b = 0;
                if (lane == 0) b = atomicAdd(src.counter(), (unsigned long long)src.chunk());
                uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
                uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
                long long bb = (long long)(((uint64_t)hi << 32) | lo);
                if (bb >= total) {
                    exhausted = true;
                } else {
                    fresh = true;
                    base2 = bb;
                    end2 = bb + src.chunk() < total ? bb + src.chunk() : total;
                }
            }
            if (mode == kIdle) {
                int r = rank_in_mask(idle);
                long long wslot = -1;   // sample slot (slot is the lane's LDS stack slot)
                if (r < avail) wslot = c_next + r;
                else if (fresh && (r - avail) < end2 - base2) wslot = base2 + (r - avail);
                if (wslot >= 0) {
                    long long i = 0, kk = 0;
                    item = src.take(wslot, i, kk);
                    const GradSettings &g = src.settings(item);
                    const int ip = src.start_prim(item, i);
                    const float sd = src.start_sd(item, i);
                    k = kk;
                    src.template load<D>(item, i, p);
                    src.bind(item, tv);
                    rng = walk_stream(src.seed(item), i, kk);
                    phase = kPrimary;
                    // include/wos.cuh:100-106 (no first-step halving here)
                    const bool outside = ip < 0 || (sd >= 0.0f && !g.double_sided);
                    if (outside || g.nsteps <= 0) {
                        src.template skip<D>(item, kk);
                        k = -1;
                    } else {
                        prim = ip;
                        d = fabsf(sd);
                        step = 0;
                        start_known = true;   // the first test uses d = |sd| as is
                        wside = sgn(sd);
                        mode = kReady;
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

        // 3. a walk step (primary or secondary): jump to the sphere and begin
        //    the query within 2d of the new point, seeded with the previous
        //    primitive (Laplace: no source term)
        if (mode == kStart) {
            const SceneView &s = src.scene(item);
            move_point<D>(p, d, rng);
            ++step;
            const float seed_d2 = prim_d2<D>(s, prim, p);
            if constexpr (STATS) {
                atomicAdd(&s.stats[1], 1ull);   // the seed primitive
                atomicAdd(&s.stats[2], 1ull);
                atomicAdd(&s.stats[3], 1ull);
            }
            if (s.n_prims == 1) {
                q.best = seed_d2;
                q.best_id = prim;
                mode = kReady;
            } else {
                query_begin_at<D>(q, s, p, bound_d2(s, 2.0f * d), prim, seed_d2, st, slot, lane_g);
                mode = kQuery;
            }
        }
        // 4. advance every pending query by up to `window` node visits
        for (int it = 0; it < src.window(); ++it) {
            if (mode == kQuery) {
                const SceneView &ts = src.trav(item, tv);
                if (query_step<D, STATS>(ts, p, q, st, slot, lane_g, ts.tol)) {
                    mode = kReady;
                    query_flush_stats<STATS>(ts, q);
                }
            }
            if (__ballot(mode == kQuery) == 0) break;
        }
    }
}

}  // namespace wos
