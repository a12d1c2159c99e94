// wos_device.h -- device-side building blocks of the MI355X walk-on-spheres
// path: counter-keyed PCG32 streams, elementary functions, closest-point
// geometry, the LBVH closest-primitive traversal and the scene view that the
// kernels receive by value.
//
// Floating point follows one canonical operation order (explicit fmaf, the
// library is built with -ffp-contract=off, IEEE division and sqrt), so a
// closest-primitive index is the same whether it comes from the brute-force
// scan or the BVH, and agrees with a CPU evaluation of the same expressions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wos {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInv2Pi = 0.15915494309189533577f;
constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kLdsStack = 8;         // traversal stack entries kept in LDS per lane
constexpr int kNodeWords2 = 4;       // float4 words per 2D internal node
constexpr int kNodeWords3 = 6;       // float4 words per 3D internal node
__host__ __device__ constexpr int node_words(int dim) { return dim == 2 ? kNodeWords2 : kNodeWords3; }

// Read-only kernel inputs seen through the constant address space: loads at
// a wave-uniform address become scalar loads (scalar cache, waited on by
// lgkmcnt), not vector loads whose wait would also wait for every vector
// load in flight.  Only for buffers no kernel writes while it runs.
typedef __attribute__((address_space(4))) const float const_f32;
__device__ __forceinline__ const const_f32 *as_constant(const float *p) { return (const const_f32 *)p; }

// base + off bytes, off an unsigned 32-bit offset (zero-extended, so the
// address needs no 64-bit arithmetic per lane)
template <typename T>
__device__ __forceinline__ const T *at_byte_offset(const T *base, uint32_t off) {
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}

// Point of walk k under W walks per point.  Launches hold fewer than 2^32
// walks (checked on the host), so this is a 32-bit unsigned division, not
// the 64-bit routine.
__device__ __forceinline__ uint32_t point_of_walk(long long k, int W) { return (uint32_t)k / (uint32_t)W; }

// ---------------------------------------------------------------------------
// random numbers
// ---------------------------------------------------------------------------

// PCG32 XSH-RR (include/sampler2.cuh:30-37).
struct Pcg32 {
    uint64_t state, inc;
    __device__ __forceinline__ uint32_t next_u32() {
        uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    // include/sampler2.cuh:20-28
    __device__ __forceinline__ void seed(uint64_t index, uint64_t initstate, uint64_t initseq) {
        state = 0;
        inc = ((initseq + index) << 1) | 1u;
        next_u32();
        state += initstate;
        next_u32();
    }
    // include/sampler2.cuh:39-44
    __device__ __forceinline__ float next_1d() {
        return __uint_as_float((next_u32() >> 9) | 0x3f800000u) - 1.0f;
    }
    // Dr.Jit next_float64: low word drawn first
    __device__ __forceinline__ double next_f64() {
        uint64_t lo = next_u32();
        uint64_t hi = next_u32();
        uint64_t v = ((lo | (hi << 32)) >> 12) | 0x3ff0000000000000ULL;
        return __longlong_as_double((long long)v) - 1.0;
    }
};

// wos/utils.py:250-256 on int32 lanes (wrapping add/shl, arithmetic shr).
__device__ __forceinline__ void tea(int32_t v0_, int32_t v1_, int32_t &o0, int32_t &o1) {
    uint32_t v0 = (uint32_t)v0_, v1 = (uint32_t)v1_, sum = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        sum += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + sum) ^ ((uint32_t)((int32_t)v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + sum) ^ ((uint32_t)((int32_t)v0 >> 5) + 0x7e95761eu);
    }
    o0 = (int32_t)v0;
    o1 = (int32_t)v1;
}

// Stream of walk k = i*W + w of point i under Solver.walk_reduce
// (wos/solver.py:54-60): TEA(seed, i) seeds a PCG32 whose sequence index is k.
__device__ __forceinline__ Pcg32 walk_stream(int32_t seed, int64_t point, int64_t k) {
    int32_t v0, v1;
    tea(seed, (int32_t)point, v0, v1);
    Pcg32 r;
    r.seed((uint64_t)k, (uint64_t)(int64_t)v0, (uint64_t)(int64_t)v1);
    return r;
}

// ---------------------------------------------------------------------------
// elementary functions: Taylor polynomials after exact / Cody-Waite reduction
// ---------------------------------------------------------------------------

__device__ __forceinline__ void sincos_poly(float x, float &s, float &c) {
    float x2 = x * x;
    float ps = fmaf(x2, -2.50521083854417187751e-08f, 2.75573192239858906526e-06f);
    ps = fmaf(x2, ps, -1.98412698412698412698e-04f);
    ps = fmaf(x2, ps, 8.33333333333333333333e-03f);
    ps = fmaf(x2, ps, -1.66666666666666666667e-01f);
    s = fmaf(x * x2, ps, x);
    float pc = fmaf(x2, 2.08767569878680989792e-09f, -2.75573192239858906526e-07f);
    pc = fmaf(x2, pc, 2.48015873015873015873e-05f);
    pc = fmaf(x2, pc, -1.38888888888888888889e-03f);
    pc = fmaf(x2, pc, 4.16666666666666666667e-02f);
    pc = fmaf(x2, pc, -0.5f);
    c = fmaf(x2, pc, 1.0f);
}

__device__ __forceinline__ void apply_quadrant(int q, float s0, float c0, float &s, float &c) {
    q &= 3;
    float a = (q & 1) ? c0 : s0;
    float b = (q & 1) ? s0 : c0;
    s = (q & 2) ? -a : a;
    c = ((q + 1) & 2) ? -b : b;
}

// sin and cos of 2*pi*u; the quadrant split of 4u is exact.
__device__ __forceinline__ void sincos_turns(float u, float &s, float &c) {
    float r = u * 4.0f;
    float q = rintf(r);
    float f = r - q;
    float x = f * 1.57079632679489661923f;
    float s0, c0;
    sincos_poly(x, s0, c0);
    apply_quadrant((int)q, s0, c0, s, c);
}

// sin x for |x| < 2^20.  k = the integer nearest x/pi, from one fma against
// 1.5 * 2^23 (the sum's lowest mantissa bit is k's parity); x = k pi + r by a
// two-part Cody-Waite reduction (pi rounded to float, then the float
// rounding of the remainder), so |r| <= pi/2 and sin x = (-1)^k sin r; sin r
// = r + r^3 p(r^2) with p of degree 3, fitted for the least maximum absolute
// error on [-pi/2, pi/2] (1.7e-8 in exact arithmetic; 1.4e-7 against the
// double-precision sine on |x| <= 1e4 as evaluated here).
__device__ __forceinline__ float sin_cw(float x) {
    const float km = fmaf(x, 0.3183098733425140380859375f, 12582912.0f);
    const float k = km - 12582912.0f;
    float r = fmaf(-k, 3.1415927410125732421875f, x);
    r = fmaf(-k, -8.742277657347585773095489e-08f, r);
    const float r2 = r * r;
    float p = fmaf(r2, 2.600053904e-06f, -1.980661473e-04f);
    p = fmaf(r2, p, 8.333017118e-03f);
    p = fmaf(r2, p, -1.666665673e-01f);
    const float sr = fmaf(r * r2, p, r);
    return __int_as_float(__float_as_int(sr) ^ (__float_as_int(km) << 31));
}

__device__ __forceinline__ float log_poly(float x) {
    uint32_t b = __float_as_uint(x);
    int e = (int)((b >> 23) & 0xff) - 127;
    float m = __uint_as_float((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356237309504880f) { m = m * 0.5f; e += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = fmaf(s2, 0.0769230769230769230769f, 0.0909090909090909090909f);
    p = fmaf(s2, p, 0.111111111111111111111f);
    p = fmaf(s2, p, 0.142857142857142857143f);
    p = fmaf(s2, p, 0.2f);
    p = fmaf(s2, p, 0.333333333333333333333f);
    float lm = fmaf(2.0f * s, s2 * p, 2.0f * s);
    float fe = (float)e;
    return fmaf(fe, 0.693147182464599609375f, fmaf(fe, -1.90465429995776e-09f, lm));
}

__device__ __forceinline__ float exp_poly(float x) {
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) return __builtin_inff();
    float k = rintf(x * 1.44269504088896340736f);
    float r = fmaf(-k, 0.693147182464599609375f, x);
    r = fmaf(-k, -1.90465429995776e-09f, r);
    float p = fmaf(r, 2.75573192239858906526e-06f, 2.48015873015873015873e-05f);
    p = fmaf(r, p, 1.98412698412698412698e-04f);
    p = fmaf(r, p, 1.38888888888888888889e-03f);
    p = fmaf(r, p, 8.33333333333333333333e-03f);
    p = fmaf(r, p, 4.16666666666666666667e-02f);
    p = fmaf(r, p, 1.66666666666666666667e-01f);
    p = fmaf(r, p, 0.5f);
    p = fmaf(r, p, 1.0f);
    p = fmaf(r, p, 1.0f);
    int ki = (int)k;
    int k1 = ki / 2, k2 = ki - k1;
    float f1 = __uint_as_float((uint32_t)(k1 + 127) << 23);
    float f2 = __uint_as_float((uint32_t)(k2 + 127) << 23);
    return (p * f1) * f2;
}

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------

__device__ __forceinline__ float dot2(float ax, float ay, float bx, float by) { return fmaf(ax, bx, ay * by); }
__device__ __forceinline__ float dot3(const float *a, const float *b) {
    return fmaf(a[0], b[0], fmaf(a[1], b[1], a[2] * b[2]));
}
__device__ __forceinline__ float clamp01(float h) { return fminf(fmaxf(h, 0.0f), 1.0f); }
__device__ __forceinline__ float sgn(float x) { return x > 0.0f ? 1.0f : -1.0f; }  // include/fwd.cuh:122-128
__device__ __forceinline__ void cross3(const float *a, const float *b, float *o) {
    o[0] = fmaf(a[1], b[2], -(a[2] * b[1]));
    o[1] = fmaf(a[2], b[0], -(a[0] * b[2]));
    o[2] = fmaf(a[0], b[1], -(a[1] * b[0]));
}

// Squared distance to segment a + h*ba, h = clamp((p-a).ba * inv, 0, 1)
// (include/util.cuh:108-114).
__device__ __forceinline__ float seg_d2(float px, float py, float ax, float ay, float bax, float bay,
                                        float inv, float &h, float &dx, float &dy) {
    float pax = px - ax, pay = py - ay;
    h = clamp01(dot2(pax, pay, bax, bay) * inv);
    dx = fmaf(-bax, h, pax);
    dy = fmaf(-bay, h, pay);
    return dot2(dx, dy, dx, dy);
}

// Ericson closest point on a triangle (include/util.cuh:22-106).  Returns the
// squared distance; bary = (weight of a, weight of b); q = closest point.
__device__ __forceinline__ float tri_closest(const float *pa, const float *pb, const float *pc,
                                             const float *x, float *q, float *bary) {
    float ab[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    float ac[3] = {pc[0] - pa[0], pc[1] - pa[1], pc[2] - pa[2]};
    float ax[3] = {x[0] - pa[0], x[1] - pa[1], x[2] - pa[2]};
    float d1 = dot3(ab, ax), d2 = dot3(ac, ax);
    float bx[3] = {x[0] - pb[0], x[1] - pb[1], x[2] - pb[2]};
    float d3 = dot3(ab, bx), d4 = dot3(ac, bx);
    float cx[3] = {x[0] - pc[0], x[1] - pc[1], x[2] - pc[2]};
    float d5 = dot3(ab, cx), d6 = dot3(ac, cx);
    float vc = fmaf(d1, d4, -(d3 * d2));
    float vb = fmaf(d5, d2, -(d1 * d6));
    float va = fmaf(d3, d6, -(d5 * d4));
    float e43 = d4 - d3, e56 = d5 - d6;
    if (d1 <= 0.0f && d2 <= 0.0f) {
        bary[0] = 1.0f; bary[1] = 0.0f;
        q[0] = pa[0]; q[1] = pa[1]; q[2] = pa[2];
    } else if (d3 >= 0.0f && d4 <= d3) {
        bary[0] = 0.0f; bary[1] = 1.0f;
        q[0] = pb[0]; q[1] = pb[1]; q[2] = pb[2];
    } else if (d6 >= 0.0f && d5 <= d6) {
        bary[0] = 0.0f; bary[1] = 0.0f;
        q[0] = pc[0]; q[1] = pc[1]; q[2] = pc[2];
    } else if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float v = d1 / (d1 - d3);
        bary[0] = 1.0f - v; bary[1] = v;
        for (int k = 0; k < 3; ++k) q[k] = fmaf(ab[k], v, pa[k]);
    } else if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float w = d2 / (d2 - d6);
        bary[0] = 1.0f - w; bary[1] = 0.0f;
        for (int k = 0; k < 3; ++k) q[k] = fmaf(ac[k], w, pa[k]);
    } else if (va <= 0.0f && e43 >= 0.0f && e56 >= 0.0f) {
        float w = e43 / (e43 + e56);
        bary[0] = 0.0f; bary[1] = 1.0f - w;
        for (int k = 0; k < 3; ++k) q[k] = fmaf(pc[k] - pb[k], w, pb[k]);
    } else {
        float denom = 1.0f / (va + vb + vc);
        float v = vb * denom, w = vc * denom;
        bary[0] = 1.0f - v - w; bary[1] = v;
        for (int k = 0; k < 3; ++k) q[k] = fmaf(ac[k], w, fmaf(ab[k], v, pa[k]));
    }
    float dd[3] = {x[0] - q[0], x[1] - q[1], x[2] - q[2]};
    return dot3(dd, dd);
}

// ---------------------------------------------------------------------------
// scene view (passed by value to every kernel)
// ---------------------------------------------------------------------------
//
// BVH node layout (binary LBVH, children's bounds stored in the parent; the
// packing kernels in wos_build.hip describe the words in detail):
//   2D node: kNodeWords2 float4, per child a capsule axis or box, plus {inv,
//            r} per child and the child ids (node index, or ~primitive for a
//            leaf whose segment is stored inline)
//   3D node: kNodeWords3 float4, per child a box and a plane slab, or the
//            leaf triangle inline with ~primitive
//   2D leaf: 2 x float4  {a.xy, (b-a).xy} {1/|b-a|^2, prim, -, -} in Morton
//            order (leaves), and again in primitive order (pleaves, whose
//            second word holds the Morton position instead of prim)
//   3D leaf: 3 x float4  {a.xyz, b.x} {b.yz, c.xy} {c.z, prim, -, -}, also
//            twice (pleaves: primitive order, Morton position in word 2)
struct SceneView {
    int dim, n_vertices, n_prims, use_bvh;
    const float *vtx;      // n_vertices * dim
    const int *idx;        // n_prims * dim
    const float *val;      // n_vertices
    const float *nrm;      // n_vertices * 2 (2D vertex normals)
    float amin[3], amax[3];
    int source_type, n_src;
    const float *src;      // source parameters (device)
    const float4 *nodes;   // n_prims - 1 internal nodes
    const float4 *leaves;  // n_prims leaves
    const float4 *pleaves; // the same leaf payload in primitive order
    float tol;             // absolute pruning margin (scale-relative)
    int max_depth;
    unsigned long long *stats;  // null, or {box tests, primitive tests, queries, walk steps}
    int n_top;             // leading nodes (depth-major order) the kernel keeps in LDS
    // near-boundary candidate grid (2D, optional): cell c lists, in
    // grid_cand[grid_start[c] .. grid_start[c + 1]), the leaves that can be
    // closest to any point of the cell; an empty range means "ask the tree"
    const uint32_t *grid_start;   // grid_nx * grid_ny + 1 entries, null when absent
    const int *grid_cand;         // leaf indices (Morton order)
    float grid_ox, grid_oy, grid_inv_cell;
    int grid_nx, grid_ny;
    // start-node grid (2D, optional): the node a query from a point of the
    // cell may start at (every primitive that can be closest lies below it)
    const int *start_node;        // st_nx * st_ny * st_nz, null when absent
    float st_ox, st_oy, st_oz, st_inv_cell;
    int st_nx, st_ny, st_nz;
    int st_packed;                // 8-byte cells (16-bit node ids, second 0: none), else int4 cells
};

// Per-lane traversal stack: the first kLdsStack entries live in LDS (lane
// stride kBlock), deeper entries spill to a global buffer (lane stride
// spill_stride).  Depth never exceeds the tree depth, which sizes the spill.
// The LDS part and the LDS copy of the top nodes are file-scope so every
// access compiles to a ds_* instruction.
static __shared__ int g_stack_node[kLdsStack * kBlock];
static __shared__ float g_stack_d2[kLdsStack * kBlock];
extern __shared__ float4 g_top_nodes[];

struct TravStack {
    int2 *spill;          // may be null when the tree fits the LDS part
    long spill_stride;
};

// Copies the first s.n_top nodes (depth-major order) to LDS; every lane of
// the workgroup must call it before any traversal.
__device__ __forceinline__ void load_top_nodes(const SceneView &s) {
    if (s.n_top > 0 && s.use_bvh) {
        const int nw = s.n_top * node_words(s.dim);
        for (int i = threadIdx.x; i < nw; i += kBlock) g_top_nodes[i] = s.nodes[i];
        __syncthreads();
    }
}

__device__ __forceinline__ float box_d2_3(const float *p, float lx, float ly, float lz, float hx, float hy,
                                          float hz) {
    float dx = fmaxf(fmaxf(lx - p[0], p[0] - hx), 0.0f);
    float dy = fmaxf(fmaxf(ly - p[1], p[1] - hy), 0.0f);
    float dz = fmaxf(fmaxf(lz - p[2], p[2] - hz), 0.0f);
    return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
}

// (d2, prim) lexicographic minimum: equal squared distances keep the
// smallest primitive index, which is what the in-order brute-force scan with
// a strict '<' returns (include/scene.cuh:102-128).  Squared distances are
// >= 0 (or NaN) and ids >= 0, so the float bits order like the values and
// the pair compares as one unsigned 64-bit key (one compare, no branch).
__device__ __forceinline__ bool better(float d2, int id, float best_d2, int best_id) {
    const uint64_t a = ((uint64_t)__float_as_uint(d2) << 32) | (uint32_t)id;
    const uint64_t b = ((uint64_t)__float_as_uint(best_d2) << 32) | (uint32_t)best_id;
    return a < b;
}

// Squared distance from p to primitive f computed from the scene arrays
// (identical to the leaf test of the same primitive).
template <int D>
__device__ __forceinline__ float prim_d2(const SceneView &s, int f, const float *p) {
    if constexpr (D == 2) {
        const float4 *lf = at_byte_offset(s.pleaves, (uint32_t)f << 5);
        const float4 g = lf[0];
        const float inv = lf[1].x;
        float h, dx, dy;
        return seg_d2(p[0], p[1], g.x, g.y, g.z, g.w, inv, h, dx, dy);
    } else {
        // the primitive-ordered copy of the leaf: three loads, no index chain
        const float4 *lf = at_byte_offset(s.pleaves, (uint32_t)f * 48u);
        const float4 w0 = lf[0], w1 = lf[1], w2 = lf[2];
        float a[3] = {w0.x, w0.y, w0.z}, b[3] = {w0.w, w1.x, w1.y}, c[3] = {w1.z, w1.w, w2.x};
        float q[3], bary[2];
        return tri_closest(a, b, c, p, q, bary);
    }
}

// Closest primitive by brute force in index order (include/scene.cuh:102-128).
template <int D>
__device__ __forceinline__ int closest_brute(const SceneView &s, const float *p, float &best_d2) {
    float best = __builtin_inff();
    int id = -1;
    for (int f = 0; f < s.n_prims; ++f) {
        float d2 = prim_d2<D>(s, f, p);
        if (d2 < best) { best = d2; id = f; }
    }
    if (s.stats) {
        atomicAdd(&s.stats[1], (unsigned long long)s.n_prims);
        atomicAdd(&s.stats[2], 1ull);
    }
    best_d2 = best;
    return id;
}

// Resumable nearest-first LBVH traversal.  The walk kernel advances every
// lane's query by a few node visits per round, so lanes with short and long
// queries stay busy; the one-shot queries loop query_step to completion.
struct Query {
    int node, sp;
    float best;        // best squared distance so far
    float thr;         // pruning threshold (a distance): children whose lower
                       // bound exceeds it cannot hold a closer primitive
    int best_id;
    unsigned nbox, nprim;   // work counters (instrumented launches only)
};

// Pruning threshold for the current best squared distance: sqrt(best) + tol,
// with the hardware square root (about 1 ulp) inflated by 2^-20.
template <int D>
__device__ __forceinline__ float prune_for(float best_d2, float tol) {
    return fmaf(__builtin_amdgcn_sqrtf(best_d2), 1.00000095367431640625f, tol);
}

// Nodes a closest-primitive query from p may start at: the cell's entry of
// the start-node grid, or the root.  A cell holds two nodes (x, y; y = -1
// when one suffices) whose subtrees together hold every primitive that can
// be closest to a point of the cell, and in z the bits of a lower bound of
// y's distance over the cell.
template <int D>
__device__ __forceinline__ int4 start_nodes_of(const SceneView &s, const float *p) {
    if (!s.start_node) return make_int4(0, -1, 0, 0);
    const float fx = (p[0] - s.st_ox) * s.st_inv_cell;
    const float fy = (p[1] - s.st_oy) * s.st_inv_cell;
    // written so that NaN coordinates fail the test
    if (!(fx >= 0.0f && fy >= 0.0f && fx < (float)s.st_nx && fy < (float)s.st_ny)) return make_int4(0, -1, 0, 0);
    // cell byte offsets fit in 32 bits: a grid holds at most (8192 + 2)^2
    // cells in 2D and (256 + 2)^3 in 3D, 16 bytes each
    uint32_t c = (uint32_t)(int)fy * (uint32_t)s.st_nx + (uint32_t)(int)fx;
    if constexpr (D == 3) {
        const float fz = (p[2] - s.st_oz) * s.st_inv_cell;
        if (!(fz >= 0.0f && fz < (float)s.st_nz)) return make_int4(0, -1, 0, 0);
        c += (uint32_t)(int)fz * (uint32_t)(s.st_nx * s.st_ny);
    }
    if (s.st_packed) {
        const uint2 w = *at_byte_offset(reinterpret_cast<const uint2 *>(s.start_node), c << 3);
        return make_int4((int)(w.x & 0xffffu), (int)(w.x >> 16), (int)w.y, 0);
    }
    return *at_byte_offset(reinterpret_cast<const int4 *>(s.start_node), c << 4);
}

// Starts a query whose answer lies within bound_d2, optionally seeded with a
// candidate primitive (seed_id, seed_d2) -- e.g. the previous step's closest
// primitive -- which tightens pruning from the root.
template <int D>
__device__ __forceinline__ void query_begin(Query &q, float bound_d2, int seed_id, float seed_d2, float tol,
                                            int start = 0) {
    q.node = start;
    q.sp = 0;
    q.best = seed_d2;
    q.best_id = seed_id;
    const float b = __builtin_amdgcn_sqrtf(bound_d2) * 1.00000095367431640625f;
    q.thr = fminf(b, prune_for<D>(seed_d2, tol));
    q.nbox = q.nprim = 0;
}

__device__ __forceinline__ void stack_push(Query &q, const TravStack &st, int slot, long spill_lane, int node,
                                           float key) {
    if (q.sp < kLdsStack) {
        g_stack_node[q.sp * kBlock + slot] = node;
        g_stack_d2[q.sp * kBlock + slot] = key;
    } else {
        st.spill[(long)(q.sp - kLdsStack) * st.spill_stride + spill_lane] = make_int2(node, __float_as_int(key));
    }
    ++q.sp;
}

// query_begin at the start nodes of p's cell: the first is visited first,
// the second (when present) waits on the stack under its cell-wide bound.
template <int D>
__device__ __forceinline__ void query_begin_from(Query &q, int4 start, float bound_d2, int seed_id, float seed_d2,
                                                 float tol, const TravStack &st, int slot, long spill_lane) {
    query_begin<D>(q, bound_d2, seed_id, seed_d2, tol, start.x);
    // no second node: -1, or 0 in packed cells (the root is never a second node)
    if (start.y > 0) stack_push(q, st, slot, spill_lane, start.y, __int_as_float(start.z));
}
template <int D>
__device__ __forceinline__ void query_begin_at(Query &q, const SceneView &s, const float *p, float bound_d2,
                                               int seed_id, float seed_d2, const TravStack &st, int slot,
                                               long spill_lane) {
    query_begin_from<D>(q, start_nodes_of<D>(s, p), bound_d2, seed_id, seed_d2, s.tol, st, slot, spill_lane);
}

// Pops until an entry survives the threshold; false when the stack is empty.
__device__ __forceinline__ bool stack_pop(Query &q, const TravStack &st, int slot, long spill_lane) {
    while (q.sp > 0) {
        --q.sp;
        int n;
        float dd;
        if (q.sp < kLdsStack) {
            n = g_stack_node[q.sp * kBlock + slot];
            dd = g_stack_d2[q.sp * kBlock + slot];
        } else {
            int2 e = st.spill[(long)(q.sp - kLdsStack) * st.spill_stride + spill_lane];
            n = e.x;
            dd = __int_as_float(e.y);
        }
        if (dd <= q.thr) {
            q.node = n;
            return true;
        }
    }
    return false;
}

// Distance lower bound of an internal child of a packed 2D node, whose word
// g is {a.xy, e.xy}: r > 0 is a capsule of radius r around the axis a + t e
// (s_axis = squared distance to the axis), r = -0.0 a box [a, a + e] (e
// rounded up).  The box distance reuses p - a from the axis distance, and
// subtracting r = -0.0 leaves it unchanged, so neither kind needs a max.
// The box arithmetic rounds within the pruning margin tol.
__device__ __forceinline__ float child_bound_2d(float px, float py, float4 g, float s_axis, float r) {
    const float ax = px - g.x, ay = py - g.y;
    const float dx = fmaxf(fmaxf(-ax, ax - g.z), 0.0f);
    const float dy = fmaxf(fmaxf(-ay, ay - g.w), 0.0f);
    const float b = fmaf(dx, dx, dy * dy);
    return __builtin_amdgcn_sqrtf(__float_as_int(r) < 0 ? b : s_axis) - r;
}

// One node visit; true when the query is complete.
//
// 2D nodes (4 x float4): words 0 and 1 hold the bound of each child as
// {a.xy, e.xy} -- a leaf's segment (e = b - a, the leaf's own words), an
// internal child's capsule axis (e = b - a) or box (lo = a, hi <= a + e) --,
// word 2 holds {inv0, r0, inv1, r1} (r = 0 leaf, r > 0 capsule radius, r =
// -0.0 box; inv = 1/|e|^2) and word 3 {id0, id1} (child node index, or
// ~primitive for a leaf).  Both children take the same arithmetic: the
// squared distance to the segment, which for a leaf is the exact primitive
// distance of a brute-force scan, and the child's lower bound, which for a
// capsule is |p - axis| - r.
template <int D, bool STATS>
__device__ __forceinline__ bool query_step(const SceneView &s, const float *p, Query &q, const TravStack &st,
                                           int slot, long spill_lane, float tol) {
    const int node = q.node;
    if constexpr (D == 2) {
        float4 n0, n1, n2, n3;
        if (node < s.n_top) {
            n0 = g_top_nodes[4 * node];
            n1 = g_top_nodes[4 * node + 1];
            n2 = g_top_nodes[4 * node + 2];
            n3 = g_top_nodes[4 * node + 3];
            // the empty asm keeps the LDS and global loads apart: merged,
            // they become one generic (flat) load for both address spaces
            __asm__ volatile("" ::: "memory");
        } else {
            // a 32-bit byte offset from the uniform base: one shift, and the
            // loads take the scalar-base + vector-offset form
            const float4 *nd = at_byte_offset(s.nodes, (uint32_t)node << 6);
            n0 = nd[0];
            n1 = nd[1];
            n2 = nd[2];
            n3 = nd[3];
        }
        const int c0 = __float_as_int(n3.x), c1 = __float_as_int(n3.y);
        const bool leaf0 = c0 < 0, leaf1 = c1 < 0;
        float h, dx, dy;
        const float s0 = seg_d2(p[0], p[1], n0.x, n0.y, n0.z, n0.w, n2.x, h, dx, dy);
        const float s1 = seg_d2(p[0], p[1], n1.x, n1.y, n1.z, n1.w, n2.z, h, dx, dy);
        if constexpr (STATS) {
            q.nbox += (leaf0 ? 0u : 1u) + (leaf1 ? 0u : 1u);
            q.nprim += (leaf0 ? 1u : 0u) + (leaf1 ? 1u : 0u);
        }
        bool u0 = leaf0 && better(s0, ~c0, q.best, q.best_id);
        if (u0) { q.best = s0; q.best_id = ~c0; }
        bool u1 = leaf1 && better(s1, ~c1, q.best, q.best_id);
        if (u1) { q.best = s1; q.best_id = ~c1; }
        // a compare and select: fminf would first canonicalise the loaded thr
        const float pr = prune_for<2>(q.best, tol);
        if ((u0 || u1) && pr < q.thr) q.thr = pr;
        const float l0 = child_bound_2d(p[0], p[1], n0, s0, n2.y), l1 = child_bound_2d(p[0], p[1], n1, s1, n2.w);
        const bool g0 = !leaf0 && l0 <= q.thr, g1 = !leaf1 && l1 <= q.thr;
        if (g0 && g1) {
            const bool near0 = l0 <= l1;
            stack_push(q, st, slot, spill_lane, near0 ? c1 : c0, near0 ? l1 : l0);
            q.node = near0 ? c0 : c1;
            return false;
        }
        if (g0 || g1) {
            q.node = g0 ? c0 : c1;
            return false;
        }
        return !stack_pop(q, st, slot, spill_lane);
    } else {
        constexpr int W = kNodeWords3;
        float4 w[6];
        if (node < s.n_top) {
#pragma unroll
            for (int k = 0; k < 6; ++k) w[k] = g_top_nodes[W * node + k];
            __asm__ volatile("" ::: "memory");
        } else {
            // 32-bit byte offset (96 B nodes; scenes hold at most 2^25 triangles)
            const float4 *nd = at_byte_offset(s.nodes, (uint32_t)node * (uint32_t)(W * sizeof(float4)));
#pragma unroll
            for (int k = 0; k < 6; ++k) w[k] = nd[k];
        }
        const int c0 = __float_as_int(w[2].w), c1 = __float_as_int(w[5].w);
        float lb[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const float4 x0 = w[3 * k], x1 = w[3 * k + 1], x2 = w[3 * k + 2];
            const int ck = k ? c1 : c0;
            if (ck < 0) {
                // leaf: the triangle, exactly as the brute-force scan tests it
                float a[3] = {x0.x, x0.y, x0.z}, b[3] = {x0.w, x1.x, x1.y}, c[3] = {x1.z, x1.w, x2.x};
                float qq[3], bary[2];
                const float d2 = tri_closest(a, b, c, p, qq, bary);
                if constexpr (STATS) q.nprim += 1u;
                if (better(d2, ~ck, q.best, q.best_id)) {
                    q.best = d2;
                    q.best_id = ~ck;
                    q.thr = fminf(q.thr, prune_for<3>(d2, tol));
                }
                lb[k] = __builtin_inff();
            } else {
                if constexpr (STATS) q.nbox += 1u;
                const float bd = box_d2_3(p, x0.x, x0.y, x0.z, x0.w, x1.x, x1.y);
                const float sd = fabsf(fmaf(x1.z, p[0], fmaf(x1.w, p[1], fmaf(x2.x, p[2], -x2.y)))) - x2.z;
                lb[k] = fmaxf(__builtin_amdgcn_sqrtf(bd), sd);
            }
        }
        const bool g0 = c0 >= 0 && lb[0] <= q.thr, g1 = c1 >= 0 && lb[1] <= q.thr;
        if (g0 && g1) {
            const bool near0 = lb[0] <= lb[1];
            stack_push(q, st, slot, spill_lane, near0 ? c1 : c0, near0 ? lb[1] : lb[0]);
            q.node = near0 ? c0 : c1;
            return false;
        }
        if (g0 || g1) {
            q.node = g0 ? c0 : c1;
            return false;
        }
        return !stack_pop(q, st, slot, spill_lane);
    }
}

// Adds a finished query's box and primitive tests to the launch counters.
template <bool STATS>
__device__ __forceinline__ void query_flush_stats(const SceneView &s, const Query &q) {
    if constexpr (STATS) {
        if (s.stats) {
            atomicAdd(&s.stats[0], (unsigned long long)q.nbox);
            atomicAdd(&s.stats[1], (unsigned long long)q.nprim);
        }
    }
}

// Closest primitive from the candidate grid: -1 when p lies outside the grid
// or in a cell without a list.  A cell of half-diagonal rho whose centre x is
// at distance D from the boundary lists every primitive within D + 2 rho (+ a
// rounding margin) of x; the closest primitive of any point p of the cell is
// within |p - x| + dist(p, boundary) <= rho + (D + rho) of x, so the
// lexicographic minimum over the list is the global one.
template <bool STATS>
__device__ __forceinline__ int grid_closest(const SceneView &s, const float *p, float &best_d2) {
    const float fx = (p[0] - s.grid_ox) * s.grid_inv_cell;
    const float fy = (p[1] - s.grid_oy) * s.grid_inv_cell;
    // written so that NaN coordinates fail the test
    if (!(fx >= 0.0f && fy >= 0.0f && fx < (float)s.grid_nx && fy < (float)s.grid_ny)) return -1;
    const int c = (int)fy * s.grid_nx + (int)fx;
    const uint32_t b = s.grid_start[c], e = s.grid_start[c + 1];
    if (b == e) return -1;
    float best = __builtin_inff();
    int best_id = 0x7fffffff;
    // four candidates per round with every load issued before the first
    // test, so a round costs two memory latencies instead of eight; indices
    // past the end repeat the last candidate, which never wins twice
    for (uint32_t j = b; j < e; j += 4) {
        int lf[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) lf[u] = s.grid_cand[j + u < e ? j + u : e - 1];
        float4 l0[4], l1[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            l0[u] = s.leaves[2 * lf[u]];
            l1[u] = s.leaves[2 * lf[u] + 1];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float h, dx, dy;
            const float d2 = seg_d2(p[0], p[1], l0[u].x, l0[u].y, l0[u].z, l0[u].w, l1[u].x, h, dx, dy);
            const int id = __float_as_int(l1[u].y);
            if (better(d2, id, best, best_id)) {
                best = d2;
                best_id = id;
            }
        }
    }
    if constexpr (STATS) {
        if (s.stats) {
            atomicAdd(&s.stats[1], (unsigned long long)(e - b));
            atomicAdd(&s.stats[2], 1ull);
        }
    }
    best_d2 = best;
    return best_id;
}

// One-shot nearest-first query.  bound_d2: squared distance within which
// the closest primitive is known to lie (infinity when unknown).  Returns the
// primitive index (-1 if none lies within the bound) and its squared distance.
template <int D>
__device__ int closest_bvh(const SceneView &s, const float *p, float bound_d2, float &best_d2,
                           const TravStack &st, int lane_slot, long spill_lane) {
    if (s.n_prims == 1) {
        best_d2 = prim_d2<D>(s, 0, p);
        if (s.stats) {
            atomicAdd(&s.stats[1], 1ull);
            atomicAdd(&s.stats[2], 1ull);
        }
        return 0;
    }
    Query q;
    query_begin_at<D>(q, s, p, bound_d2, 0x7fffffff, __builtin_inff(), st, lane_slot, spill_lane);
    while (!query_step<D, true>(s, p, q, st, lane_slot, spill_lane, s.tol)) {
    }
    query_flush_stats<true>(s, q);
    if (s.stats) atomicAdd(&s.stats[2], 1ull);
    best_d2 = q.best;
    // a leaf is accepted whatever the bound; beyond it, the answer is not
    // guaranteed to be the closest, so report nothing and let the caller widen
    return (q.best_id == 0x7fffffff || q.best > bound_d2) ? -1 : q.best_id;
}

// Closest primitive with an optional distance bound; falls back to an
// unbounded search when nothing lies inside the bound.
template <int D>
__device__ __forceinline__ int closest_prim(const SceneView &s, const float *p, float bound_d2, float &best_d2,
                                            const TravStack &st, int lane_slot, long spill_lane) {
    if (!s.use_bvh) return closest_brute<D>(s, p, best_d2);
    if constexpr (D == 2) {
        if (s.grid_start) {
            int g = grid_closest<true>(s, p, best_d2);
            if (g >= 0) return g;
        }
    }
    int id = closest_bvh<D>(s, p, bound_d2, best_d2, st, lane_slot, spill_lane);
    if (id < 0 && bound_d2 < __builtin_inff())
        id = closest_bvh<D>(s, p, __builtin_inff(), best_d2, st, lane_slot, spill_lane);
    return id;
}

// ---------------------------------------------------------------------------
// scene queries on a known closest primitive
// ---------------------------------------------------------------------------

struct Rec {
    int valid;
    float p[3], n[3];
    float t;      // 2D segment parameter
    float uv[2];  // 3D barycentrics of (b, c)
    float d;
    int prim;
};

// include/scene.cuh:148-178 (3D: only the sign of n.(p-x) is used, so the
// normal is not normalised).
template <int D>
__device__ __forceinline__ float signed_dist(const SceneView &s, int f, const float *p) {
    if constexpr (D == 2) {
        int i0 = s.idx[2 * f], i1 = s.idx[2 * f + 1];
        float ax = s.vtx[2 * i0], ay = s.vtx[2 * i0 + 1];
        float bax = s.vtx[2 * i1] - ax, bay = s.vtx[2 * i1 + 1] - ay;
        float inv = 1.0f / dot2(bax, bay, bax, bay);
        float h, dx, dy;
        float d2 = seg_d2(p[0], p[1], ax, ay, bax, bay, inv, h, dx, dy);
        float nax = s.nrm[2 * i0], nay = s.nrm[2 * i0 + 1];
        float nx = fmaf(h, s.nrm[2 * i1] - nax, nax), ny = fmaf(h, s.nrm[2 * i1 + 1] - nay, nay);
        return sqrtf(d2) * sgn(dot2(nx, ny, dx, dy));
    } else {
        int i0 = s.idx[3 * f], i1 = s.idx[3 * f + 1], i2 = s.idx[3 * f + 2];
        float a[3] = {s.vtx[3 * i0], s.vtx[3 * i0 + 1], s.vtx[3 * i0 + 2]};
        float b[3] = {s.vtx[3 * i1], s.vtx[3 * i1 + 1], s.vtx[3 * i1 + 2]};
        float c[3] = {s.vtx[3 * i2], s.vtx[3 * i2 + 1], s.vtx[3 * i2 + 2]};
        float q[3], bary[2];
        float d2 = tri_closest(a, b, c, p, q, bary);
        float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        float ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        float n[3];
        cross3(ab, ac, n);
        float px[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
        return sqrtf(d2) * sgn(dot3(n, px));
    }
}

// include/scene.cuh:180-223.  side = 0: the normal points toward p's side of
// the primitive (the sign of n.(p - a)), as the reference's closestPoint
// computes it.  side = +-1: the normal points to the side of the boundary
// where the sdf has that sign (the face orientation the sdf uses: (b-a).y,
// -(b-a).x in 2D, (b-a) x (c-a) in 3D) -- the record of a walk that ends on
// the boundary, whose final point can lie within float rounding of the
// primitive (walk exits concentrate at depth 0), where the sign test is
// arbitrary; a walk never leaves the side it started on.
template <int D>
__device__ __forceinline__ void closest_record(const SceneView &s, int f, const float *p, Rec &r,
                                               float side = 0.0f) {
    r.valid = 1;
    r.prim = f;
    if constexpr (D == 2) {
        int i0 = s.idx[2 * f], i1 = s.idx[2 * f + 1];
        float ax = s.vtx[2 * i0], ay = s.vtx[2 * i0 + 1];
        float bax = s.vtx[2 * i1] - ax, bay = s.vtx[2 * i1 + 1] - ay;
        float inv = 1.0f / dot2(bax, bay, bax, bay);
        float h, dx, dy;
        float d2 = seg_d2(p[0], p[1], ax, ay, bax, bay, inv, h, dx, dy);
        float pax = p[0] - ax, pay = p[1] - ay;
        float l = sqrtf(dot2(bay, -bax, bay, -bax));
        float nx = bay / l, ny = -bax / l;
        float sg = side != 0.0f ? side : sgn(dot2(nx, ny, pax, pay));
        r.p[0] = fmaf(h, bax, ax);
        r.p[1] = fmaf(h, bay, ay);
        r.p[2] = 0.0f;
        r.n[0] = sg * nx;
        r.n[1] = sg * ny;
        r.n[2] = 0.0f;
        r.t = h;
        r.uv[0] = r.uv[1] = 0.0f;
        r.d = sqrtf(d2);
    } else {
        int i0 = s.idx[3 * f], i1 = s.idx[3 * f + 1], i2 = s.idx[3 * f + 2];
        float a[3] = {s.vtx[3 * i0], s.vtx[3 * i0 + 1], s.vtx[3 * i0 + 2]};
        float b[3] = {s.vtx[3 * i1], s.vtx[3 * i1 + 1], s.vtx[3 * i1 + 2]};
        float c[3] = {s.vtx[3 * i2], s.vtx[3 * i2 + 1], s.vtx[3 * i2 + 2]};
        float q[3], bary[2];
        float d2 = tri_closest(a, b, c, p, q, bary);
        float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        float ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        float n[3];
        cross3(ab, ac, n);
        float l = sqrtf(dot3(n, n));
        n[0] = n[0] / l; n[1] = n[1] / l; n[2] = n[2] / l;
        float pa[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
        float sg = side != 0.0f ? side : sgn(dot3(n, pa));
        for (int k = 0; k < 3; ++k) { r.p[k] = q[k]; r.n[k] = sg * n[k]; }
        r.uv[0] = bary[1];
        r.uv[1] = 1.0f - bary[0] - bary[1];
        r.t = 0.0f;
        r.d = sqrtf(d2);
    }
}

// include/scene.cuh:251-264 (3D: b*u + c*v + a*(1-u-v), wos/utils.py:13-14)
template <int D>
__device__ __forceinline__ float dirichlet_at(const SceneView &s, int f, float t, float u, float v) {
    if constexpr (D == 2) {
        float a = s.val[s.idx[2 * f]], b = s.val[s.idx[2 * f + 1]];
        return fmaf(t, b - a, a);
    } else {
        float a = s.val[s.idx[3 * f]], b = s.val[s.idx[3 * f + 1]], c = s.val[s.idx[3 * f + 2]];
        return fmaf(b, u, fmaf(c, v, a * (1.0f - u - v)));
    }
}

// include/scene.cuh:278-301, plus type 3: sum_k c_k sin(ax_k x + ay_k y + ph_k)
__device__ __forceinline__ float source_at(const SceneView &s, float x, float y) {
    const const_f32 *q = as_constant(s.src);
    if (s.source_type == 1) {
        float ex = x - q[0], ey = y - q[1];
        float dx = ex * ex / (2.0f * q[2] * q[2]);
        float dy = ey * ey / (2.0f * q[3] * q[3]);
        return q[4] * exp_poly(-dx - dy);
    } else if (s.source_type == 2) {
        float phi = fmaf(q[0], x, fmaf(q[1], y, q[2]));
        return fmaf(q[3], sin_cw(phi), q[4]);
    } else if (s.source_type == 3) {
        float f = 0.0f;
        int K = s.n_src / 4;
        for (int k = 0; k < K; ++k) {
            float phi = fmaf(q[4 * k], x, fmaf(q[4 * k + 1], y, q[4 * k + 2]));
            f = fmaf(q[4 * k + 3], sin_cw(phi), f);
        }
        return f;
    }
    return 0.0f;
}

// ---------------------------------------------------------------------------
// interior samples of the source term (include/wos.cuh:44-88, 115-139)
// ---------------------------------------------------------------------------

// Tabulated CDF of the radial density r log(1/r) on [0, 1]: entry i is
// F((i + 1) / kGreenCdfSize) with F(x) = x^2 (1 - 2 ln x), rounded to float
// (the closed form of include/tabulated_G_cdf.cuh's table).  Every
// translation unit whose kernels sample it holds its own copy, filled from
// the host by green_cdf_host() before the first launch on a device.
constexpr int kGreenCdfSize = 10000;

// Bin of u in the table, as the reference's binary search returns it
// (include/wos.cuh:49-58): the smallest l with u < cdf[l], or the last bin
// when there is none.  Found without the search: u = F(x) is, in b = -2 ln x,
// e^-b (1 + b) = u, whose series solution b ~ t + t^2/3 + t^3/36 (t =
// sqrt(-2 ln u)) after one Newton step on ln(1 + b) - b - ln u gives x =
// e^(-b/2) within one bin of the answer for every u that next_1d can draw;
// the table entries either side confirm it (they are the pmf and the
// previous CDF value the sample needs anyway), and a guess that misses by
// more than two bins falls back to the binary search.  cdf_lo / cdf_hi:
// cdf[l - 1] (0 for l = 0) and cdf[l].
__device__ __forceinline__ int green_bin(const float *cdf, float u, float &cdf_lo, float &cdf_hi) {
    const float lu = log_poly(fmaxf(u, 1e-30f));
    const float t = sqrtf(-2.0f * lu);
    float b = t * fmaf(t, fmaf(t, 0.02777777777777777778f, 0.33333333333333333333f), 1.0f);
    const float g = log_poly(1.0f + b) - b - lu;
    b = fmaf(g * (1.0f + b), 1.0f / fmaxf(b, 1e-30f), b);
    const float x = exp_poly(-0.5f * b);
    int l = (int)(x * (float)kGreenCdfSize);
    l = l < 0 ? 0 : (l > kGreenCdfSize - 1 ? kGreenCdfSize - 1 : l);
    for (int it = 0; it < 3; ++it) {
        cdf_lo = l > 0 ? cdf[l - 1] : 0.0f;
        cdf_hi = cdf[l];
        if (l > 0 && u < cdf_lo) { --l; continue; }
        if (!(u < cdf_hi) && l < kGreenCdfSize - 1) { ++l; continue; }
        return l;
    }
    int lo = 0, hi = kGreenCdfSize - 1;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (u < cdf[mid]) hi = mid; else lo = mid + 1;
    }
    cdf_lo = lo > 0 ? cdf[lo - 1] : 0.0f;
    cdf_hi = cdf[lo];
    return lo;
}

// One interior sample of the disk of radius d around the walk position from
// the draws (u0, u1): the offset (dx, dy), the Green's function G of the
// disk at it (include/wos.cuh:44-47) and the weight w = Jacobian / pdf, so
// that the step adds f(x + offset) G w.  Uniform in the disk (w = pi d^2),
// or, with `is`, radius from the tabulated CDF with the draw reused inside
// its bin and angle 2 pi u1 (include/wos.cuh:49-88; w = |offset| / pdf).
// Returns false where the reference adds nothing (pdf <= 1e-5).
__device__ __forceinline__ bool interior_sample(float u0, float u1, float d, bool is, const float *cdf, float &dx,
                                                float &dy, float &G, float &w) {
    float pdf;
    if (is) {
        float lo, hi;
        const int l = green_bin(cdf, u0, lo, hi);
        const float pmf = hi - lo;
        const float ur = (u0 - lo) / pmf;
        const float rad = d * ((float)l + ur) / (float)kGreenCdfSize;
        pdf = pmf * (float)kGreenCdfSize / d / (2.0f * kPi);
        float sn, cs;
        sincos_turns(u1, sn, cs);
        dx = rad * cs;
        dy = rad * sn;
        const float r = sqrtf(dot2(dx, dy, dx, dy));
        G = log_poly(d / fmaxf(r, 1e-4f)) * kInv2Pi;
        w = r / pdf;
    } else {
        float sn, cs;
        sincos_turns(u0, sn, cs);
        const float r = sqrtf(u1);
        dx = (r * cs) * d;
        dy = (r * sn) * d;
        const float area = kPi * (d * d);
        pdf = 1.0f / area;
        const float rr = fmaxf(sqrtf(dot2(dx, dy, dx, dy)), 1e-4f);
        G = log_poly(d / rr) * kInv2Pi;
        w = area;
    }
    return pdf > 1e-5f;
}

// ---------------------------------------------------------------------------
// walk steps (include/wos.cuh:14-42, 115-142)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float bound_d2(const SceneView &s, float r) {
    float b = fmaf(r, 1.00001f, 2.0f * s.tol);
    return b * b;
}

template <int D>
__device__ __forceinline__ void load_point(const float *p, long i, float *q) {
    q[0] = p[D * i];
    q[1] = p[D * i + 1];
    q[2] = (D == 3) ? p[D * i + 2] : 0.0f;
}

// include/wos.cuh:115-139: one interior sample, f G w added to contrib
template <int D>
__device__ __forceinline__ void source_step(const SceneView &s, const float *p, float d, bool is, const float *cdf,
                                            Pcg32 &rng, float &contrib) {
    const float u0 = rng.next_1d(), u1 = rng.next_1d();
    float dyx, dyy, G, w;
    if (interior_sample(u0, u1, d, is, cdf, dyx, dyy, G, w))
        contrib = fmaf(source_at(s, p[0] + dyx, p[1] + dyy) * G, w, contrib);
}

template <int D>
__device__ __forceinline__ void move_point(float *p, float d, Pcg32 &rng) {
    if constexpr (D == 2) {
        float u = rng.next_1d();
        float sn, cs;
        sincos_turns(u, sn, cs);
        p[0] = fmaf(cs, d, p[0]);
        p[1] = fmaf(sn, d, p[1]);
    } else {
        float u0 = rng.next_1d(), u1 = rng.next_1d();
        float z = 1.0f - 2.0f * u0;
        float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float sn, cs;
        sincos_turns(u1, sn, cs);
        p[0] = fmaf(r * cs, d, p[0]);
        p[1] = fmaf(r * sn, d, p[1]);
        p[2] = fmaf(z, d, p[2]);
    }
}

// Adds a valid walk's Fourier basis sums bas[q] (q < K <= NMAX) to its
// point's 64-bit fixed-point accumulators acc[q * stride], in units of
// 1 / scale: integer additions, so the sums do not depend on the order in
// which walks end.
template <int NMAX>
__device__ __forceinline__ void add_basis(unsigned long long *acc, long stride, int K, double scale,
                                          const float *bas) {
#pragma unroll
    for (int q = 0; q < NMAX; ++q) {
        if (q < K) {
            const long long x = __double2ll_rn((double)bas[q] * scale);
            if (x != 0) atomicAdd(acc + (long)q * stride, (unsigned long long)x);
        }
    }
}

// ---------------------------------------------------------------------------
// vertex-value adjoint in 64-bit fixed point
// ---------------------------------------------------------------------------
// A walk that ends on primitive f adds g_i / W times its interpolation
// weights (non-negative, summing to 1) to f's vertex values, so no vertex
// gathers more than sum_i |g_i|.  The per-walk float terms are rounded to
// integer multiples of 1 / scale, scale = 2^(61 - ceil(log2 sum |g|)), and
// summed as 64-bit integers: the sums do not depend on the order of the
// additions, so the gradient is bit-reproducible.

// scale_out[0] = that scale for grad_out[0..n) (one workgroup; FP64 sum in
// a fixed order).
__device__ __forceinline__ void value_scale_block(const float *grad_out, long n, double *scale_out) {
    __shared__ double part[256];
    double acc = 0.0;
    for (long i = threadIdx.x; i < n; i += blockDim.x) acc += fabs((double)grad_out[i]);
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int e = 0;
        (void)frexp(part[0], &e);   // sum < 2^e
        scale_out[0] = part[0] > 0.0 ? ldexp(1.0, 61 - e) : 1.0;
    }
}

// One term of the value adjoint, rounded to the fixed-point unit.
__device__ __forceinline__ unsigned long long value_term(float x, double scale) {
    return (unsigned long long)__double2ll_rn((double)x * scale);
}

// Lane modes of the persistent walk kernels: no work, a step to take, a
// query in progress, a query just completed, a finished walk whose outputs
// are not yet written.
enum : int { kIdle = 0, kStart = 1, kQuery = 2, kReady = 3, kDone = 4 };

// wave helpers ---------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ int rank_in_mask(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

}  // namespace wos
