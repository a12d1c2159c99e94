/*
 * wos_oracle.c -- CPU restatement of the zihay/diff-wos walk-on-spheres hot
 * path.  TEST INFRASTRUCTURE ONLY (see wos_oracle.h): it checks the HIP path
 * and times the CPU baseline; the product never uses it.
 *
 * Build: oracle/Makefile (gcc -std=c11 -O3 -ffp-contract=off -mavx2 -mfma).
 * Contraction is off so that every fused multiply-add below is an explicit
 * fmaf(); the device code writes the same expressions in the same order.
 */
#define _GNU_SOURCE
#include "wos_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define PCG32_MULT 0x5851f42d4c957f2dULL
#define OR_PI_F    3.14159265358979323846f
#define OR_INV2PI_F 0.15915494309189533577f

/* ======================================================================= */
/* random numbers                                                           */
/* ======================================================================= */

/* include/sampler2.cuh:30-37 (PCG32 XSH-RR, O'Neill 2014) */
uint32_t or_pcg32_next_u32(or_pcg32 *r) {
    uint64_t old = r->state;
    r->state = old * PCG32_MULT + r->inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31u));
}

/* include/sampler2.cuh:20-28 */
void or_pcg32_seed(or_pcg32 *r, uint64_t index, uint64_t initstate, uint64_t initseq) {
    r->state = 0;
    r->inc = ((initseq + index) << 1) | 1u;
    or_pcg32_next_u32(r);
    r->state += initstate;
    or_pcg32_next_u32(r);
}

/* include/sampler2.cuh:39-44: 23 random mantissa bits -> [0,1) */
float or_pcg32_next_1d(or_pcg32 *r) {
    uint32_t v = (or_pcg32_next_u32(r) >> 9) | 0x3f800000u;
    float f;
    memcpy(&f, &v, 4);
    return f - 1.0f;
}

/* Dr.Jit PCG32.next_float64: 64-bit word (first draw low), 52 mantissa bits. */
double or_pcg32_next_f64(or_pcg32 *r) {
    uint64_t lo = or_pcg32_next_u32(r);
    uint64_t hi = or_pcg32_next_u32(r);
    uint64_t v = ((lo | (hi << 32)) >> 12) | 0x3ff0000000000000ULL;
    double d;
    memcpy(&d, &v, 8);
    return d - 1.0;
}

void or_pcg32_stream(uint64_t index, uint64_t initstate, uint64_t initseq, int n, uint32_t *out) {
    or_pcg32 r;
    or_pcg32_seed(&r, index, initstate, initseq);
    for (int i = 0; i < n; ++i) out[i] = or_pcg32_next_u32(&r);
}

/* wos/utils.py:250-256.  Int32 arithmetic: + and << wrap, >> is arithmetic. */
static inline uint32_t asr32(uint32_t v, int s) { return (uint32_t)((int32_t)v >> s); }
void or_tea(int32_t v0_, int32_t v1_, int rounds, int32_t *o0, int32_t *o1) {
    uint32_t v0 = (uint32_t)v0_, v1 = (uint32_t)v1_, sum = 0;
    for (int i = 0; i < rounds; ++i) {
        sum += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + sum) ^ (asr32(v1, 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + sum) ^ (asr32(v0, 5) + 0x7e95761eu);
    }
    *o0 = (int32_t)v0;
    *o1 = (int32_t)v1;
}

/* Per-walk stream of Solver.walk_reduce (wos/solver.py:54-60): point i of
 * sample k = i*W + w gets TEA(seed, i) as (initstate, initseq); Dr.Jit's
 * PCG32(size, initstate, initseq) seeds element k with index k.  The Int32 ->
 * UInt64 conversion sign-extends. */
static void walk_stream(int64_t seed, int64_t point, int64_t k, or_pcg32 *r) {
    int32_t v0, v1;
    or_tea((int32_t)seed, (int32_t)point, 4, &v0, &v1);
    or_pcg32_seed(r, (uint64_t)k, (uint64_t)(int64_t)v0, (uint64_t)(int64_t)v1);
}

/* ======================================================================= */
/* elementary functions: Taylor polynomials after exact/Cody-Waite          */
/* reduction.  Same constants and order as the device code.                 */
/* ======================================================================= */

static inline void sincos_poly(float x, float *s, float *c) {
    /* x in [-pi/4, pi/4]; Taylor to x^11 (sin) and x^12 (cos) */
    float x2 = x * x;
    float ps = fmaf(x2, -2.50521083854417187751e-08f, 2.75573192239858906526e-06f);
    ps = fmaf(x2, ps, -1.98412698412698412698e-04f);
    ps = fmaf(x2, ps, 8.33333333333333333333e-03f);
    ps = fmaf(x2, ps, -1.66666666666666666667e-01f);
    *s = fmaf(x * x2, ps, x);
    float pc = fmaf(x2, 2.08767569878680989792e-09f, -2.75573192239858906526e-07f);
    pc = fmaf(x2, pc, 2.48015873015873015873e-05f);
    pc = fmaf(x2, pc, -1.38888888888888888889e-03f);
    pc = fmaf(x2, pc, 4.16666666666666666667e-02f);
    pc = fmaf(x2, pc, -0.5f);
    *c = fmaf(x2, pc, 1.0f);
}

static inline void quadrant(int q, float s0, float c0, float *s, float *c) {
    switch (q & 3) {
    case 0: *s = s0; *c = c0; break;
    case 1: *s = c0; *c = -s0; break;
    case 2: *s = -s0; *c = -c0; break;
    default: *s = -c0; *c = s0; break;
    }
}

/* sin and cos of 2*pi*u.  4u is exact, so the quadrant split is exact. */
void or_sincos_turns(float u, float *s, float *c) {
    float r = u * 4.0f;
    float q = rintf(r);
    float f = r - q;                         /* [-0.5, 0.5] */
    float x = f * 1.57079632679489661923f;   /* [-pi/4, pi/4] */
    float s0, c0;
    sincos_poly(x, &s0, &c0);
    quadrant((int)q, s0, c0, s, c);
}

/* sin(x) for |x| < 2^20: k = the integer nearest x/pi from one fma against
 * 1.5 * 2^23 (the sum's lowest mantissa bit is k's parity), x = k pi + r by a
 * two-part Cody-Waite reduction (pi rounded to float, then the float rounding
 * of the remainder), |r| <= pi/2, sin x = (-1)^k sin r with sin r = r + r^3
 * p(r^2), p of degree 3 fitted for the least maximum absolute error on
 * [-pi/2, pi/2].  The device sin_cw performs the same operations. */
float or_sinf(float x) {
    const float km = fmaf(x, 0.3183098733425140380859375f, 12582912.0f);
    const float k = km - 12582912.0f;
    float r = fmaf(-k, 3.1415927410125732421875f, x);
    r = fmaf(-k, -8.742277657347585773095489e-08f, r);
    const float r2 = r * r;
    float p = fmaf(r2, 2.600053904e-06f, -1.980661473e-04f);
    p = fmaf(r2, p, 8.333017118e-03f);
    p = fmaf(r2, p, -1.666665673e-01f);
    const float sr = fmaf(r * r2, p, r);
    uint32_t b, kb;
    memcpy(&b, &sr, 4);
    memcpy(&kb, &km, 4);
    b ^= kb << 31;
    float out;
    memcpy(&out, &b, 4);
    return out;
}

/* log(x), x > 0 finite: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
 * log m = 2 atanh(s), s = (m-1)/(m+1), series to s^13. */
float or_logf(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    int e = (int)((b >> 23) & 0xff) - 127;
    uint32_t mb = (b & 0x007fffffu) | 0x3f800000u;
    float m;
    memcpy(&m, &mb, 4);
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

/* exp(x): x = k ln2 + r, |r| <= ln2/2, Taylor to r^9, scale by 2^k. */
float or_expf(float x) {
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) return INFINITY;
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
    /* 2^k in two factors keeps k = -127..128 representable */
    int k1 = ki / 2, k2 = ki - k1;
    uint32_t b1 = (uint32_t)(k1 + 127) << 23, b2 = (uint32_t)(k2 + 127) << 23;
    float f1, f2;
    memcpy(&f1, &b1, 4);
    memcpy(&f2, &b2, 4);
    return (p * f1) * f2;
}

/* ======================================================================= */
/* geometry (canonical forms)                                               */
/* ======================================================================= */

static inline float dot2(float ax, float ay, float bx, float by) { return fmaf(ax, bx, ay * by); }
static inline float dot3(const float *a, const float *b) { return fmaf(a[0], b[0], fmaf(a[1], b[1], a[2] * b[2])); }
static inline float clamp01(float h) { return fminf(fmaxf(h, 0.0f), 1.0f); }
static inline float signf_(float x) { return x > 0.0f ? 1.0f : -1.0f; } /* include/fwd.cuh:122-128 */

/* include/util.cuh:108-114: squared distance and parameter h.  The inverse
 * squared length is formed once per segment. */
static inline float seg_d2(float px, float py, const float *a, const float *b, float *h_out,
                           float *dx_out, float *dy_out) {
    float bax = b[0] - a[0], bay = b[1] - a[1];
    float inv = 1.0f / dot2(bax, bay, bax, bay);
    float pax = px - a[0], pay = py - a[1];
    float h = clamp01(dot2(pax, pay, bax, bay) * inv);
    float dx = fmaf(-bax, h, pax), dy = fmaf(-bay, h, pay);
    if (h_out) *h_out = h;
    if (dx_out) { *dx_out = dx; *dy_out = dy; }
    return dot2(dx, dy, dx, dy);
}

/* include/util.cuh:22-106 (Ericson, RTCD 5.1.5).  Returns squared distance;
 * bary = (weight of a, weight of b); pt = closest point. */
static float tri_closest(const float *pa, const float *pb, const float *pc, const float *x,
                         float *pt, float *bary) {
    float ab[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    float ac[3] = {pc[0] - pa[0], pc[1] - pa[1], pc[2] - pa[2]};
    float ax[3] = {x[0] - pa[0], x[1] - pa[1], x[2] - pa[2]};
    float d1 = dot3(ab, ax), d2 = dot3(ac, ax);
    float q[3];
    if (d1 <= 0.0f && d2 <= 0.0f) {
        bary[0] = 1.0f; bary[1] = 0.0f;
        q[0] = pa[0]; q[1] = pa[1]; q[2] = pa[2];
        goto done;
    }
    float bx[3] = {x[0] - pb[0], x[1] - pb[1], x[2] - pb[2]};
    float d3 = dot3(ab, bx), d4 = dot3(ac, bx);
    if (d3 >= 0.0f && d4 <= d3) {
        bary[0] = 0.0f; bary[1] = 1.0f;
        q[0] = pb[0]; q[1] = pb[1]; q[2] = pb[2];
        goto done;
    }
    float cx[3] = {x[0] - pc[0], x[1] - pc[1], x[2] - pc[2]};
    float d5 = dot3(ab, cx), d6 = dot3(ac, cx);
    if (d6 >= 0.0f && d5 <= d6) {
        bary[0] = 0.0f; bary[1] = 0.0f;
        q[0] = pc[0]; q[1] = pc[1]; q[2] = pc[2];
        goto done;
    }
    float vc = fmaf(d1, d4, -(d3 * d2));
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float v = d1 / (d1 - d3);
        bary[0] = 1.0f - v; bary[1] = v;
        for (int k = 0; k < 3; ++k) q[k] = fmaf(ab[k], v, pa[k]);
        goto done;
    }
    float vb = fmaf(d5, d2, -(d1 * d6));
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float w = d2 / (d2 - d6);
        bary[0] = 1.0f - w; bary[1] = 0.0f;
        for (int k = 0; k < 3; ++k) q[k] = fmaf(ac[k], w, pa[k]);
        goto done;
    }
    float va = fmaf(d3, d6, -(d5 * d4));
    float e43 = d4 - d3, e56 = d5 - d6;
    if (va <= 0.0f && e43 >= 0.0f && e56 >= 0.0f) {
        float w = e43 / (e43 + e56);
        bary[0] = 0.0f; bary[1] = 1.0f - w;
        for (int k = 0; k < 3; ++k) q[k] = fmaf(pc[k] - pb[k], w, pb[k]);
        goto done;
    }
    {
        float denom = 1.0f / (va + vb + vc);
        float v = vb * denom, w = vc * denom;
        bary[0] = 1.0f - v - w; bary[1] = v;
        for (int k = 0; k < 3; ++k) q[k] = fmaf(ac[k], w, fmaf(ab[k], v, pa[k]));
    }
done:;
    float dd[3] = {x[0] - q[0], x[1] - q[1], x[2] - q[2]};
    if (pt) { pt[0] = q[0]; pt[1] = q[1]; pt[2] = q[2]; }
    return dot3(dd, dd);
}

static inline void cross3(const float *a, const float *b, float *o) {
    o[0] = fmaf(a[1], b[2], -(a[2] * b[1]));
    o[1] = fmaf(a[2], b[0], -(a[0] * b[2]));
    o[2] = fmaf(a[0], b[1], -(a[1] * b[0]));
}

/* ======================================================================= */
/* scene                                                                    */
/* ======================================================================= */

void or_scene_init(or_scene *s) {
    int D = s->dim;
    for (int k = 0; k < 3; ++k) { s->aabb_min[k] = FLT_MAX; s->aabb_max[k] = -FLT_MAX; }
    for (int v = 0; v < s->n_vertices; ++v)
        for (int k = 0; k < D; ++k) {
            float x = s->vertices[v * D + k];
            s->aabb_min[k] = fminf(s->aabb_min[k], x);
            s->aabb_max[k] = fmaxf(s->aabb_max[k], x);
        }
    if (D == 2 && s->normals) {
        /* include/scene.cuh:369-380: sum of adjacent unit face normals (dy, -dx), normalised */
        memset(s->normals, 0, sizeof(float) * 2 * (size_t)s->n_vertices);
        for (int f = 0; f < s->n_prims; ++f) {
            int i0 = s->indices[2 * f], i1 = s->indices[2 * f + 1];
            float dx = s->vertices[2 * i1] - s->vertices[2 * i0];
            float dy = s->vertices[2 * i1 + 1] - s->vertices[2 * i0 + 1];
            float l = sqrtf(dot2(dy, -dx, dy, -dx));
            float nx = dy / l, ny = -dx / l;
            s->normals[2 * i0] += nx; s->normals[2 * i0 + 1] += ny;
            s->normals[2 * i1] += nx; s->normals[2 * i1 + 1] += ny;
        }
        for (int v = 0; v < s->n_vertices; ++v) {
            float nx = s->normals[2 * v], ny = s->normals[2 * v + 1];
            float l = sqrtf(dot2(nx, ny, nx, ny));
            if (l > 0.0f) { s->normals[2 * v] = nx / l; s->normals[2 * v + 1] = ny / l; }
        }
    }
}

static inline const float *V(const or_scene *s, int v) { return s->vertices + (size_t)v * s->dim; }

static float prim_d2(const or_scene *s, int f, const float *p) {
    if (s->dim == 2) {
        const int *ix = s->indices + 2 * (size_t)f;
        return seg_d2(p[0], p[1], V(s, ix[0]), V(s, ix[1]), NULL, NULL, NULL);
    } else {
        const int *ix = s->indices + 3 * (size_t)f;
        float bary[2];
        return tri_closest(V(s, ix[0]), V(s, ix[1]), V(s, ix[2]), p, NULL, bary);
    }
}

/* include/scene.cuh:102-128.  Strict '<' over increasing index keeps the
 * smallest index among equal squared distances.  With an LBVH attached the
 * tree answers (include/scene.cuh:130-146) with the same result. */
int or_closest_index(const or_scene *s, const float *p, float *out_d2) {
    if (s->bvh) return or_bvh_closest(s, s->bvh, p, out_d2);
    float best = INFINITY;
    int id = -1;
    for (int f = 0; f < s->n_prims; ++f) {
        float d2 = prim_d2(s, f, p);
        if (d2 < best) { best = d2; id = f; }
    }
    if (out_d2) *out_d2 = best;
    return id;
}

/* include/scene.cuh:148-178 */
float or_sdf(const or_scene *s, const float *p) {
    int f = or_closest_index(s, p, NULL);
    if (f < 0) return INFINITY;
    if (s->dim == 2) {
        const int *ix = s->indices + 2 * (size_t)f;
        float h, dx, dy;
        float d2 = seg_d2(p[0], p[1], V(s, ix[0]), V(s, ix[1]), &h, &dx, &dy);
        const float *na = s->normals + 2 * (size_t)ix[0], *nb = s->normals + 2 * (size_t)ix[1];
        float nx = fmaf(h, nb[0] - na[0], na[0]), ny = fmaf(h, nb[1] - na[1], na[1]);
        return sqrtf(d2) * signf_(dot2(nx, ny, dx, dy));
    } else {
        const int *ix = s->indices + 3 * (size_t)f;
        const float *a = V(s, ix[0]), *b = V(s, ix[1]), *c = V(s, ix[2]);
        float pt[3], bary[2];
        float d2 = tri_closest(a, b, c, p, pt, bary);
        float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        float ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        float n[3];
        cross3(ab, ac, n);
        float px[3] = {p[0] - pt[0], p[1] - pt[1], p[2] - pt[2]};
        /* only the sign of n.(p - x) matters; the unit normal of the
         * reference has the same sign */
        return sqrtf(d2) * signf_(dot3(n, px));
    }
}

/* include/scene.cuh:180-223.  side = 0: n toward p's side of the primitive
 * (sign of n.(p - a)), as closestPoint computes it; side = +-1: n toward the
 * side where the sdf has that sign (face orientation (b-a).y, -(b-a).x in 2D,
 * (b-a) x (c-a) in 3D) -- the record that ends a walk, whose final point may
 * lie within float rounding of the primitive, where the sign test is
 * arbitrary (a walk stays on the side it started on). */
static void record_from_prim_side(const or_scene *s, int f, const float *p, or_record *rec, float side) {
    memset(rec, 0, sizeof(*rec));
    rec->valid = 1;
    rec->prim_id = f;
    if (s->dim == 2) {
        const int *ix = s->indices + 2 * (size_t)f;
        const float *a = V(s, ix[0]), *b = V(s, ix[1]);
        float h;
        float d2 = seg_d2(p[0], p[1], a, b, &h, NULL, NULL);
        float bax = b[0] - a[0], bay = b[1] - a[1];
        float pax = p[0] - a[0], pay = p[1] - a[1];
        float l = sqrtf(dot2(bay, -bax, bay, -bax));
        float nx = bay / l, ny = -bax / l;
        float sg = side != 0.0f ? side : signf_(dot2(nx, ny, pax, pay));
        rec->p[0] = fmaf(h, bax, a[0]);
        rec->p[1] = fmaf(h, bay, a[1]);
        rec->n[0] = sg * nx;
        rec->n[1] = sg * ny;
        rec->t = h;
        rec->d = sqrtf(d2);
    } else {
        const int *ix = s->indices + 3 * (size_t)f;
        const float *a = V(s, ix[0]), *b = V(s, ix[1]), *c = V(s, ix[2]);
        float pt[3], bary[2];
        float d2 = tri_closest(a, b, c, p, pt, bary);
        float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        float ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        float n[3];
        cross3(ab, ac, n);
        float l = sqrtf(dot3(n, n));
        n[0] /= l; n[1] /= l; n[2] /= l;
        float pa[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
        float sg = side != 0.0f ? side : signf_(dot3(n, pa));
        for (int k = 0; k < 3; ++k) { rec->p[k] = pt[k]; rec->n[k] = sg * n[k]; }
        rec->uv[0] = bary[1];
        rec->uv[1] = 1.0f - bary[0] - bary[1];
        rec->d = sqrtf(d2);
    }
}

static void record_from_prim(const or_scene *s, int f, const float *p, or_record *rec) {
    record_from_prim_side(s, f, p, rec, 0.0f);
}

/* include/scene.cuh:180-223 */
void or_closest_point(const or_scene *s, const float *p, or_record *rec) {
    int f = or_closest_index(s, p, NULL);
    if (f < 0) { memset(rec, 0, sizeof(*rec)); rec->prim_id = -1; return; }
    record_from_prim(s, f, p, rec);
}

/* include/scene.cuh:225-239 */
float or_largest_inscribed_ball(const or_scene *s, const or_record *rec, float epsilon) {
    if (!rec->valid) return 0.0f;
    int D = s->dim;
    float b = INFINITY;
    for (int k = 0; k < D; ++k) b = fminf(b, fabsf(s->aabb_max[k] - s->aabb_min[k]));
    b = b / 2.0f;
    float a = epsilon;
    for (int i = 0; i < 10; ++i) {
        float m = (a + b) / 2.0f;
        float q[3] = {0, 0, 0};
        for (int k = 0; k < D; ++k) q[k] = fmaf(rec->n[k], m, rec->p[k]);
        float d = fabsf(or_sdf(s, q));
        if (fabsf(d - m) < 1e-5f) a = m; else b = m;
    }
    return a;
}

/* include/scene.cuh:251-264 (3D: b*u + c*v + a*(1-u-v), wos/utils.py:13-14) */
float or_dirichlet(const or_scene *s, const or_record *rec) {
    if (!rec->valid || rec->prim_id < 0) return 0.0f;
    if (s->dim == 2) {
        const int *ix = s->indices + 2 * (size_t)rec->prim_id;
        float a = s->values[ix[0]], b = s->values[ix[1]];
        return fmaf(rec->t, b - a, a);
    } else {
        const int *ix = s->indices + 3 * (size_t)rec->prim_id;
        float a = s->values[ix[0]], b = s->values[ix[1]], c = s->values[ix[2]];
        float u = rec->uv[0], v = rec->uv[1];
        return fmaf(b, u, fmaf(c, v, a * (1.0f - u - v)));
    }
}

/* include/scene.cuh:278-301; type 3 = Fourier sum of sinusoids whose
 * coefficients are differentiable (params: K x (ax, ay, phase, coeff)). */
float or_source(const or_scene *s, float x, float y) {
    const float *q = s->source_params;
    switch (s->source_type) {
    case 1: {
        float ex = x - q[0], ey = y - q[1];
        float dx = ex * ex / (2.0f * q[2] * q[2]);
        float dy = ey * ey / (2.0f * q[3] * q[3]);
        return q[4] * or_expf(-dx - dy);
    }
    case 2: {
        float phi = fmaf(q[0], x, fmaf(q[1], y, q[2]));
        return fmaf(q[3], or_sinf(phi), q[4]);
    }
    case 3: {
        float f = 0.0f;
        int K = s->n_source_params / 4;
        for (int k = 0; k < K; ++k) {
            float phi = fmaf(q[4 * k], x, fmaf(q[4 * k + 1], y, q[4 * k + 2]));
            f = fmaf(q[4 * k + 3], or_sinf(phi), f);
        }
        return f;
    }
    default:
        return 0.0f;
    }
}

/* Tabulated CDF of the radial density r log(1/r) on [0, 1]
 * (include/tabulated_G_cdf.cuh:5-7): entry i = F((i + 1) / N) with the
 * closed form F(x) = x^2 (1 - 2 ln x) = (integral_0^x r ln(1/r) dr) /
 * (integral_0^1 r ln(1/r) dr), evaluated in double, rounded to float. */
#define OR_GREEN_N 10000
static float g_green_cdf[OR_GREEN_N];
static pthread_once_t g_green_once = PTHREAD_ONCE_INIT;
static void green_init(void) {
    for (int i = 0; i < OR_GREEN_N; ++i) {
        double x = (double)(i + 1) / (double)OR_GREEN_N;
        g_green_cdf[i] = (float)(x * x * (1.0 - 2.0 * log(x)));
    }
}
const float *or_green_cdf(int *n) {
    pthread_once(&g_green_once, green_init);
    if (n) *n = OR_GREEN_N;
    return g_green_cdf;
}

/* WoSDevice::sample_G_radius (include/wos.cuh:49-75): binary search for the
 * smallest bin l with u < cdf[l] (clamped to the last bin), the draw reused
 * inside the bin, radius R (l + u') / N, radial pdf pmf N / R. */
float or_sample_green_radius(float u, float R, float *pdf, int *bin) {
    const float *cdf = or_green_cdf(NULL);
    int l = 0, r = OR_GREEN_N - 1;
    while (l < r) {
        int mid = (l + r) / 2;
        if (u < cdf[mid]) r = mid; else l = mid + 1;
    }
    float lo = l > 0 ? cdf[l - 1] : 0.0f;
    float pmf = cdf[l] - lo;
    float ur = (u - lo) / pmf;
    if (bin) *bin = l;
    *pdf = pmf * (float)OR_GREEN_N / R;
    return R * ((float)l + ur) / (float)OR_GREEN_N;
}

/* One interior sample of the disk of radius d (include/wos.cuh:115-139):
 * offset (dx, dy), G of the disk there (include/wos.cuh:44-47) and weight w
 * = Jacobian / pdf, so the step adds f(p + offset) G w.  Uniform in the disk
 * (include/wos.cuh:28-42), or, with is, by sample_G_2D (include/wos.cuh:
 * 77-88: radius from the table, angle 2 pi u1, pdf / 2 pi, Jacobian |offset|).
 * Returns 0 where the reference adds nothing (pdf <= 1e-5). */
int or_interior_sample(float u0, float u1, float d, int is, float *dx, float *dy, float *G, float *w) {
    float pdf;
    if (is) {
        float pr;
        float rad = or_sample_green_radius(u0, d, &pr, NULL);
        pdf = pr / (2.0f * OR_PI_F);
        float sn, cs;
        or_sincos_turns(u1, &sn, &cs);
        *dx = rad * cs;
        *dy = rad * sn;
        float r = sqrtf(dot2(*dx, *dy, *dx, *dy));
        *G = or_logf(d / fmaxf(r, 1e-4f)) * OR_INV2PI_F;
        *w = r / pdf;
    } else {
        float sn, cs;
        or_sincos_turns(u0, &sn, &cs);
        float r = sqrtf(u1);
        *dx = (r * cs) * d;
        *dy = (r * sn) * d;
        float area = OR_PI_F * (d * d);
        pdf = 1.0f / area;
        float rr = fmaxf(sqrtf(dot2(*dx, *dy, *dx, *dy)), 1e-4f);
        *G = or_logf(d / rr) * OR_INV2PI_F;
        *w = area;
    }
    return pdf > 1e-5f;
}

/* include/wos.cuh:97-145 */
void or_walk(const or_scene *s, const or_walk_params *w, const float *p0, or_pcg32 *rng,
             or_record *rec, float *basis) {
    int D = s->dim;
    float p[3] = {p0[0], p0[1], D == 3 ? p0[2] : 0.0f};
    float contrib = 0.0f;
    int K = (s->source_type == 3) ? s->n_source_params / 4 : 0;
    if (basis) for (int k = 0; k < K; ++k) basis[k] = 0.0f;
    memset(rec, 0, sizeof(*rec));
    rec->prim_id = -1;
    float d = or_sdf(s, p);
    if (d >= 0.0f && !w->double_sided) return;
    const float side = signf_(d);   /* the exit record's normal points to this side */
    d = fabsf(d);
    if (w->prevent_fd_artifacts) d *= 0.5f;
    int has_source = (D == 2) && s->source_type != 0;
    for (int i = 0; i < w->nsteps; ++i) {
        if (d < w->epsilon) {
            int f = or_closest_index(s, p, NULL);
            record_from_prim_side(s, f, p, rec, side);
            rec->contrib = contrib;
            return;
        }
        if (has_source) {
            /* include/wos.cuh:115-139: one interior sample, f G w */
            float u0 = or_pcg32_next_1d(rng), u1 = or_pcg32_next_1d(rng);
            float dyx, dyy, G, wt;
            if (or_interior_sample(u0, u1, d, w->use_IS_for_greens, &dyx, &dyy, &G, &wt)) {
                float yx = p[0] + dyx, yy = p[1] + dyy;
                contrib = fmaf(or_source(s, yx, yy) * G, wt, contrib);
                if (basis) {
                    const float *q = s->source_params;
                    for (int k = 0; k < K; ++k) {
                        float phi = fmaf(q[4 * k], yx, fmaf(q[4 * k + 1], yy, q[4 * k + 2]));
                        basis[k] = fmaf(or_sinf(phi) * G, wt, basis[k]);
                    }
                }
            }
        }
        if (D == 2) {
            float u = or_pcg32_next_1d(rng);
            float sn, cs;
            or_sincos_turns(u, &sn, &cs);
            p[0] = fmaf(cs, d, p[0]);
            p[1] = fmaf(sn, d, p[1]);
        } else {
            float u0 = or_pcg32_next_1d(rng), u1 = or_pcg32_next_1d(rng);
            float z = 1.0f - 2.0f * u0;
            float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
            float sn, cs;
            or_sincos_turns(u1, &sn, &cs);
            p[0] = fmaf(r * cs, d, p[0]);
            p[1] = fmaf(r * sn, d, p[1]);
            p[2] = fmaf(z, d, p[2]);
        }
        float d2;
        or_closest_index(s, p, &d2);
        d = sqrtf(d2);
    }
    /* not terminated: invalid record (include/wos.cuh:144) */
    rec->prim_id = -1;
}

/* WoSCUDA / WoSWithSourceCUDA single_walk value: valid ? g(exit) + contrib : 0 */
float or_walk_value(const or_scene *s, const or_record *rec) {
    if (!rec->valid) return 0.0f;
    return or_dirichlet(s, rec) + rec->contrib;
}

/* ---- threaded walk_reduce ---------------------------------------------- */
typedef struct {
    const or_scene *s;
    const or_walk_params *w;
    const float *points;
    int n_points;
    int64_t seed;
    double *mean, *var, *basis;
    float *walk_val, *rec_t;
    int *rec_prim;
    int next;               /* shared point counter */
    pthread_mutex_t lock;
} reduce_job;

/* Exponent e of the per-point basis sums, kept as 64-bit integers in units of
 * 2^-e: the same bound as the library's (one interior sample adds b G w with
 * |b| <= 1, |G w| <= R^2 (1 + ln(R / 1e-4)) < 16 R^2 for R <= 178.5, the
 * largest ball the pdf > 1e-5 guard of include/wos.cuh:136 samples; at most
 * nsteps samples per walk, nwalks walks), so the integer sum stays below 2^61
 * and does not depend on the order of the additions. */
int or_basis_exponent(int nwalks, int nsteps) {
    const double r = 178.5;
    const double per_sample = r * r * 16.0;
    const double bound = (double)(nwalks > 1 ? nwalks : 1) * (double)(nsteps > 1 ? nsteps : 1) * per_sample;
    int e = 0;
    (void)frexp(bound, &e);
    return 61 - e;
}

static void reduce_point(reduce_job *J, int i) {
    const or_scene *s = J->s;
    int W = J->w->nwalks, D = s->dim;
    int K = (s->source_type == 3) ? s->n_source_params / 4 : 0;
    float basis[64];
    long long bsum[64];
    const double bscale = ldexp(1.0, or_basis_exponent(W, J->w->nsteps));
    for (int k = 0; k < K; ++k) bsum[k] = 0;
    double sum = 0.0, sum2 = 0.0;
    for (int w = 0; w < W; ++w) {
        int64_t k = (int64_t)i * W + w;
        or_pcg32 r;
        walk_stream(J->seed, i, k, &r);
        or_record rec;
        or_walk(s, J->w, J->points + (size_t)i * D, &r, &rec, K ? basis : NULL);
        float v = or_walk_value(s, &rec);
        sum += v;
        sum2 += (double)v * v;
        if (J->walk_val) J->walk_val[k] = v;
        if (J->rec_prim) J->rec_prim[k] = rec.valid ? rec.prim_id : -1;
        if (J->rec_t) {
            if (D == 2) J->rec_t[k] = rec.t;
            else { J->rec_t[2 * k] = rec.uv[0]; J->rec_t[2 * k + 1] = rec.uv[1]; }
        }
        if (rec.valid) for (int q = 0; q < K; ++q) bsum[q] += llrint((double)basis[q] * bscale);
    }
    double m = sum / W;
    if (J->mean) J->mean[i] = m;
    if (J->var) J->var[i] = W > 1 ? (sum2 - sum * m) / (W - 1) : 0.0;
    if (J->basis) for (int q = 0; q < K; ++q) J->basis[(size_t)i * K + q] = (double)bsum[q] / bscale;
}

static void *reduce_worker(void *arg) {
    reduce_job *J = (reduce_job *)arg;
    for (;;) {
        pthread_mutex_lock(&J->lock);
        int start = J->next;
        J->next += 16;
        pthread_mutex_unlock(&J->lock);
        if (start >= J->n_points) break;
        int end = start + 16 < J->n_points ? start + 16 : J->n_points;
        for (int i = start; i < end; ++i) reduce_point(J, i);
    }
    return NULL;
}

static void run_threads(void *(*fn)(void *), void *job, int nthreads) {
    if (nthreads <= 1) { fn(job); return; }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, fn, job);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
}

/* wos/solver.py:49-65 */
void or_walk_reduce(const or_scene *s, const or_walk_params *w, const float *points,
                    int n_points, int64_t seed, int nthreads, double *mean, double *var,
                    float *walk_val, int *rec_prim, float *rec_t, double *basis) {
    reduce_job J;
    memset(&J, 0, sizeof(J));
    J.s = s; J.w = w; J.points = points; J.n_points = n_points; J.seed = seed;
    J.mean = mean; J.var = var; J.walk_val = walk_val; J.rec_prim = rec_prim;
    J.rec_t = rec_t; J.basis = basis;
    pthread_mutex_init(&J.lock, NULL);
    run_threads(reduce_worker, &J, nthreads);
    pthread_mutex_destroy(&J.lock);
}

/* Reverse mode of value = lerp/interpolate(values at exit) (include/scene.cuh:251-264) */
void or_adjoint_values(const or_scene *s, int n_points, int nwalks, const int *rec_prim,
                       const float *rec_t, const double *grad_out, double *grad_values) {
    int D = s->dim;
    for (int i = 0; i < n_points; ++i) {
        double g = grad_out[i] / nwalks;
        for (int w = 0; w < nwalks; ++w) {
            int64_t k = (int64_t)i * nwalks + w;
            int f = rec_prim[k];
            if (f < 0) continue;
            if (D == 2) {
                const int *ix = s->indices + 2 * (size_t)f;
                double t = rec_t[k];
                grad_values[ix[0]] += g * (1.0 - t);
                grad_values[ix[1]] += g * t;
            } else {
                const int *ix = s->indices + 3 * (size_t)f;
                float u = rec_t[2 * k], v = rec_t[2 * k + 1];
                grad_values[ix[0]] += g * (double)(1.0f - u - v);
                grad_values[ix[1]] += g * u;
                grad_values[ix[2]] += g * v;
            }
        }
    }
}

/* ---- shape gradient ------------------------------------------------------- */

/* WoS.single_walk at a secondary point (wos/wos.py:46-54 through the CUDA
 * walk), used inside normal_derivative: |sdf| < eps2 -> boundary value of
 * the closest point (wos/wos.py:96-98).  The CUDA walk draws no random
 * numbers in that case, so evaluating the walk unconditionally (as Dr.Jit's
 * select does) consumes the same stream. */
static float secondary_u(const or_scene *s, const or_walk_params *w2, const float *y, or_pcg32 *rng) {
    float sd = or_sdf(s, y);
    or_record rec;
    or_walk(s, w2, y, rng, &rec, NULL);
    if (fabsf(sd) < w2->epsilon) {
        or_record c;
        or_closest_point(s, y, &c);
        return or_dirichlet(s, &c);
    }
    return or_walk_value(s, &rec);
}

/* wos/wos.py:59-107 (2D) / wos/wos3d.py:97-135 (3D) normal derivative, then
 * tangent derivative (wos/scene.py:451-460, wos/wos3d.py:137-156).
 * grad_u receives the spatial gradient estimate; returns 0 if the sample
 * does not contribute. */
static int shape_sample(const or_scene *s, const or_grad_params *g, const float *p,
                        or_pcg32 *rng, or_record *rec, float *grad_u) {
    int D = s->dim;
    or_walk_params w1 = {g->nwalks, g->nsteps, g->epsilon, g->double_sided, 0, 0};
    or_walk_params w2 = {g->nwalks, g->nsteps, g->epsilon2, g->double_sided, 0, 0};
    float d0 = or_sdf(s, p);
    int inside = g->double_sided ? 1 : (d0 < 0.0f);
    or_walk(s, &w1, p, rng, rec, NULL);
    if (!rec->valid) return 0;
    float n[3] = {-rec->n[0], -rec->n[1], D == 3 ? -rec->n[2] : 0.0f};
    float u_ref = or_dirichlet(s, rec);
    float R = or_largest_inscribed_ball(s, rec, 1e-3f);
    float c[3];
    for (int k = 0; k < D; ++k) c[k] = fmaf(-n[k], R, rec->p[k]);
    float grad = 0.0f;
    if (D == 2) {
        double th = or_pcg32_next_f64(rng);   /* theta / 2pi */
        float t = (float)th;
        float lo = g->clamping * OR_INV2PI_F;
        t = fminf(fmaxf(t, lo), 1.0f - lo);
        float sn, cs;
        or_sincos_turns(t, &sn, &cs);
        float pd[2] = {-n[1], n[0]};
        for (int side = 0; side < 2; ++side) {
            float ss = side ? -sn : sn;
            float y[2];
            y[0] = fmaf(R, fmaf(n[0], cs, pd[0] * ss), c[0]);
            y[1] = fmaf(R, fmaf(n[1], cs, pd[1] * ss), c[1]);
            float u = secondary_u(s, &w2, y, rng);
            float P = 1.0f / (cs - 1.0f);
            float num = g->control_variates ? (u - u_ref) : u;
            grad = grad + P * num / R;
        }
    } else {
        double z64 = 1.0 - 2.0 * or_pcg32_next_f64(rng);
        double ph64 = or_pcg32_next_f64(rng);   /* phi / 2pi */
        float ct = fminf(fmaxf((float)z64, -1.0f), 1.0f - g->clamping);
        float up[3];
        if (fabsf(n[2]) < 0.9f) { up[0] = 0; up[1] = 0; up[2] = 1; }
        else { up[0] = 1; up[1] = 0; up[2] = 0; }
        float pd[3];
        cross3(n, up, pd);
        float l = sqrtf(dot3(pd, pd));
        pd[0] /= l; pd[1] /= l; pd[2] /= l;
        float qd[3];
        cross3(n, pd, qd);
        float r = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        float sn, cs;
        or_sincos_turns((float)ph64, &sn, &cs);
        float om = 1.0f - ct;
        float P = -1.0f / (1.41421356237309504880f * (om * sqrtf(om)));
        for (int side = 0; side < 2; ++side) {
            float ss = side ? -sn : sn, cc = side ? -cs : cs;   /* phi + pi */
            float y[3];
            for (int k = 0; k < 3; ++k)
                y[k] = fmaf(R, fmaf(ct, n[k], r * fmaf(cc, pd[k], ss * qd[k])), c[k]);
            float u = secondary_u(s, &w2, y, rng);
            grad = grad + P * (u - u_ref) / R;
        }
    }
    grad = grad / 2.0f;
    for (int k = 0; k < D; ++k) grad_u[k] = grad * n[k];
    /* tangential part */
    if (D == 2) {
        const int *ix = s->indices + 2 * (size_t)rec->prim_id;
        const float *v0 = V(s, ix[0]), *v1 = V(s, ix[1]);
        float ex = v1[0] - v0[0], ey = v1[1] - v0[1];
        float len = sqrtf(dot2(ex, ey, ex, ey));
        float dv = (s->values[ix[1]] - s->values[ix[0]]) / len;
        grad_u[0] += dv * (ex / len);
        grad_u[1] += dv * (ey / len);
    } else {
        const int *ix = s->indices + 3 * (size_t)rec->prim_id;
        const float *a = V(s, ix[0]), *b = V(s, ix[1]), *cc = V(s, ix[2]);
        float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        float ac[3] = {cc[0] - a[0], cc[1] - a[1], cc[2] - a[2]};
        float lab = sqrtf(dot3(ab, ab));
        float t[3] = {ab[0] / lab, ab[1] / lab, ab[2] / lab};
        float fn[3];
        cross3(ab, ac, fn);
        float lf = sqrtf(dot3(fn, fn));
        fn[0] /= lf; fn[1] /= lf; fn[2] /= lf;
        float nn[3];
        cross3(fn, t, nn);
        float ln = sqrtf(dot3(nn, nn));
        nn[0] /= ln; nn[1] /= ln; nn[2] /= ln;
        float va = s->values[ix[0]], vb = s->values[ix[1]], vc = s->values[ix[2]];
        float gt = (vb - va) / lab;
        float gn = ((vc - va) - gt * dot3(ac, t)) / dot3(ac, nn);
        for (int k = 0; k < 3; ++k) grad_u[k] += fmaf(gt, t[k], gn * nn[k]);
    }
    return inside;
}

typedef struct {
    const or_scene *s;
    const or_grad_params *g;
    const float *points;
    int n_points;
    int64_t seed;
    const double *grad_out;
    double *grad_vertices;   /* shared, guarded */
    double *sample_grad_u;
    int next;
    pthread_mutex_t lock;
} shape_job;

static void *shape_worker(void *arg) {
    shape_job *J = (shape_job *)arg;
    const or_scene *s = J->s;
    int D = s->dim, W = J->g->nwalks;
    double *acc = (double *)calloc((size_t)s->n_vertices * D, sizeof(double));
    for (;;) {
        pthread_mutex_lock(&J->lock);
        int start = J->next;
        J->next += 4;
        pthread_mutex_unlock(&J->lock);
        if (start >= J->n_points) break;
        int end = start + 4 < J->n_points ? start + 4 : J->n_points;
        for (int i = start; i < end; ++i) {
            double gi = J->grad_out ? J->grad_out[i] / W : 0.0;
            for (int w = 0; w < W; ++w) {
                int64_t k = (int64_t)i * W + w;
                or_pcg32 r;
                walk_stream(J->seed, i, k, &r);
                or_record rec;
                float gu[3] = {0, 0, 0};
                int ok = shape_sample(s, J->g, J->points + (size_t)i * D, &r, &rec, gu);
                if (J->sample_grad_u)
                    for (int q = 0; q < D; ++q) J->sample_grad_u[k * D + q] = ok ? gu[q] : 0.0;
                if (!ok) continue;
                /* res = -(grad_u . (p - detach(p))), p = lerp/interpolate of vertices */
                if (D == 2) {
                    const int *ix = s->indices + 2 * (size_t)rec.prim_id;
                    double t = rec.t;
                    for (int q = 0; q < 2; ++q) {
                        acc[2 * (size_t)ix[0] + q] += -gi * (1.0 - t) * gu[q];
                        acc[2 * (size_t)ix[1] + q] += -gi * t * gu[q];
                    }
                } else {
                    const int *ix = s->indices + 3 * (size_t)rec.prim_id;
                    double u = rec.uv[0], v = rec.uv[1];
                    double wa = (double)(1.0f - rec.uv[0] - rec.uv[1]);
                    for (int q = 0; q < 3; ++q) {
                        acc[3 * (size_t)ix[0] + q] += -gi * wa * gu[q];
                        acc[3 * (size_t)ix[1] + q] += -gi * u * gu[q];
                        acc[3 * (size_t)ix[2] + q] += -gi * v * gu[q];
                    }
                }
            }
        }
    }
    pthread_mutex_lock(&J->lock);
    if (J->grad_vertices)
        for (size_t q = 0; q < (size_t)s->n_vertices * D; ++q) J->grad_vertices[q] += acc[q];
    pthread_mutex_unlock(&J->lock);
    free(acc);
    return NULL;
}

void or_shape_gradient(const or_scene *s, const or_grad_params *g, const float *points,
                       int n_points, int64_t seed, const double *grad_out, int nthreads,
                       double *grad_vertices, double *sample_grad_u) {
    shape_job J;
    memset(&J, 0, sizeof(J));
    J.s = s; J.g = g; J.points = points; J.n_points = n_points; J.seed = seed;
    J.grad_out = grad_out; J.grad_vertices = grad_vertices; J.sample_grad_u = sample_grad_u;
    pthread_mutex_init(&J.lock, NULL);
    run_threads(shape_worker, &J, nthreads);
    pthread_mutex_destroy(&J.lock);
}

/* ======================================================================= */
/* LBVH restatement                                                         */
/* ======================================================================= */

static inline uint32_t expand2(uint32_t x) {   /* 16 bits -> even bit positions */
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
static inline uint32_t expand3(uint32_t x) {   /* 10 bits -> every third bit */
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

uint32_t or_morton(int dim, const float *u) {
    if (dim == 2) {
        uint32_t x = (uint32_t)fminf(fmaxf(u[0] * 65536.0f, 0.0f), 65535.0f);
        uint32_t y = (uint32_t)fminf(fmaxf(u[1] * 65536.0f, 0.0f), 65535.0f);
        return expand2(x) | (expand2(y) << 1);
    } else {
        uint32_t x = (uint32_t)fminf(fmaxf(u[0] * 1024.0f, 0.0f), 1023.0f);
        uint32_t y = (uint32_t)fminf(fmaxf(u[1] * 1024.0f, 0.0f), 1023.0f);
        uint32_t z = (uint32_t)fminf(fmaxf(u[2] * 1024.0f, 0.0f), 1023.0f);
        return (expand3(x) << 2) | (expand3(y) << 1) | expand3(z);
    }
}

typedef struct { uint32_t code; int id; } code_pair;
static int cmp_pair(const void *a, const void *b) {
    const code_pair *x = (const code_pair *)a, *y = (const code_pair *)b;
    if (x->code != y->code) return x->code < y->code ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id);
}

static inline int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }
static int delta(const uint32_t *codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a == b) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(a ^ b);
}

/* Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees and
 * k-d Trees" (HPG 2012), Figure 4. */
int or_lbvh_build(const or_scene *s, uint32_t *codes_sorted, int *prim_sorted, int *children,
                  int *parent_internal, int *parent_leaf, int *max_depth) {
    int n = s->n_prims, D = s->dim;
    if (n <= 0) return -1;
    float cmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, cmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float *cent = (float *)malloc(sizeof(float) * 3 * (size_t)n);
    for (int f = 0; f < n; ++f) {
        for (int k = 0; k < D; ++k) {
            float lo = FLT_MAX, hi = -FLT_MAX;
            for (int c = 0; c < D; ++c) {
                float x = s->vertices[(size_t)s->indices[(size_t)f * D + c] * D + k];
                lo = fminf(lo, x);
                hi = fmaxf(hi, x);
            }
            float m = (lo + hi) * 0.5f;
            cent[3 * (size_t)f + k] = m;
            cmin[k] = fminf(cmin[k], m);
            cmax[k] = fmaxf(cmax[k], m);
        }
    }
    float inv[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) {
        float e = cmax[k] - cmin[k];
        inv[k] = e > 0.0f ? 1.0f / e : 0.0f;
    }
    code_pair *pairs = (code_pair *)malloc(sizeof(code_pair) * (size_t)n);
    for (int f = 0; f < n; ++f) {
        float u[3] = {0.5f, 0.5f, 0.5f};
        for (int k = 0; k < D; ++k) u[k] = inv[k] > 0.0f ? (cent[3 * (size_t)f + k] - cmin[k]) * inv[k] : 0.5f;
        pairs[f].code = or_morton(D, u);
        pairs[f].id = f;
    }
    qsort(pairs, (size_t)n, sizeof(code_pair), cmp_pair);
    uint32_t *codes = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    for (int j = 0; j < n; ++j) {
        codes[j] = pairs[j].code;
        if (codes_sorted) codes_sorted[j] = pairs[j].code;
        if (prim_sorted) prim_sorted[j] = pairs[j].id;
    }
    if (parent_leaf) for (int j = 0; j < n; ++j) parent_leaf[j] = -1;
    if (parent_internal && n > 1) parent_internal[0] = -1;
    for (int i = 0; i < n - 1; ++i) {
        int dd = delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1) > 0 ? 1 : -1;
        int dmin = delta(codes, n, i, i - dd);
        int lmax = 2;
        while (delta(codes, n, i, i + lmax * dd) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (delta(codes, n, i, i + (l + t) * dd) > dmin) l += t;
        int j = i + l * dd;
        int dnode = delta(codes, n, i, j);
        int lo = 0, hi = l - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (delta(codes, n, i, i + mid * dd) > dnode) lo = mid; else hi = mid - 1;
        }
        int gamma = i + lo * dd + (dd < 0 ? -1 : 0);
        int first = i < j ? i : j, last = i < j ? j : i;
        int left = (first == gamma) ? ~gamma : gamma;
        int right = (last == gamma + 1) ? ~(gamma + 1) : gamma + 1;
        children[2 * i] = left;
        children[2 * i + 1] = right;
        if (left < 0) { if (parent_leaf) parent_leaf[~left] = i; }
        else if (parent_internal) parent_internal[left] = i;
        if (right < 0) { if (parent_leaf) parent_leaf[~right] = i; }
        else if (parent_internal) parent_internal[right] = i;
    }
    if (max_depth && parent_leaf && parent_internal) {
        int md = 0;
        for (int j = 0; j < n; ++j) {
            int dep = 0, x = parent_leaf[j];
            while (x >= 0) { dep++; x = parent_internal[x]; }
            if (dep > md) md = dep;
        }
        *max_depth = md;
    }
    free(cent);
    free(pairs);
    free(codes);
    return 0;
}

/* ---- CPU LBVH queries ----------------------------------------------------- */
struct or_bvh {
    int n;              /* primitives */
    int *children;      /* 2 (n - 1): internal node, or ~sorted position of a leaf */
    int *prim_sorted;   /* sorted position -> primitive */
    float *box;         /* per internal node: lo[3], hi[3] */
    float tol;          /* pruning margin (scale-relative) */
};

static void bvh_box_of(const or_scene *s, const or_bvh *b, int c, float *lo, float *hi) {
    int D = s->dim;
    for (int k = 0; k < 3; ++k) { lo[k] = 0.0f; hi[k] = 0.0f; }
    if (c < 0) {
        int f = b->prim_sorted[~c];
        for (int k = 0; k < D; ++k) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
        for (int v = 0; v < D; ++v)
            for (int k = 0; k < D; ++k) {
                float x = V(s, s->indices[(size_t)f * D + v])[k];
                lo[k] = fminf(lo[k], x);
                hi[k] = fmaxf(hi[k], x);
            }
    } else {
        for (int k = 0; k < 3; ++k) { lo[k] = b->box[6 * (size_t)c + k]; hi[k] = b->box[6 * (size_t)c + 3 + k]; }
    }
}

/* boxes bottom-up: children before parents in a post-order walk */
static void bvh_fit(const or_scene *s, or_bvh *b, int node, int depth) {
    float lo[3], hi[3], l2[3], h2[3];
    for (int j = 0; j < 2; ++j) {
        int c = b->children[2 * (size_t)node + j];
        if (c >= 0 && depth < 4096) bvh_fit(s, b, c, depth + 1);
    }
    bvh_box_of(s, b, b->children[2 * (size_t)node], lo, hi);
    bvh_box_of(s, b, b->children[2 * (size_t)node + 1], l2, h2);
    for (int k = 0; k < 3; ++k) {
        b->box[6 * (size_t)node + k] = fminf(lo[k], l2[k]);
        b->box[6 * (size_t)node + 3 + k] = fmaxf(hi[k], h2[k]);
    }
}

or_bvh *or_bvh_create(const or_scene *s) {
    int n = s->n_prims;
    if (n <= 0) return NULL;
    or_bvh *b = (or_bvh *)calloc(1, sizeof(or_bvh));
    b->n = n;
    b->prim_sorted = (int *)malloc(sizeof(int) * (size_t)n);
    int m = n > 1 ? n - 1 : 1;
    b->children = (int *)malloc(sizeof(int) * 2 * (size_t)m);
    b->box = (float *)calloc(6 * (size_t)m, sizeof(float));
    int *pi = (int *)malloc(sizeof(int) * (size_t)m), *pl = (int *)malloc(sizeof(int) * (size_t)n);
    int depth = 0;
    if (or_lbvh_build(s, NULL, b->prim_sorted, b->children, pi, pl, &depth) != 0) {
        free(pi); free(pl); or_bvh_destroy(b);
        return NULL;
    }
    free(pi);
    free(pl);
    float scale = 1.0f;
    for (int k = 0; k < s->dim; ++k) scale = fmaxf(scale, fmaxf(fabsf(s->aabb_min[k]), fabsf(s->aabb_max[k])));
    b->tol = 1e-6f * scale;
    if (n > 1) bvh_fit(s, b, 0, 0);
    return b;
}

void or_bvh_destroy(or_bvh *b) {
    if (!b) return;
    free(b->children);
    free(b->prim_sorted);
    free(b->box);
    free(b);
}

/* distance from p to a box, in double (a lower bound of every contained
 * primitive's distance up to the pruning margin) */
static double box_dist(const float *lo, const float *hi, const float *p, int D) {
    double s2 = 0.0;
    for (int k = 0; k < D; ++k) {
        double d = fmax(fmax((double)lo[k] - p[k], (double)p[k] - hi[k]), 0.0);
        s2 += d * d;
    }
    return sqrt(s2);
}

int or_bvh_closest(const or_scene *s, const or_bvh *b, const float *p, float *out_d2) {
    float best = INFINITY;
    int id = INT32_MAX;
    if (b->n == 1) {
        best = prim_d2(s, 0, p);
        id = 0;
    } else {
        int stack_node[256];
        double stack_key[256];
        int sp = 0;
        stack_node[sp] = 0;
        stack_key[sp++] = 0.0;
        while (sp > 0) {
            --sp;
            int node = stack_node[sp];
            if (stack_key[sp] > sqrt((double)best) + b->tol) continue;
            int cn[2];
            double key[2];
            for (int j = 0; j < 2; ++j) {
                int c = b->children[2 * (size_t)node + j];
                cn[j] = -1;
                key[j] = INFINITY;
                if (c < 0) {
                    int f = b->prim_sorted[~c];
                    float d2 = prim_d2(s, f, p);
                    if (d2 < best || (d2 == best && f < id)) { best = d2; id = f; }
                } else {
                    float lo[3], hi[3];
                    bvh_box_of(s, b, c, lo, hi);
                    cn[j] = c;
                    key[j] = box_dist(lo, hi, p, s->dim);
                }
            }
            /* far child first, so the near one is popped next */
            int near = key[0] <= key[1] ? 0 : 1;
            for (int t = 0; t < 2; ++t) {
                int j = t == 0 ? 1 - near : near;
                if (cn[j] >= 0 && key[j] <= sqrt((double)best) + b->tol) {
                    if (sp == 256) {   /* deeper than any tree of 32-bit codes: answer by the scan */
                        const or_scene s2 = {s->dim, s->n_vertices, s->n_prims, s->vertices, s->indices,
                                             s->values, s->normals, {0, 0, 0}, {0, 0, 0}, 0, NULL, 0, NULL};
                        return or_closest_index(&s2, p, out_d2);
                    }
                    stack_node[sp] = cn[j];
                    stack_key[sp++] = key[j];
                }
            }
        }
    }
    if (out_d2) *out_d2 = best;
    return id == INT32_MAX ? -1 : id;
}

/* ======================================================================= */
/* batch helpers                                                            */
/* ======================================================================= */
void or_batch_elementary(int fn, const float *in, int n, float *out) {
    for (int i = 0; i < n; ++i) {
        float s, c;
        switch (fn) {
        case 0: or_sincos_turns(in[i], &s, &c); out[i] = s; break;
        case 1: or_sincos_turns(in[i], &s, &c); out[i] = c; break;
        case 2: out[i] = or_sinf(in[i]); break;
        case 3: out[i] = or_logf(in[i]); break;
        default: out[i] = or_expf(in[i]); break;
        }
    }
}

void or_batch_closest(const or_scene *s, const float *p, int n, int *idx, float *d2) {
    for (int i = 0; i < n; ++i) {
        float q[3] = {p[(size_t)i * s->dim], p[(size_t)i * s->dim + 1], s->dim == 3 ? p[(size_t)i * 3 + 2] : 0.0f};
        float dd;
        idx[i] = or_closest_index(s, q, &dd);
        if (d2) d2[i] = dd;
    }
}

void or_batch_sdf(const or_scene *s, const float *p, int n, float *out) {
    for (int i = 0; i < n; ++i) {
        float q[3] = {p[(size_t)i * s->dim], p[(size_t)i * s->dim + 1], s->dim == 3 ? p[(size_t)i * 3 + 2] : 0.0f};
        out[i] = or_sdf(s, q);
    }
}

static void store_rec(const or_scene *s, const or_record *r, int i, int *valid, float *rp, float *rn, float *rt,
                      int *prim) {
    int D = s->dim;
    if (valid) valid[i] = r->valid;
    for (int k = 0; k < D; ++k) {
        if (rp) rp[(size_t)i * D + k] = r->valid ? r->p[k] : 0.0f;
        if (rn) rn[(size_t)i * D + k] = r->valid ? r->n[k] : 0.0f;
    }
    if (rt) {
        if (D == 2) rt[i] = r->valid ? r->t : 0.0f;
        else { rt[2 * (size_t)i] = r->valid ? r->uv[0] : 0.0f; rt[2 * (size_t)i + 1] = r->valid ? r->uv[1] : 0.0f; }
    }
    if (prim) prim[i] = r->valid ? r->prim_id : -1;
}

void or_batch_closest_point(const or_scene *s, const float *p, int n, int *valid, float *rp, float *rn, float *rt,
                            int *prim) {
    for (int i = 0; i < n; ++i) {
        float q[3] = {p[(size_t)i * s->dim], p[(size_t)i * s->dim + 1], s->dim == 3 ? p[(size_t)i * 3 + 2] : 0.0f};
        or_record r;
        or_closest_point(s, q, &r);
        store_rec(s, &r, i, valid, rp, rn, rt, prim);
    }
}

void or_batch_lib(const or_scene *s, const int *valid, const float *p, const float *nrm, int n, float eps,
                  float *out) {
    int D = s->dim;
    for (int i = 0; i < n; ++i) {
        or_record r;
        memset(&r, 0, sizeof(r));
        r.valid = valid ? valid[i] : 1;
        for (int k = 0; k < D; ++k) { r.p[k] = p[(size_t)i * D + k]; r.n[k] = nrm[(size_t)i * D + k]; }
        out[i] = or_largest_inscribed_ball(s, &r, eps);
    }
}

void or_batch_dirichlet(const or_scene *s, const int *valid, const int *prim, const float *t, int n, float *out) {
    for (int i = 0; i < n; ++i) {
        or_record r;
        memset(&r, 0, sizeof(r));
        r.valid = valid ? valid[i] : 1;
        r.prim_id = prim[i];
        if (s->dim == 2) r.t = t[i];
        else { r.uv[0] = t[2 * (size_t)i]; r.uv[1] = t[2 * (size_t)i + 1]; }
        out[i] = (r.prim_id >= 0 && r.prim_id < s->n_prims) ? or_dirichlet(s, &r) : 0.0f;
    }
}

void or_batch_single_walk(const or_scene *s, const or_walk_params *w, const float *p, int n, uint64_t *state,
                          uint64_t *inc, int *valid, float *rp, float *rn, float *rt, int *prim, float *contrib) {
    int D = s->dim;
    for (int i = 0; i < n; ++i) {
        or_pcg32 g = {state[i], inc[i]};
        or_record r;
        float q[3] = {p[(size_t)i * D], p[(size_t)i * D + 1], D == 3 ? p[(size_t)i * 3 + 2] : 0.0f};
        or_walk(s, w, q, &g, &r, NULL);
        store_rec(s, &r, i, valid, rp, rn, rt, prim);
        if (contrib) contrib[i] = r.valid ? r.contrib : 0.0f;
        state[i] = g.state;
        inc[i] = g.inc;
    }
}
