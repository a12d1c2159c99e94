// wos_build.hip -- scene construction on the GPU: vertex normals, bounds, and
// the LBVH (Morton codes, stable LDS radix sort, Karras 2012 hierarchy,
// bottom-up refit into the traversal layout of wos_device.h).
//
// Replaces Scene<N>::configure_normal / configure (include/scene.cuh:368-453)
// and the lbvh::bvh constructor it calls (include/scene.cuh:452).
#include "wos_build.h"

namespace wos {

// ---- float <-> order-preserving uint (for atomic min/max of floats) -------
__device__ __forceinline__ uint32_t f2o(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t o) {
    uint32_t b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(b);
}

__global__ void k_iota(int *p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// obox[0..2] = ordered min, obox[3..5] = ordered max (initialised to 0xffffffff / 0)
__global__ void k_vertex_bounds(const float *vtx, int nv, int D, uint32_t *obox) {
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x)
        for (int k = 0; k < D; ++k) {
            float x = vtx[(long)v * D + k];
            lo[k] = fminf(lo[k], x);
            hi[k] = fmaxf(hi[k], x);
        }
    for (int k = 0; k < D; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], off));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off));
        }
        if (__lane_id() == 0) {
            atomicMin(&obox[k], f2o(lo[k]));
            atomicMax(&obox[3 + k], f2o(hi[k]));
        }
    }
}

// Primitive centroid = centre of its AABB; centroid bounds into cbox.
// err bit 1: degenerate primitive (zero-length segment / zero-area triangle).
__global__ void k_prim_centroids(const float *vtx, const int *idx, int n, int nv, int D, float *cent,
                                 uint32_t *cbox, uint32_t *err) {
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
        int ids[3];
        bool bad = false;
        for (int c = 0; c < D; ++c) {
            ids[c] = idx[(long)f * D + c];
            if (ids[c] < 0 || ids[c] >= nv) bad = true;
        }
        if (bad) { atomicOr(err, 2u); continue; }
        for (int k = 0; k < D; ++k) {
            float a = __builtin_inff(), b = -__builtin_inff();
            for (int c = 0; c < D; ++c) {
                float x = vtx[(long)ids[c] * D + k];
                a = fminf(a, x);
                b = fmaxf(b, x);
            }
            float m = (a + b) * 0.5f;
            cent[(long)f * 3 + k] = m;
            lo[k] = fminf(lo[k], m);
            hi[k] = fmaxf(hi[k], m);
        }
        if (D == 2) {
            float ex = vtx[2 * ids[1]] - vtx[2 * ids[0]], ey = vtx[2 * ids[1] + 1] - vtx[2 * ids[0] + 1];
            if (dot2(ex, ey, ex, ey) == 0.0f) atomicOr(err, 1u);
        } else {
            float a[3], b[3], c[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = vtx[3 * ids[0] + k];
                b[k] = vtx[3 * ids[1] + k] - a[k];
                c[k] = vtx[3 * ids[2] + k] - a[k];
            }
            float nn[3];
            cross3(b, c, nn);
            if (dot3(nn, nn) == 0.0f) atomicOr(err, 1u);
        }
    }
    for (int k = 0; k < D; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], off));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off));
        }
        if (__lane_id() == 0) {
            atomicMin(&cbox[k], f2o(lo[k]));
            atomicMax(&cbox[3 + k], f2o(hi[k]));
        }
    }
}

__device__ __forceinline__ uint32_t expand2(uint32_t x) {
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
__device__ __forceinline__ uint32_t expand3(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// 2D: 16 bits per axis; 3D: 10 bits per axis.
__global__ void k_morton(const float *cent, int n, int D, const uint32_t *cbox, uint32_t *codes, int *vals) {
    float cmin[3], inv[3];
    for (int k = 0; k < D; ++k) {
        cmin[k] = o2f(cbox[k]);
        float e = o2f(cbox[3 + k]) - cmin[k];
        inv[k] = e > 0.0f ? 1.0f / e : 0.0f;
    }
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
        float u[3] = {0.5f, 0.5f, 0.5f};
        for (int k = 0; k < D; ++k) u[k] = inv[k] > 0.0f ? (cent[(long)f * 3 + k] - cmin[k]) * inv[k] : 0.5f;
        uint32_t code;
        if (D == 2) {
            uint32_t x = (uint32_t)fminf(fmaxf(u[0] * 65536.0f, 0.0f), 65535.0f);
            uint32_t y = (uint32_t)fminf(fmaxf(u[1] * 65536.0f, 0.0f), 65535.0f);
            code = expand2(x) | (expand2(y) << 1);
        } else {
            uint32_t x = (uint32_t)fminf(fmaxf(u[0] * 1024.0f, 0.0f), 1023.0f);
            uint32_t y = (uint32_t)fminf(fmaxf(u[1] * 1024.0f, 0.0f), 1023.0f);
            uint32_t z = (uint32_t)fminf(fmaxf(u[2] * 1024.0f, 0.0f), 1023.0f);
            code = (expand3(x) << 2) | (expand3(y) << 1) | expand3(z);
        }
        codes[f] = code;
        vals[f] = f;
    }
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.  A tile is kRsItems x kBlock keys
// read in (item, wave, lane) order; the scatter ranks keys of equal digit by
// that order, so the sort is stable and equal codes stay in index order.
// ---------------------------------------------------------------------------
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsItems * kBlock;

__global__ __launch_bounds__(kBlock) void k_rs_hist(const uint32_t *keys, int n, int shift, int nblocks,
                                                    uint32_t *hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    long base = (long)blockIdx.x * kRsTile;
    for (int it = 0; it < kRsItems; ++it) {
        long e = base + it * kBlock + threadIdx.x;
        if (e < n) atomicAdd(&h[(keys[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(long)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of m entries in one workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t *hist, long m) {
    __shared__ uint32_t part[1024];
    long per = (m + 1023) / 1024;
    long lo = threadIdx.x * per, hi = lo + per < m ? lo + per : m;
    uint32_t s = 0;
    for (long i = lo; i < hi; ++i) s += hist[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (long i = lo; i < hi; ++i) {
        uint32_t v = hist[i];
        hist[i] = run;
        run += v;
    }
}

// Exclusive scan of a long array in three passes: per-tile sums, one
// workgroup scanning the tile sums, then per-tile scans plus their offsets.
constexpr int kScanTile = 4096;   // elements per workgroup of 1024 threads

__device__ __forceinline__ uint32_t block_exclusive_1024(uint32_t v, uint32_t *sh, uint32_t &total) {
    const int t = threadIdx.x, lane = __lane_id(), w = t >> 6;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    if (t < 16) {
        uint32_t sv = sh[t];
        for (int off = 1; off < 16; off <<= 1) {
            uint32_t y = __shfl_up(sv, off, 16);
            if (t >= off) sv += y;
        }
        sh[16 + t] = sv;
    }
    __syncthreads();
    total = sh[16 + 15];
    uint32_t base = w ? sh[16 + w - 1] : 0u;
    uint32_t excl = base + x - v;
    __syncthreads();
    return excl;
}

__global__ __launch_bounds__(1024) void k_scan_tile_sums(const uint32_t *a, long n, uint32_t *sums) {
    __shared__ uint32_t sh[32];
    const long base = (long)blockIdx.x * kScanTile;
    uint32_t v = 0;
    for (int j = 0; j < kScanTile / 1024; ++j) {
        long i = base + (long)j * 1024 + threadIdx.x;
        if (i < n) v += a[i];
    }
    uint32_t tot;
    (void)block_exclusive_1024(v, sh, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(uint32_t *a, long n, const uint32_t *offs) {
    __shared__ uint32_t sh[32];
    const long base = (long)blockIdx.x * kScanTile;
    uint32_t run = offs[blockIdx.x];
    for (int j = 0; j < kScanTile / 1024; ++j) {
        long i = base + (long)j * 1024 + threadIdx.x;
        uint32_t v = i < n ? a[i] : 0u;
        uint32_t tot;
        uint32_t e = block_exclusive_1024(v, sh, tot);
        if (i < n) a[i] = run + e;
        run += tot;
    }
}

// In-place exclusive scan of n entries; `scratch` holds ceil(n / kScanTile)
// entries.
static void exclusive_scan_u32(uint32_t *a, long n, uint32_t *scratch, hipStream_t stream) {
    const long tiles = (n + kScanTile - 1) / kScanTile;
    k_scan_tile_sums<<<(int)tiles, 1024, 0, stream>>>(a, n, scratch);
    k_rs_scan<<<1, 1024, 0, stream>>>(scratch, tiles);
    k_scan_tiles<<<(int)tiles, 1024, 0, stream>>>(a, n, scratch);
}

__global__ __launch_bounds__(kBlock) void k_rs_scatter(const uint32_t *keys, const int *vals, int n, int shift,
                                                       int nblocks, const uint32_t *offs, uint32_t *okeys,
                                                       int *ovals) {
    __shared__ uint32_t run[256];        // running count per digit within the tile
    __shared__ uint32_t wcnt[4][256];    // per-wave counts of the current item
    const int w = threadIdx.x >> 6, lane = __lane_id();
    run[threadIdx.x] = 0;
    for (int q = 0; q < 4; ++q) wcnt[q][threadIdx.x] = 0;
    __syncthreads();
    long base = (long)blockIdx.x * kRsTile;
    for (int it = 0; it < kRsItems; ++it) {
        long e = base + it * kBlock + threadIdx.x;
        bool ok = e < n;
        uint32_t key = ok ? keys[e] : 0u;
        int val = ok ? vals[e] : 0;
        uint32_t dg = (key >> shift) & 255u;
        // lanes of this wave holding the same digit
        uint64_t peers = __ballot(ok);
        for (int b = 0; b < 8; ++b) {
            uint64_t bal = __ballot((dg >> b) & 1u);
            peers &= ((dg >> b) & 1u) ? bal : ~bal;
        }
        int lrank = rank_in_mask(peers);
        int leader = __ffsll((unsigned long long)peers) - 1;
        if (ok && lane == leader) wcnt[w][dg] = (uint32_t)__popcll(peers);
        __syncthreads();
        // digit-owner thread: prefix over waves, advance running count
        {
            int d = threadIdx.x;
            uint32_t c0 = wcnt[0][d], c1 = wcnt[1][d], c2 = wcnt[2][d], c3 = wcnt[3][d];
            uint32_t r = run[d];
            wcnt[0][d] = r;
            wcnt[1][d] = r + c0;
            wcnt[2][d] = r + c0 + c1;
            wcnt[3][d] = r + c0 + c1 + c2;
            run[d] = r + c0 + c1 + c2 + c3;
        }
        __syncthreads();
        if (ok) {
            uint32_t pos = offs[(long)dg * nblocks + blockIdx.x] + wcnt[w][dg] + (uint32_t)lrank;
            okeys[pos] = key;
            ovals[pos] = val;
        }
        __syncthreads();
        for (int q = 0; q < 4; ++q) wcnt[q][threadIdx.x] = 0;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Karras 2012 hierarchy (one thread per internal node)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int kdelta(const uint32_t *c, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = c[i], b = c[j];
    if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clz((int)(a ^ b));
}

__global__ void k_karras(const uint32_t *codes, int n, int *children, int *par_int, int *par_leaf) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += gridDim.x * blockDim.x) {
        int d = kdelta(codes, n, i, i + 1) - kdelta(codes, n, i, i - 1) > 0 ? 1 : -1;
        int dmin = kdelta(codes, n, i, i - d);
        int lmax = 2;
        while (kdelta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (kdelta(codes, n, i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        int dnode = kdelta(codes, n, i, j);
        int lo = 0, hi = l - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (kdelta(codes, n, i, i + mid * d) > dnode) lo = mid; else hi = mid - 1;
        }
        int gamma = i + lo * d + (d < 0 ? -1 : 0);
        int first = min(i, j), last = max(i, j);
        int left = (first == gamma) ? ~gamma : gamma;
        int right = (last == gamma + 1) ? ~(gamma + 1) : gamma + 1;
        children[2 * i] = left;
        children[2 * i + 1] = right;
        if (left < 0) par_leaf[~left] = i; else par_int[left] = i;
        if (right < 0) par_leaf[~right] = i; else par_int[right] = i;
        if (i == 0) par_int[0] = -1;
    }
}

// Leaf payload in Morton order; in 2D also a copy in primitive order after
// it (leaves + 2 n), which step queries read to seed themselves with the
// previous primitive.
__global__ void k_leaves(const int *sorted_prim, int n, int D, const float *vtx, const int *idx, float4 *leaves) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        int f = sorted_prim[j];
        if (D == 2) {
            int i0 = idx[2 * f], i1 = idx[2 * f + 1];
            float ax = vtx[2 * i0], ay = vtx[2 * i0 + 1];
            float bax = vtx[2 * i1] - ax, bay = vtx[2 * i1 + 1] - ay;
            float inv = 1.0f / dot2(bax, bay, bax, bay);
            leaves[2 * j] = make_float4(ax, ay, bax, bay);
            leaves[2 * j + 1] = make_float4(inv, __int_as_float(f), 0.0f, 0.0f);
            leaves[2 * (long)n + 2 * f] = make_float4(ax, ay, bax, bay);
            leaves[2 * (long)n + 2 * f + 1] = make_float4(inv, __int_as_float(j), 0.0f, 0.0f);
        } else {
            int i0 = idx[3 * f], i1 = idx[3 * f + 1], i2 = idx[3 * f + 2];
            const float4 w0 = make_float4(vtx[3 * i0], vtx[3 * i0 + 1], vtx[3 * i0 + 2], vtx[3 * i1]);
            const float4 w1 = make_float4(vtx[3 * i1 + 1], vtx[3 * i1 + 2], vtx[3 * i2], vtx[3 * i2 + 1]);
            leaves[3 * j] = w0;
            leaves[3 * j + 1] = w1;
            leaves[3 * j + 2] = make_float4(vtx[3 * i2 + 2], __int_as_float(f), 0.0f, 0.0f);
            leaves[3 * (long)n + 3 * f] = w0;
            leaves[3 * (long)n + 3 * f + 1] = w1;
            leaves[3 * (long)n + 3 * f + 2] = make_float4(vtx[3 * i2 + 2], __int_as_float(j), 0.0f, 0.0f);
        }
    }
}

// 3D child boxes: six floats per child in words 0..2 of the node.
__device__ __forceinline__ void write_child_box(float4 *nodes, int node, int slot, const float *lo, const float *hi) {
    float *f = reinterpret_cast<float *>(nodes + 4 * node);
    int o = slot * 6;
    f[o + 0] = lo[0]; f[o + 1] = lo[1]; f[o + 2] = lo[2];
    f[o + 3] = hi[0]; f[o + 4] = hi[1]; f[o + 5] = hi[2];
}

__device__ __forceinline__ void read_child_box(const float4 *nodes, int node, int slot, float *lo, float *hi) {
    const float *f = reinterpret_cast<const float *>(nodes + 4 * node);
    int o = slot * 6;
    lo[0] = f[o + 0]; lo[1] = f[o + 1]; lo[2] = f[o + 2];
    hi[0] = f[o + 3]; hi[1] = f[o + 4]; hi[2] = f[o + 5];
}

// Writes the children word (the last of W words) of every node.
__global__ void k_node_children(const int *children, int n, int W, float4 *nodes) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += gridDim.x * blockDim.x) {
        nodes[W * i + W - 1] =
            make_float4(__int_as_float(children[2 * i]), __int_as_float(children[2 * i + 1]), 0.0f, 0.0f);
    }
}

// 2D bounding capsules: the set of points within r of the segment [a, b].
// The union bound of two capsules takes the farthest pair of their four
// endpoints as its axis (so endpoints are always exact vertex coordinates)
// and the largest endpoint distance to that axis plus the endpoint's own
// radius as its radius -- distance to a segment is convex, so the endpoints
// bound the whole child capsule.  Computed in double and rounded up.
__device__ double seg_dist_d(double px, double py, double ax, double ay, double bx, double by) {
    double dx = bx - ax, dy = by - ay, qx = px - ax, qy = py - ay;
    double dd = dx * dx + dy * dy;
    double h = dd > 0.0 ? (qx * dx + qy * dy) / dd : 0.0;
    h = h < 0.0 ? 0.0 : (h > 1.0 ? 1.0 : h);
    double ex = qx - h * dx, ey = qy - h * dy;
    return sqrt(ex * ex + ey * ey);
}

__device__ void capsule_union(const float4 c0, float r0, const float4 c1, float r1, float4 &c, float &r) {
    const float ex[4] = {c0.x, c0.z, c1.x, c1.z}, ey[4] = {c0.y, c0.w, c1.y, c1.w};
    const float er[4] = {r0, r0, r1, r1};
    int bi = 0, bj = 1;
    double best = -1.0;
    for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j) {
            double dx = (double)ex[j] - ex[i], dy = (double)ey[j] - ey[i];
            double d2 = dx * dx + dy * dy;
            if (d2 > best) { best = d2; bi = i; bj = j; }
        }
    double rr = 0.0;
    for (int k = 0; k < 4; ++k) {
        double d = seg_dist_d(ex[k], ey[k], ex[bi], ey[bi], ex[bj], ey[bj]) + er[k];
        rr = d > rr ? d : rr;
    }
    c = make_float4(ex[bi], ey[bi], ex[bj], ey[bj]);
    r = __double2float_ru(rr * (1.0 + 1e-12));
}

// Bottom-up bounds: the second child to arrive at a node forms the union
// and carries on to the parent.  Release/acquire at agent scope orders the
// child stores before the counter update across XCDs.
//
// 2D build layout (kBuildWords2 words per node): words 0, 1 the capsule
// axes {a, b} of the children, word 2 their radii, words 3, 4 their boxes
// {lo, hi}, word 5 the children.
constexpr int kBuildWords2 = 6;
// Query radius (x scene scale) at which the 2D packer compares the dilated
// areas of a child's capsule and box; walk queries are dominated by steps far
// from the boundary, and larger values measured marginally faster.
constexpr float kBoundR0 = 0.3f;

__global__ void k_refit_2d(int n, const int *children, const int *par_int, const int *par_leaf, const float *vtx,
                           const int *idx, const int *sorted_prim, float4 *nodes, uint32_t *flags) {
    const int W = kBuildWords2;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        int f = sorted_prim[j];
        int i0 = idx[2 * f], i1 = idx[2 * f + 1];
        float ax = vtx[2 * i0], ay = vtx[2 * i0 + 1], bx = vtx[2 * i1], by = vtx[2 * i1 + 1];
        float4 cap = make_float4(ax, ay, bx, by);
        float4 box = make_float4(fminf(ax, bx), fminf(ay, by), fmaxf(ax, bx), fmaxf(ay, by));
        float r = 0.0f;
        int child = ~j;
        int node = par_leaf[j];
        while (node >= 0) {
            int slot = (children[2 * node] == child) ? 0 : 1;
            nodes[W * node + slot] = cap;
            reinterpret_cast<float *>(nodes + W * node + 2)[slot] = r;
            nodes[W * node + 3 + slot] = box;
            __threadfence();
            uint32_t old = __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (old == 0u) break;
            __threadfence();
            float4 rw = nodes[W * node + 2];
            capsule_union(nodes[W * node], rw.x, nodes[W * node + 1], rw.y, cap, r);
            float4 b0 = nodes[W * node + 3], b1 = nodes[W * node + 4];
            box = make_float4(fminf(b0.x, b1.x), fminf(b0.y, b1.y), fmaxf(b0.z, b1.z), fmaxf(b0.w, b1.w));
            child = node;
            node = par_int[node];
        }
    }
}

// 3D build layout (4 words per node): both child boxes in words 0..2, the
// children in word 3.  Alongside, nsum[2 node + slot] accumulates the
// area-weighted normal sum of each child's triangles (the slab direction).
__global__ void k_refit_3d(int n, const int *children, const int *par_int, const int *par_leaf, const float *vtx,
                           const int *idx, const int *sorted_prim, float4 *nodes, float4 *nsum, uint32_t *flags) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        int f = sorted_prim[j];
        float v[3][3];
        float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
        float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
        for (int c = 0; c < 3; ++c) {
            int vi = idx[3L * f + c];
            for (int k = 0; k < 3; ++k) {
                v[c][k] = vtx[3L * vi + k];
                lo[k] = fminf(lo[k], v[c][k]);
                hi[k] = fmaxf(hi[k], v[c][k]);
            }
        }
        float e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
        float e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
        float4 S = make_float4(0.5f * (e1[1] * e2[2] - e1[2] * e2[1]), 0.5f * (e1[2] * e2[0] - e1[0] * e2[2]),
                               0.5f * (e1[0] * e2[1] - e1[1] * e2[0]), 0.0f);
        int child = ~j;
        int node = par_leaf[j];
        while (node >= 0) {
            int slot = (children[2 * node] == child) ? 0 : 1;
            write_child_box(nodes, node, slot, lo, hi);
            nsum[2 * node + slot] = S;
            __threadfence();
            uint32_t old = __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (old == 0u) break;
            __threadfence();
            float l0[3], h0[3], l1[3], h1[3];
            read_child_box(nodes, node, 0, l0, h0);
            read_child_box(nodes, node, 1, l1, h1);
            for (int k = 0; k < 3; ++k) {
                lo[k] = fminf(l0[k], l1[k]);
                hi[k] = fmaxf(h0[k], h1[k]);
            }
            float4 s0 = nsum[2 * node], s1 = nsum[2 * node + 1];
            S = make_float4(s0.x + s1.x, s0.y + s1.y, s0.z + s1.z, 0.0f);
            child = node;
            node = par_int[node];
        }
    }
}

// Unit slab direction of a normal sum (false when the sum vanishes).
__device__ __forceinline__ bool slab_normal(float4 S, float *n) {
    float l = sqrtf(fmaf(S.x, S.x, fmaf(S.y, S.y, S.z * S.z)));
    if (!(l > 0.0f)) return false;
    float il = 1.0f / l;
    n[0] = S.x * il;
    n[1] = S.y * il;
    n[2] = S.z * il;
    return true;
}

// Order-preserving float <-> uint32 map for atomic min / max.
__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Empty ranges: min = +max ordinal, max = 0.
__global__ void k_fill_slab_range(uint32_t *mm, long n_slots) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n_slots; i += (long)gridDim.x * blockDim.x) {
        mm[2 * i] = 0xffffffffu;
        mm[2 * i + 1] = 0u;
    }
}

// Extent of every child's triangles along its slab direction: each leaf
// walks up its ancestors and folds its three vertices into the range of the
// child slot it came through (mm[2 (2 node + slot)] = min, [+1] = max).
__global__ void k_slab_extent(int n, const int *children, const int *par_int, const int *par_leaf,
                              const float *vtx, const int *idx, const int *sorted_prim, const float4 *nsum,
                              uint32_t *mm) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        int f = sorted_prim[j];
        float v[3][3];
        for (int c = 0; c < 3; ++c) {
            int vi = idx[3L * f + c];
            for (int k = 0; k < 3; ++k) v[c][k] = vtx[3L * vi + k];
        }
        int child = ~j;
        int node = par_leaf[j];
        while (node >= 0) {
            int slot = (children[2 * node] == child) ? 0 : 1;
            float nn[3];
            if (slab_normal(nsum[2 * node + slot], nn)) {
                float lo = __builtin_inff(), hi = -__builtin_inff();
                for (int c = 0; c < 3; ++c) {
                    float t = fmaf(nn[0], v[c][0], fmaf(nn[1], v[c][1], nn[2] * v[c][2]));
                    lo = fminf(lo, t);
                    hi = fmaxf(hi, t);
                }
                atomicMin(&mm[2 * (2 * node + slot)], f2ord(lo));
                atomicMax(&mm[2 * (2 * node + slot) + 1], f2ord(hi));
            }
            child = node;
            node = par_int[node];
        }
    }
}

__global__ void k_depth(int n, const int *par_int, const int *par_leaf, uint32_t *depth) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        uint32_t dep = 0;
        int x = par_leaf[j];
        while (x >= 0) { ++dep; x = par_int[x]; }
        atomicMax(depth, dep);
    }
}

// Depth of every internal node (root 0) and a histogram of depths.
__global__ void k_internal_depth(int n_int, const int *par_int, uint32_t *depth, uint32_t *hist) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_int; i += gridDim.x * blockDim.x) {
        uint32_t dep = 0;
        int x = par_int[i];
        while (x >= 0) { ++dep; x = par_int[x]; }
        depth[i] = dep;
        atomicAdd(&hist[dep < 63 ? dep : 63], 1u);
    }
}

// pos[old node] = new node (the stable depth-sorted order).
__global__ void k_scatter_pos(const int *sorted_old, int n_int, int *pos) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n_int; j += gridDim.x * blockDim.x)
        pos[sorted_old[j]] = j;
}

// Copies every node to its depth-major slot, renumbering internal children.
__global__ void k_reorder_nodes(const float4 *src, const int *pos, int n_int, int W, float4 *dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_int; i += gridDim.x * blockDim.x) {
        int j = pos[i];
        for (int k = 0; k < W - 1; ++k) dst[W * j + k] = src[W * i + k];
        float4 c = src[W * i + W - 1];
        int l = __float_as_int(c.x), r = __float_as_int(c.y);
        if (l >= 0) l = pos[l];
        if (r >= 0) r = pos[r];
        dst[W * j + W - 1] = make_float4(__int_as_float(l), __int_as_float(r), 0.0f, 0.0f);
    }
}

// 2D node layout used by the traversal (4 x float4):
//   word k (k = 0, 1): bound of child k as {a.xy, e.xy} -- a leaf's segment
//                      (its leaf word: a, e = b - a), or for an internal
//                      child whichever of its capsule axis (a, e = b - a) and
//                      its box (a = lo, e = hi - lo rounded up) has the
//                      smaller area
//   word 2:            {inv0, r0, inv1, r1}: r = 0 for a leaf (inv as
//                      computed for the leaf), r > 0 a capsule radius (inv =
//                      1/|e|^2, 0 for a point), r = -0.0 a box
//   word 3:            {id0, id1}: internal child index, or ~primitive
// A leaf's distance comes from the same arithmetic as a brute-force scan, so
// both children go through one code path and a node fetch carries everything
// a visit needs.  Capsule radii grow by 2 tol (tol = 1e-6 max(1,
// |coordinate|)) to cover the float evaluation of the bound.
__global__ void k_pack_nodes_2d(const float4 *src, const float4 *leaves, int n_int, const uint32_t *meta,
                                float4 *dst) {
    const int W = kBuildWords2;
    float m = 1.0f;
    for (int k = 0; k < 2; ++k) {
        uint32_t lo = meta[k], hi = meta[3 + k];
        uint32_t blo = (lo & 0x80000000u) ? (lo & 0x7fffffffu) : ~lo;
        uint32_t bhi = (hi & 0x80000000u) ? (hi & 0x7fffffffu) : ~hi;
        m = fmaxf(m, fmaxf(fabsf(__uint_as_float(blo)), fabsf(__uint_as_float(bhi))));
    }
    const float slack = 2e-6f * m;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_int; i += gridDim.x * blockDim.x) {
        float4 rw = src[W * i + 2], c = src[W * i + W - 1];
        int ch[2] = {__float_as_int(c.x), __float_as_int(c.y)};
        float r[2] = {rw.x, rw.y};
        float aux[4];
        int ids[2];
        for (int k = 0; k < 2; ++k) {
            float4 g = src[W * i + k];
            if (ch[k] < 0) {
                float4 l1 = leaves[2 * (~ch[k]) + 1];
                dst[4 * i + k] = leaves[2 * (~ch[k])];
                aux[2 * k] = l1.x;
                aux[2 * k + 1] = 0.0f;
                ids[k] = ~__float_as_int(l1.y);
                continue;
            }
            // keep the bound whose dilation by a ball of radius r0 (a
            // typical query distance) has the smaller area
            float4 bx = src[W * i + 3 + k];
            float bw = bx.z - bx.x, bh = bx.w - bx.y;
            float rr = r[k] + slack;
            float len = sqrtf(dot2(g.z - g.x, g.w - g.y, g.z - g.x, g.w - g.y));
            const float r0 = kBoundR0 * m;
            float bx_cost = fmaf(bw, bh, 2.0f * (bw + bh) * r0);
            float cap_cost = fmaf(2.0f * rr, len, kPi * rr * rr) + (2.0f * len + 2.0f * kPi * rr) * r0;
            if (cap_cost < bx_cost) {
                float ex = g.z - g.x, ey = g.w - g.y;
                float dd = dot2(ex, ey, ex, ey);
                dst[4 * i + k] = make_float4(g.x, g.y, ex, ey);
                aux[2 * k] = dd > 0.0f ? 1.0f / dd : 0.0f;
                aux[2 * k + 1] = rr;
            } else {
                dst[4 * i + k] = make_float4(bx.x, bx.y, nextafterf(bw, __builtin_inff()),
                                             nextafterf(bh, __builtin_inff()));
                aux[2 * k] = 0.0f;
                aux[2 * k + 1] = -0.0f;
            }
            ids[k] = ch[k];
        }
        dst[4 * i + 2] = make_float4(aux[0], aux[1], aux[2], aux[3]);
        dst[4 * i + 3] = make_float4(__int_as_float(ids[0]), __int_as_float(ids[1]), 0.0f, 0.0f);
    }
}

// 3D node layout used by the traversal (kNodeWords3 float4), one block of
// three words per child:
//   internal child: {lo.xyz, hi.x}, {hi.yz, n.xy}, {n.z, c, h, id}: its box
//                   and its slab |n.x - c| <= h (n the unit area-weighted
//                   normal of its triangles); the lower bound of a visit is
//                   max(box distance, |n.p - c| - h)
//   leaf child:     {a.xyz, b.x}, {b.yz, c.xy}, {c.z, 0, 0, ~primitive}
// Slab half-widths grow by 4e-6 max(1, |coordinate|) to cover the float
// evaluation of n.x at build and query time.
__global__ void k_pack_nodes_3d(const float4 *src, const int *old_of_new, const float4 *nsum, const uint32_t *mm,
                                const float4 *leaves, int n_int, const uint32_t *meta, float4 *dst) {
    const int V = kNodeWords3;
    float m = 1.0f;
    for (int k = 0; k < 3; ++k) {
        uint32_t lo = meta[k], hi = meta[3 + k];
        uint32_t blo = (lo & 0x80000000u) ? (lo & 0x7fffffffu) : ~lo;
        uint32_t bhi = (hi & 0x80000000u) ? (hi & 0x7fffffffu) : ~hi;
        m = fmaxf(m, fmaxf(fabsf(__uint_as_float(blo)), fabsf(__uint_as_float(bhi))));
    }
    const float slack = 4e-6f * m;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_int; i += gridDim.x * blockDim.x) {
        const int old = old_of_new[i];
        const float *bx = reinterpret_cast<const float *>(src + 4 * i);
        float4 c = src[4 * i + 3];
        int ch[2] = {__float_as_int(c.x), __float_as_int(c.y)};
        for (int k = 0; k < 2; ++k) {
            float4 *o = dst + V * i + 3 * k;
            if (ch[k] < 0) {
                int j = ~ch[k];
                float4 l0 = leaves[3 * j], l1 = leaves[3 * j + 1], l2 = leaves[3 * j + 2];
                o[0] = l0;
                o[1] = l1;
                o[2] = make_float4(l2.x, 0.0f, 0.0f, __int_as_float(~__float_as_int(l2.y)));
                continue;
            }
            const float *b = bx + 6 * k;
            float nn[3] = {0.0f, 0.0f, 0.0f}, cc = 0.0f, hh = 3.0e38f;
            if (slab_normal(nsum[2 * old + k], nn)) {
                float lo = ord2f(mm[2 * (2 * old + k)]), hi = ord2f(mm[2 * (2 * old + k) + 1]);
                cc = 0.5f * (lo + hi);
                hh = 0.5f * (hi - lo) + slack;
            } else {
                nn[0] = nn[1] = nn[2] = 0.0f;
            }
            o[0] = make_float4(b[0], b[1], b[2], b[3]);
            o[1] = make_float4(b[4], b[5], nn[0], nn[1]);
            o[2] = make_float4(nn[2], cc, hh, __int_as_float(ch[k]));
        }
    }
}

// 2D vertex normals: sum of the unit normals (dy, -dx) of adjacent segments
// (include/scene.cuh:369-380).  Two contributions per vertex sum exactly in
// either order.
__global__ void k_face_normals_2d(const float *vtx, const int *idx, int n, float *nrm) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
        int i0 = idx[2 * f], i1 = idx[2 * f + 1];
        float dx = vtx[2 * i1] - vtx[2 * i0], dy = vtx[2 * i1 + 1] - vtx[2 * i0 + 1];
        float l = sqrtf(dot2(dy, -dx, dy, -dx));
        float nx = dy / l, ny = -dx / l;
        atomicAdd(&nrm[2 * i0], nx);
        atomicAdd(&nrm[2 * i0 + 1], ny);
        atomicAdd(&nrm[2 * i1], nx);
        atomicAdd(&nrm[2 * i1 + 1], ny);
    }
}

// 3D angle-weighted vertex normals (include/scene.cuh:381-408), summed in
// 64-bit fixed point (units of 2^-40; a vertex's sum of angle-weighted unit
// normals stays below valence * pi, far inside the range) so that the result
// does not depend on the order in which the faces add, then converted.
constexpr double kNormalScale = 1099511627776.0;   // 2^40
__global__ void k_face_normals_3d(const float *vtx, const int *idx, int n, unsigned long long *acc) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
        int ids[3] = {idx[3 * f], idx[3 * f + 1], idx[3 * f + 2]};
        float p[3][3];
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) p[c][k] = vtx[3 * ids[c] + k];
        float e1[3], e2[3], fn[3];
        for (int k = 0; k < 3; ++k) { e1[k] = p[1][k] - p[0][k]; e2[k] = p[2][k] - p[0][k]; }
        cross3(e1, e2, fn);
        float l = sqrtf(dot3(fn, fn));
        for (int k = 0; k < 3; ++k) fn[k] /= l;
        for (int c = 0; c < 3; ++c) {
            float a[3], b[3], cr[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = p[(c + 1) % 3][k] - p[c][k];
                b[k] = p[(c + 2) % 3][k] - p[c][k];
            }
            cross3(a, b, cr);
            float ang = atan2f(sqrtf(dot3(cr, cr)), dot3(a, b));
            for (int k = 0; k < 3; ++k)
                atomicAdd(&acc[3 * ids[c] + k],
                          (unsigned long long)__double2ll_rn((double)(ang * fn[k]) * kNormalScale));
        }
    }
}

__global__ void k_normals_from_fixed(const unsigned long long *acc, long m, float *nrm) {
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (long)gridDim.x * blockDim.x)
        nrm[j] = (float)((double)(long long)acc[j] / kNormalScale);
}

__global__ void k_normalize(float *nrm, int nv, int D) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
        float *q = nrm + (long)v * D;
        float l2 = (D == 2) ? dot2(q[0], q[1], q[0], q[1]) : dot3(q, q);
        float l = sqrtf(l2);
        if (l > 0.0f)
            for (int k = 0; k < D; ++k) q[k] = q[k] / l;
    }
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
static int grid_for(long n, int cap = 4096) {
    long g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

#define BCHK(x)                                                   \
    do {                                                          \
        hipError_t e_ = (x);                                      \
        if (e_ != hipSuccess) {                                   \
            set_error("%s: %s", #x, hipGetErrorString(e_));       \
            return -1;                                            \
        }                                                         \
    } while (0)

int build_scene_gpu(BuildInputs &in, hipStream_t stream, BuildOutputs &out) {
    const int D = in.dim, n = in.n_prims, nv = in.n_vertices;
    // scratch: obox(6) cbox(6) err(1) depth(1)
    uint32_t *meta = nullptr;
    float *cent = nullptr;
    uint32_t *keys[2] = {nullptr, nullptr};
    int *vals[2] = {nullptr, nullptr};
    uint32_t *hist = nullptr;
    int *children = nullptr, *par_int = nullptr, *par_leaf = nullptr;
    uint32_t *flags = nullptr;
    int nblocks = (n + kRsTile - 1) / kRsTile;
    int rc = -1;
    uint32_t hmeta[14 + 64];
    uint32_t *dhist = nullptr, *ndepth = nullptr;
    int *dvals[2] = {nullptr, nullptr}, *pos = nullptr;
    uint32_t *dkeys = nullptr;
    float4 *tmp_nodes = nullptr, *work_nodes = nullptr, *nsum = nullptr;
    uint32_t *slab_mm = nullptr;

    BCHK(hipMallocAsync((void **)&meta, (14 + 64) * sizeof(uint32_t), stream));
    dhist = meta + 14;
    BCHK(hipMallocAsync((void **)&cent, sizeof(float) * 3 * (size_t)n, stream));
    for (int b = 0; b < 2; ++b) {
        BCHK(hipMallocAsync((void **)&keys[b], sizeof(uint32_t) * (size_t)n, stream));
        BCHK(hipMallocAsync((void **)&vals[b], sizeof(int) * (size_t)n, stream));
    }
    BCHK(hipMallocAsync((void **)&hist, sizeof(uint32_t) * 256 * (size_t)nblocks, stream));
    BCHK(hipMallocAsync((void **)&children, sizeof(int) * 2 * (size_t)(n > 1 ? n - 1 : 1), stream));
    BCHK(hipMallocAsync((void **)&par_int, sizeof(int) * (size_t)(n > 1 ? n - 1 : 1), stream));
    BCHK(hipMallocAsync((void **)&par_leaf, sizeof(int) * (size_t)n, stream));
    BCHK(hipMallocAsync((void **)&flags, sizeof(uint32_t) * (size_t)(n > 1 ? n - 1 : 1), stream));

    // meta: [0..2] vmin, [3..5] vmax, [6..8] cmin, [9..11] cmax, [12] err, [13] depth
    k_fill_u32<<<1, 64, 0, stream>>>(meta, 0xffffffffu, 3);
    k_fill_u32<<<1, 64, 0, stream>>>(meta + 3, 0u, 3);
    k_fill_u32<<<1, 64, 0, stream>>>(meta + 6, 0xffffffffu, 3);
    k_fill_u32<<<1, 64, 0, stream>>>(meta + 9, 0u, 5);
    k_fill_u32<<<grid_for(n), kBlock, 0, stream>>>(reinterpret_cast<uint32_t *>(par_leaf), 0xffffffffu, n);
    if (n > 1) k_fill_u32<<<grid_for(n - 1), kBlock, 0, stream>>>(flags, 0u, n - 1);
    k_vertex_bounds<<<grid_for(nv, 256), kBlock, 0, stream>>>(in.vtx, nv, D, meta);
    k_prim_centroids<<<grid_for(n, 256), kBlock, 0, stream>>>(in.vtx, in.idx, n, nv, D, cent, meta + 6, meta + 12);
    k_morton<<<grid_for(n), kBlock, 0, stream>>>(cent, n, D, meta + 6, keys[0], vals[0]);
    for (int pass = 0; pass < 4; ++pass) {
        int src = pass & 1, dst = src ^ 1;
        k_rs_hist<<<nblocks, kBlock, 0, stream>>>(keys[src], n, 8 * pass, nblocks, hist);
        k_rs_scan<<<1, 1024, 0, stream>>>(hist, 256L * nblocks);
        k_rs_scatter<<<nblocks, kBlock, 0, stream>>>(keys[src], vals[src], n, 8 * pass, nblocks, hist, keys[dst],
                                                     vals[dst]);
    }
    // after 4 passes the result is back in buffer 0
    BCHK(hipMemcpyAsync(out.sorted_prim, vals[0], sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, stream));
    BCHK(hipMemcpyAsync(out.sorted_codes, keys[0], sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, stream));
    k_leaves<<<grid_for(n), kBlock, 0, stream>>>(vals[0], n, D, in.vtx, in.idx, out.leaves);
    if (n > 1) {
        const int ni = n - 1;
        // 2D builds capsules and boxes in a wider scratch layout and packs
        // them at the end; 3D builds its final layout in place
        const int W = (D == 2) ? kBuildWords2 : 4;
        BCHK(hipMallocAsync((void **)&work_nodes, sizeof(float4) * W * (size_t)ni, stream));
        float4 *work = work_nodes;
        k_karras<<<grid_for(n - 1), kBlock, 0, stream>>>(keys[0], n, children, par_int, par_leaf);
        k_node_children<<<grid_for(n - 1), kBlock, 0, stream>>>(children, n, W, work);
        if (D == 2) {
            k_refit_2d<<<grid_for(n), kBlock, 0, stream>>>(n, children, par_int, par_leaf, in.vtx, in.idx, vals[0],
                                                         work, flags);
        } else {
            BCHK(hipMallocAsync((void **)&nsum, sizeof(float4) * 2 * (size_t)ni, stream));
            BCHK(hipMallocAsync((void **)&slab_mm, sizeof(uint32_t) * 4 * (size_t)ni, stream));
            k_refit_3d<<<grid_for(n), kBlock, 0, stream>>>(n, children, par_int, par_leaf, in.vtx, in.idx, vals[0],
                                                         work, nsum, flags);
            k_fill_slab_range<<<grid_for(2L * ni), kBlock, 0, stream>>>(slab_mm, 2L * ni);
            k_slab_extent<<<grid_for(n), kBlock, 0, stream>>>(n, children, par_int, par_leaf, in.vtx, in.idx,
                                                            vals[0], nsum, slab_mm);
        }
        k_depth<<<grid_for(n), kBlock, 0, stream>>>(n, par_int, par_leaf, meta + 13);
        // depth-major node order: the top levels become a contiguous prefix
        // (root stays 0), which the traversal can keep in LDS
        const int nb = (ni + kRsTile - 1) / kRsTile;
        BCHK(hipMallocAsync((void **)&ndepth, sizeof(uint32_t) * (size_t)ni, stream));
        BCHK(hipMallocAsync((void **)&dkeys, sizeof(uint32_t) * (size_t)ni, stream));
        for (int b = 0; b < 2; ++b) BCHK(hipMallocAsync((void **)&dvals[b], sizeof(int) * (size_t)ni, stream));
        BCHK(hipMallocAsync((void **)&pos, sizeof(int) * (size_t)ni, stream));
        BCHK(hipMallocAsync((void **)&tmp_nodes, sizeof(float4) * W * (size_t)ni, stream));
        k_fill_u32<<<1, 64, 0, stream>>>(dhist, 0u, 64);
        k_internal_depth<<<grid_for(ni), kBlock, 0, stream>>>(ni, par_int, ndepth, dhist);
        k_iota<<<grid_for(ni), kBlock, 0, stream>>>(dvals[0], ni);
        uint32_t *hist2 = nullptr;
        BCHK(hipMallocAsync((void **)&hist2, sizeof(uint32_t) * 256 * (size_t)nb, stream));
        k_rs_hist<<<nb, kBlock, 0, stream>>>(ndepth, ni, 0, nb, hist2);
        k_rs_scan<<<1, 1024, 0, stream>>>(hist2, 256L * nb);
        k_rs_scatter<<<nb, kBlock, 0, stream>>>(ndepth, dvals[0], ni, 0, nb, hist2, dkeys, dvals[1]);
        k_scatter_pos<<<grid_for(ni), kBlock, 0, stream>>>(dvals[1], ni, pos);
        k_reorder_nodes<<<grid_for(ni), kBlock, 0, stream>>>(work, pos, ni, W, tmp_nodes);
        if (D == 2)
            k_pack_nodes_2d<<<grid_for(ni), kBlock, 0, stream>>>(tmp_nodes, out.leaves, ni, meta, out.nodes);
        else
            k_pack_nodes_3d<<<grid_for(ni), kBlock, 0, stream>>>(tmp_nodes, dvals[1], nsum, slab_mm, out.leaves, ni,
                                                               meta, out.nodes);
        (void)hipFreeAsync(hist2, stream);
        if (out.children) BCHK(hipMemcpyAsync(out.children, children, sizeof(int) * 2 * (size_t)(n - 1),
                                              hipMemcpyDeviceToDevice, stream));
    }
    if (D == 2) {
        k_fill_u32<<<grid_for(2L * nv), kBlock, 0, stream>>>(reinterpret_cast<uint32_t *>(out.normals), 0u, 2L * nv);
        k_face_normals_2d<<<grid_for(n), kBlock, 0, stream>>>(in.vtx, in.idx, n, out.normals);
    } else {
        unsigned long long *nacc = nullptr;
        BCHK(hipMallocAsync((void **)&nacc, sizeof(unsigned long long) * 3 * (size_t)nv, stream));
        k_fill_u32<<<grid_for(6L * nv), kBlock, 0, stream>>>(reinterpret_cast<uint32_t *>(nacc), 0u, 6L * nv);
        k_face_normals_3d<<<grid_for(n), kBlock, 0, stream>>>(in.vtx, in.idx, n, nacc);
        k_normals_from_fixed<<<grid_for(3L * nv), kBlock, 0, stream>>>(nacc, 3L * nv, out.normals);
        (void)hipFreeAsync(nacc, stream);
    }
    k_normalize<<<grid_for(nv), kBlock, 0, stream>>>(out.normals, nv, D);
    BCHK(hipGetLastError());
    BCHK(hipMemcpyAsync(hmeta, meta, sizeof(hmeta), hipMemcpyDeviceToHost, stream));
    BCHK(hipStreamSynchronize(stream));
    for (int k = 0; k < 3; ++k) {
        uint32_t lo = hmeta[k], hi = hmeta[3 + k];
        uint32_t blo = (lo & 0x80000000u) ? (lo & 0x7fffffffu) : ~lo;
        uint32_t bhi = (hi & 0x80000000u) ? (hi & 0x7fffffffu) : ~hi;
        memcpy(&out.amin[k], &blo, 4);
        memcpy(&out.amax[k], &bhi, 4);
        if (k >= D) out.amin[k] = out.amax[k] = 0.0f;
    }
    out.error_bits = hmeta[12];
    {
        uint32_t acc = 0;
        out.top_counts[0] = 0;
        for (int k = 0; k < 16; ++k) {
            acc += (n > 1) ? hmeta[14 + k] : 0u;
            out.top_counts[k + 1] = (int)acc;
        }
    }
    out.max_depth = (int)hmeta[13];
    rc = 0;
    (void)hipFreeAsync(meta, stream);
    (void)hipFreeAsync(cent, stream);
    for (int b = 0; b < 2; ++b) {
        (void)hipFreeAsync(keys[b], stream);
        (void)hipFreeAsync(vals[b], stream);
    }
    (void)hipFreeAsync(hist, stream);
    (void)hipFreeAsync(children, stream);
    (void)hipFreeAsync(par_int, stream);
    (void)hipFreeAsync(par_leaf, stream);
    (void)hipFreeAsync(flags, stream);
    if (ndepth) (void)hipFreeAsync(ndepth, stream);
    if (dkeys) (void)hipFreeAsync(dkeys, stream);
    for (int b = 0; b < 2; ++b)
        if (dvals[b]) (void)hipFreeAsync(dvals[b], stream);
    if (pos) (void)hipFreeAsync(pos, stream);
    if (tmp_nodes) (void)hipFreeAsync(tmp_nodes, stream);
    if (work_nodes) (void)hipFreeAsync(work_nodes, stream);
    if (nsum) (void)hipFreeAsync(nsum, stream);
    if (slab_mm) (void)hipFreeAsync(slab_mm, stream);
    return rc;
}

// ---------------------------------------------------------------------------
// near-boundary candidate grid (2D)
// ---------------------------------------------------------------------------
constexpr int kGridStack = 64;   // private traversal stack of the grid builder
constexpr int kGridBlock = 8;    // fine cells per coarse block side

// Sum of segment lengths in a fixed order (one workgroup), so the cell size
// and with it the grid layout are reproducible.
__global__ __launch_bounds__(1024) void k_segment_length_sum(const float *vtx, const int *idx, int n,
                                                              double *out) {
    __shared__ double part[1024];
    double acc = 0.0;
    for (int f = threadIdx.x; f < n; f += 1024) {
        int i0 = idx[2 * f], i1 = idx[2 * f + 1];
        double dx = (double)vtx[2 * i1] - vtx[2 * i0], dy = (double)vtx[2 * i1 + 1] - vtx[2 * i0 + 1];
        acc += sqrt(dx * dx + dy * dy);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
}

__global__ void k_leaf_of_prim(const int *sorted_prim, int n, int *leaf_of_prim) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        leaf_of_prim[sorted_prim[j]] = j;
}

// Child k of a packed 2D node: squared distance to its segment / capsule
// axis, and whether it is a leaf.
__device__ __forceinline__ float child_axis_d2(float px, float py, float4 g, float inv) {
    float h, dx, dy;
    return seg_d2(px, py, g.x, g.y, g.z, g.w, inv, h, dx, dy);
}

// Squared distance from (px, py) to the boundary over the packed 2D nodes.
// Pruning can only drop subtrees no closer than the current best, so the
// result never undershoots the true distance by more than rounding.
__device__ float grid_nearest_d2(const float4 *nodes, const float4 *leaves, int n_prims, float px, float py) {
    float h, dx, dy;
    if (n_prims == 1) {
        float4 l0 = leaves[0], l1 = leaves[1];
        return seg_d2(px, py, l0.x, l0.y, l0.z, l0.w, l1.x, h, dx, dy);
    }
    int stack[kGridStack];
    float sd[kGridStack];
    int sp = 0, node = 0;
    float best = __builtin_inff();   // squared
    for (;;) {
        float4 n0 = nodes[4 * node], n1 = nodes[4 * node + 1], n2 = nodes[4 * node + 2], n3 = nodes[4 * node + 3];
        int c0 = __float_as_int(n3.x), c1 = __float_as_int(n3.y);
        float s0 = child_axis_d2(px, py, n0, n2.x), s1 = child_axis_d2(px, py, n1, n2.z);
        if (c0 < 0) best = fminf(best, s0);
        if (c1 < 0) best = fminf(best, s1);
        // capsule lower bounds, compared in the distance domain
        float lb0 = child_bound_2d(px, py, n0, s0, n2.y), lb1 = child_bound_2d(px, py, n1, s1, n2.w);
        float rb = sqrtf(best);
        bool g0 = c0 >= 0 && lb0 < rb, g1 = c1 >= 0 && lb1 < rb;
        if (g0 && g1) {
            bool near0 = lb0 <= lb1;
            stack[sp] = near0 ? c1 : c0;
            sd[sp] = near0 ? lb1 : lb0;
            ++sp;
            node = near0 ? c0 : c1;
            continue;
        }
        if (g0 || g1) {
            node = g0 ? c0 : c1;
            continue;
        }
        bool found = false;
        while (sp > 0) {
            --sp;
            if (sd[sp] < sqrtf(best)) {
                node = stack[sp];
                found = true;
                break;
            }
        }
        if (!found) return best;
    }
}

// Leaves whose segment lies within r (not squared) of (px, py): counted, or
// written to out.  Returns the count, or -1 once it exceeds cap.
template <bool FILL>
__device__ int grid_range(const float4 *nodes, const float4 *leaves, const int *leaf_of_prim, int n_prims,
                          float px, float py, float r, int cap, int *out) {
    float h, dx, dy;
    int cnt = 0;
    const float r2 = r * r;
    if (n_prims == 1) {
        float4 l0 = leaves[0], l1 = leaves[1];
        if (seg_d2(px, py, l0.x, l0.y, l0.z, l0.w, l1.x, h, dx, dy) <= r2) {
            if (FILL) out[0] = 0;
            cnt = 1;
        }
        return cnt > cap ? -1 : cnt;
    }
    int stack[kGridStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        int node = stack[--sp];
        float4 n2 = nodes[4 * node + 2], n3 = nodes[4 * node + 3];
        for (int k = 0; k < 2; ++k) {
            float4 g = nodes[4 * node + k];
            int id = __float_as_int(k ? n3.y : n3.x);
            float d2 = child_axis_d2(px, py, g, k ? n2.z : n2.x);
            if (id < 0) {
                if (d2 <= r2) {
                    if (cnt >= cap) return -1;
                    if (FILL) out[cnt] = leaf_of_prim[~id];
                    ++cnt;
                }
            } else if (child_bound_2d(px, py, g, d2, k ? n2.w : n2.y) <= r) {
                stack[sp++] = id;
            }
        }
    }
    return cnt;
}

struct GridGeom {
    float ox, oy, cell;
    int nx, ny;
    float band, margin;
    int cap;
};

__device__ __forceinline__ void cell_center(const GridGeom &g, float cx, float cy, float &x, float &y) {
    x = fmaf(cx, g.cell, g.ox);
    y = fmaf(cy, g.cell, g.oy);
}

// Coarse blocks of kGridBlock^2 cells: a block whose centre is farther than
// band + its half-diagonal from the boundary holds no cell of the band.
__global__ void k_grid_blocks(const float4 *nodes, const float4 *leaves, int n_prims, GridGeom g, int nbx, int nby,
                              uint8_t *active) {
    const float half_diag = 0.70710678f * kGridBlock * g.cell;
    for (long b = blockIdx.x * (long)blockDim.x + threadIdx.x; b < (long)nbx * nby; b += (long)gridDim.x * blockDim.x) {
        int bx = (int)(b % nbx), by = (int)(b / nbx);
        float x, y;
        cell_center(g, bx * kGridBlock + 0.5f * kGridBlock, by * kGridBlock + 0.5f * kGridBlock, x, y);
        float d = sqrtf(grid_nearest_d2(nodes, leaves, n_prims, x, y));
        active[b] = d - half_diag <= g.band * 1.0001f + g.margin ? 1 : 0;
    }
}

template <bool FILL>
__global__ void k_grid_cells(const float4 *nodes, const float4 *leaves, const int *leaf_of_prim, int n_prims,
                             GridGeom g, int nbx, const uint8_t *active, uint32_t *counts, int *cand,
                             uint32_t n_cand) {
    const float two_rho = 1.41421356f * g.cell;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < (long)g.nx * g.ny; c += (long)gridDim.x * blockDim.x) {
        int cx = (int)(c % g.nx), cy = (int)(c / g.nx);
        if (FILL) {
            // ranges outside the list (impossible for a correct scan) are skipped
            if (counts[c + 1] == counts[c] || counts[c + 1] < counts[c] || counts[c + 1] > n_cand) continue;
        } else if (!active[(cy / kGridBlock) * nbx + cx / kGridBlock]) {
            counts[c] = 0;
            continue;
        }
        float x, y;
        cell_center(g, cx + 0.5f, cy + 0.5f, x, y);
        float d = sqrtf(grid_nearest_d2(nodes, leaves, n_prims, x, y));
        if (!FILL && d > g.band) {
            counts[c] = 0;
            continue;
        }
        float r = (d + two_rho) * 1.00001f + g.margin;
        // the fill pass repeats the count pass exactly; its capacity is the
        // counted length, so it cannot write past the cell's range
        int n = grid_range<FILL>(nodes, leaves, leaf_of_prim, n_prims, x, y, r,
                                 FILL ? (int)(counts[c + 1] - counts[c]) : g.cap, FILL ? cand + counts[c] : nullptr);
        if (!FILL) counts[c] = n < 0 ? 0u : (uint32_t)n;
    }
}

int build_grid_2d(const GridInputs &in, const GridSettings &gs, hipStream_t stream, GridOutputs &out) {
    out = GridOutputs{};
    if (in.n_prims < 1 || gs.cell_factor <= 0.0f || gs.band_cells <= 0 || gs.max_cand <= 0) return 0;
    // the range traversal holds at most depth + 2 entries; deeper (degenerate)
    // trees keep to the tree queries
    if (in.max_depth + 2 > kGridStack) return 0;
    double *dsum = nullptr;
    int *leaf_of_prim = nullptr;
    uint8_t *active = nullptr;
    uint32_t *counts = nullptr, *scan_tmp = nullptr;
    int *cand = nullptr;
    int rc = -1;
    double sum = 0.0;
    uint32_t total = 0;
    GridGeom g;
    long ncells = 0;
    int nbx = 0, nby = 0;
    BCHK(hipMallocAsync((void **)&dsum, sizeof(double), stream));
    k_segment_length_sum<<<1, 1024, 0, stream>>>(in.vtx, in.idx, in.n_prims, dsum);
    BCHK(hipMemcpyAsync(&sum, dsum, sizeof(double), hipMemcpyDeviceToHost, stream));
    BCHK(hipStreamSynchronize(stream));
    {
        double cell = gs.cell_factor * sum / in.n_prims;
        double band = gs.band_cells * cell;
        double w = (double)in.amax[0] - in.amin[0] + 2.0 * band + 2.0 * cell;
        double h = (double)in.amax[1] - in.amin[1] + 2.0 * band + 2.0 * cell;
        double cells = (w / cell) * (h / cell);
        if (cells > (double)gs.max_cells) {
            cell *= sqrt(cells / (double)gs.max_cells);
            band = gs.band_cells * cell;
            w = (double)in.amax[0] - in.amin[0] + 2.0 * band + 2.0 * cell;
            h = (double)in.amax[1] - in.amin[1] + 2.0 * band + 2.0 * cell;
        }
        g.cell = (float)cell;
        g.ox = (float)(in.amin[0] - band - cell);
        g.oy = (float)(in.amin[1] - band - cell);
        g.nx = (int)ceil(w / cell);
        g.ny = (int)ceil(h / cell);
        g.band = (float)band;
        g.margin = 8.0f * in.tol;   // rounding of cell centres, cell lookup and distances
        g.cap = gs.max_cand;
    }
    ncells = (long)g.nx * g.ny;
    nbx = (g.nx + kGridBlock - 1) / kGridBlock;
    nby = (g.ny + kGridBlock - 1) / kGridBlock;
    BCHK(hipMallocAsync((void **)&leaf_of_prim, sizeof(int) * (size_t)in.n_prims, stream));
    BCHK(hipMallocAsync((void **)&active, (size_t)nbx * nby, stream));
    BCHK(hipMalloc((void **)&counts, sizeof(uint32_t) * (size_t)(ncells + 1)));
    k_leaf_of_prim<<<grid_for(in.n_prims), kBlock, 0, stream>>>(in.sorted_prim, in.n_prims, leaf_of_prim);
    k_grid_blocks<<<grid_for((long)nbx * nby), kBlock, 0, stream>>>(in.nodes, in.leaves, in.n_prims, g, nbx, nby,
                                                                      active);
    k_fill_u32<<<1, 64, 0, stream>>>(counts + ncells, 0u, 1);
    k_grid_cells<false><<<grid_for(ncells), kBlock, 0, stream>>>(in.nodes, in.leaves, leaf_of_prim, in.n_prims, g,
                                                                  nbx, active, counts, nullptr, 0u);
    BCHK(hipMallocAsync((void **)&scan_tmp, sizeof(uint32_t) * (size_t)((ncells + 1) / kScanTile + 1), stream));
    exclusive_scan_u32(counts, ncells + 1, scan_tmp, stream);
    BCHK(hipMemcpyAsync(&total, counts + ncells, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    BCHK(hipStreamSynchronize(stream));
    BCHK(hipMalloc((void **)&cand, sizeof(int) * (size_t)(total > 0 ? total : 1)));
    if (total > 0)
        k_grid_cells<true><<<grid_for(ncells), kBlock, 0, stream>>>(in.nodes, in.leaves, leaf_of_prim, in.n_prims,
                                                                     g, nbx, active, counts, cand, total);
    BCHK(hipGetLastError());
    out.start = counts;
    out.cand = cand;
    out.n_cand = total;
    out.ox = g.ox;
    out.oy = g.oy;
    out.inv_cell = 1.0f / g.cell;
    out.cell = g.cell;
    out.nx = g.nx;
    out.ny = g.ny;
    counts = nullptr;
    cand = nullptr;
    rc = 0;
    (void)hipFreeAsync(dsum, stream);
    (void)hipFreeAsync(leaf_of_prim, stream);
    (void)hipFreeAsync(active, stream);
    if (scan_tmp) (void)hipFreeAsync(scan_tmp, stream);
    if (counts) (void)hipFree(counts);
    if (cand) (void)hipFree(cand);
    return rc;
}

// ---------------------------------------------------------------------------
// start-node grid (2D)
// ---------------------------------------------------------------------------
// For every cell, the deepest node N such that every primitive that can be
// closest to a point of the cell lies in N's subtree.  For a cell with centre
// x, half-diagonal rho and boundary distance D at x, those primitives lie
// within R = D + 2 rho of x (as for the candidate grid); descending from the
// root while exactly one child's bound comes within R of x keeps them all,
// so a query from any point of the cell may start its traversal at N.
// Descends from node while exactly one child's bound comes within R of
// (x, y): every primitive within R of the point lies below the node reached.
// Leaves never become start nodes (a start node is an internal node).  lb:
// the bound of the node reached at (x, y), as its parent stores it (left
// as passed in when the descent does not move).
__device__ int descend_single(const float4 *nodes, int node, float x, float y, float R, bool &split, float &lb) {
    for (;;) {
        const float4 *nd = nodes + 4 * node;
        float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
        int c0 = __float_as_int(n3.x), c1 = __float_as_int(n3.y);
        float s0 = child_axis_d2(x, y, n0, n2.x), s1 = child_axis_d2(x, y, n1, n2.z);
        float lb0 = c0 < 0 ? sqrtf(s0) : child_bound_2d(x, y, n0, s0, n2.y);
        float lb1 = c1 < 0 ? sqrtf(s1) : child_bound_2d(x, y, n1, s1, n2.w);
        bool p0 = lb0 <= R, p1 = lb1 <= R;
        // both children within R: a split point when both are internal
        split = p0 && p1 && c0 >= 0 && c1 >= 0;
        if (p0 == p1) return node;
        int next = p0 ? c0 : c1;
        if (next < 0) return node;
        node = next;
        lb = p0 ? lb0 : lb1;
    }
}

// Per cell: two start nodes {a, b, key} (b = -1: one suffices).  The
// descent stops at the first node with both children within D + 2 rho of
// the cell centre; when both are internal the cell starts at both children
// instead, each descended the same way, so a query skips the levels where
// only one side can hold its answer.  key is a lower bound of b's bound over
// the whole cell (the bound is 1-Lipschitz: its value at the centre minus
// rho, less the float margin), with which a query can prune b unvisited.
__global__ void k_start_cells_2d(const float4 *nodes, const float4 *leaves, int n_prims, GridGeom g, int4 *start,
                                 uint2 *packed) {
    const float two_rho = 1.41421356f * g.cell;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < (long)g.nx * g.ny; c += (long)gridDim.x * blockDim.x) {
        float x, y;
        cell_center(g, (float)(c % g.nx) + 0.5f, (float)(c / g.nx) + 0.5f, x, y);
        int4 out = make_int4(0, -1, 0, 0);
        if (n_prims > 1) {
            const float R = (sqrtf(grid_nearest_d2(nodes, leaves, n_prims, x, y)) + two_rho) * 1.00001f + g.margin;
            bool split = false;
            float lb = 0.0f;
            const int node = descend_single(nodes, 0, x, y, R, split, lb);
            out.x = node;
            if (split) {
                const float4 *nd = nodes + 4 * node;
                const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
                float lb0 = child_bound_2d(x, y, n0, child_axis_d2(x, y, n0, n2.x), n2.y);
                float lb1 = child_bound_2d(x, y, n1, child_axis_d2(x, y, n1, n2.z), n2.w);
                bool s2 = false;
                out.x = descend_single(nodes, __float_as_int(n3.x), x, y, R, s2, lb0);
                out.y = descend_single(nodes, __float_as_int(n3.y), x, y, R, s2, lb1);
                out.z = __float_as_int(fmaxf(0.0f, (lb1 - 0.5f * two_rho) * 0.99999f - g.margin));
            }
        }
        if (packed)
            packed[c] = make_uint2((uint32_t)out.x | ((out.y < 0 ? 0u : (uint32_t)out.y) << 16), (uint32_t)out.z);
        else
            start[c] = out;
    }
}

int build_start_grid_2d(const GridInputs &in, int resolution, hipStream_t stream, StartGridOutputs &out) {
    out = StartGridOutputs{};
    if (in.n_prims < 2 || resolution == 0) return 0;
    GridGeom g;
    double w = (double)in.amax[0] - in.amin[0], h = (double)in.amax[1] - in.amin[1];
    if (resolution < 0) {
        // two cells per mean segment length
        double *dsum = nullptr, sum = 0.0;
        BCHK(hipMallocAsync((void **)&dsum, sizeof(double), stream));
        k_segment_length_sum<<<1, 1024, 0, stream>>>(in.vtx, in.idx, in.n_prims, dsum);
        BCHK(hipMemcpyAsync(&sum, dsum, sizeof(double), hipMemcpyDeviceToHost, stream));
        BCHK(hipFreeAsync(dsum, stream));
        BCHK(hipStreamSynchronize(stream));
        const double mean = sum / in.n_prims;
        const double r = mean > 0.0 ? 2.0 * (w > h ? w : h) / mean : 64.0;
        resolution = (int)std::min(std::max(std::lround(r), 64L), (long)kStartGridAutoMax2);
    }
    double cell = (w > h ? w : h) / resolution;
    if (!(cell > 0.0)) return 0;
    g.cell = (float)cell;
    g.ox = (float)(in.amin[0] - cell);
    g.oy = (float)(in.amin[1] - cell);
    g.nx = (int)ceil(w / cell) + 2;
    g.ny = (int)ceil(h / cell) + 2;
    g.band = 0.0f;
    g.margin = 8.0f * in.tol;
    g.cap = 0;
    // 8-byte cells when every internal node id (0 .. n_prims - 2) fits 16
    // bits; a second node is never the root, so 0 marks "none"
    const bool packed = in.n_prims - 2 <= 0xffff;
    int *start = nullptr;
    BCHK(hipMalloc((void **)&start, (packed ? sizeof(uint2) : sizeof(int4)) * (size_t)g.nx * g.ny));
    k_start_cells_2d<<<grid_for((long)g.nx * g.ny), kBlock, 0, stream>>>(
        in.nodes, in.leaves, in.n_prims, g, packed ? nullptr : reinterpret_cast<int4 *>(start),
        packed ? reinterpret_cast<uint2 *>(start) : nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        (void)hipFree(start);
        set_error("start grid: %s", hipGetErrorString(e));
        return -1;
    }
    out.start = start;
    out.ox = g.ox;
    out.oy = g.oy;
    out.inv_cell = 1.0f / g.cell;
    out.nx = g.nx;
    out.ny = g.ny;
    out.nz = 1;
    out.packed = packed ? 1 : 0;
    return 0;
}

// ---------------------------------------------------------------------------
// start-node grid (3D)
// ---------------------------------------------------------------------------
// Lower bound of an internal child block of a packed 3D node (see
// k_pack_nodes_3d) and the triangle distance of a leaf block.
__device__ __forceinline__ float child_bound_3d(const float *p, float4 x0, float4 x1, float4 x2) {
    float bd = box_d2_3(p, x0.x, x0.y, x0.z, x0.w, x1.x, x1.y);
    float sd = fabsf(fmaf(x1.z, p[0], fmaf(x1.w, p[1], fmaf(x2.x, p[2], -x2.y)))) - x2.z;
    return fmaxf(sqrtf(bd), sd);
}
__device__ __forceinline__ float leaf_block_d2(const float *p, float4 x0, float4 x1, float4 x2) {
    float a[3] = {x0.x, x0.y, x0.z}, b[3] = {x0.w, x1.x, x1.y}, c[3] = {x1.z, x1.w, x2.x};
    float q[3], bary[2];
    return tri_closest(a, b, c, p, q, bary);
}

// Squared distance from p to the boundary over the packed 3D nodes.
__device__ float nearest_d2_3d(const float4 *nodes, const float4 *leaves, int n_prims, const float *p) {
    if (n_prims == 1) {
        float4 l2 = leaves[2];
        return leaf_block_d2(p, leaves[0], leaves[1], make_float4(l2.x, 0.0f, 0.0f, 0.0f));
    }
    int stack[kGridStack];
    float sd[kGridStack];
    int sp = 0, node = 0;
    float best = __builtin_inff();   // squared
    for (;;) {
        const float4 *nd = nodes + kNodeWords3 * node;
        float lb[2];
        int ch[2];
        for (int k = 0; k < 2; ++k) {
            float4 x0 = nd[3 * k], x1 = nd[3 * k + 1], x2 = nd[3 * k + 2];
            ch[k] = __float_as_int(x2.w);
            if (ch[k] < 0) {
                best = fminf(best, leaf_block_d2(p, x0, x1, x2));
                lb[k] = __builtin_inff();
            } else {
                lb[k] = child_bound_3d(p, x0, x1, x2);
            }
        }
        float rb = sqrtf(best);
        bool g0 = ch[0] >= 0 && lb[0] < rb, g1 = ch[1] >= 0 && lb[1] < rb;
        if (g0 && g1) {
            bool near0 = lb[0] <= lb[1];
            stack[sp] = near0 ? ch[1] : ch[0];
            sd[sp] = near0 ? lb[1] : lb[0];
            ++sp;
            node = near0 ? ch[0] : ch[1];
            continue;
        }
        if (g0 || g1) {
            node = g0 ? ch[0] : ch[1];
            continue;
        }
        bool found = false;
        while (sp > 0) {
            --sp;
            if (sd[sp] < sqrtf(best)) {
                node = stack[sp];
                found = true;
                break;
            }
        }
        if (!found) return best;
    }
}

struct GridGeom3 {
    float ox, oy, oz, cell, margin;
    int nx, ny, nz;
};

// As k_start_cells_2d, for cubic cells (half-diagonal sqrt(3)/2 cell).
// 3D counterpart of descend_single: the child bounds are the packed box and
// slab (or, for a leaf, the triangle distance).
__device__ int descend_single_3d(const float4 *nodes, int node, const float *p, float R, bool &split, float &lb) {
    for (;;) {
        const float4 *nd = nodes + kNodeWords3 * node;
        bool pass[2];
        int ch[2];
        float lbs[2];
        for (int k = 0; k < 2; ++k) {
            float4 x0 = nd[3 * k], x1 = nd[3 * k + 1], x2 = nd[3 * k + 2];
            ch[k] = __float_as_int(x2.w);
            lbs[k] = ch[k] < 0 ? sqrtf(leaf_block_d2(p, x0, x1, x2)) : child_bound_3d(p, x0, x1, x2);
            pass[k] = lbs[k] <= R;
        }
        split = pass[0] && pass[1] && ch[0] >= 0 && ch[1] >= 0;
        if (pass[0] == pass[1]) return node;
        int next = pass[0] ? ch[0] : ch[1];
        if (next < 0) return node;
        node = next;
        lb = pass[0] ? lbs[0] : lbs[1];
    }
}

// Per cell {a, b, key, 0} as in 2D (k_start_cells_2d): the first node with
// both children within D + 2 rho of the cell centre is replaced by both
// children, each descended, when both are internal.
__global__ void k_start_cells_3d(const float4 *nodes, const float4 *leaves, int n_prims, GridGeom3 g, int4 *start) {
    const float two_rho = 1.7320508f * g.cell;
    const long nxy = (long)g.nx * g.ny;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < nxy * g.nz; c += (long)gridDim.x * blockDim.x) {
        long cz = c / nxy, r = c % nxy;
        float p[3] = {fmaf((float)(r % g.nx) + 0.5f, g.cell, g.ox), fmaf((float)(r / g.nx) + 0.5f, g.cell, g.oy),
                      fmaf((float)cz + 0.5f, g.cell, g.oz)};
        int4 out = make_int4(0, -1, 0, 0);
        if (n_prims > 1) {
            const float R = (sqrtf(nearest_d2_3d(nodes, leaves, n_prims, p)) + two_rho) * 1.00001f + g.margin;
            bool split = false;
            float lb = 0.0f;
            const int node = descend_single_3d(nodes, 0, p, R, split, lb);
            out.x = node;
            if (split) {
                const float4 *nd = nodes + kNodeWords3 * node;
                float lb0 = child_bound_3d(p, nd[0], nd[1], nd[2]);
                float lb1 = child_bound_3d(p, nd[3], nd[4], nd[5]);
                bool s2 = false;
                out.x = descend_single_3d(nodes, __float_as_int(nd[2].w), p, R, s2, lb0);
                out.y = descend_single_3d(nodes, __float_as_int(nd[5].w), p, R, s2, lb1);
                out.z = __float_as_int(fmaxf(0.0f, (lb1 - 0.5f * two_rho) * 0.99999f - g.margin));
            }
        }
        start[c] = out;
    }
}

int build_start_grid_3d(const GridInputs3 &in, int resolution, hipStream_t stream, StartGridOutputs &out) {
    out = StartGridOutputs{};
    if (in.n_prims < 2 || resolution <= 0) return 0;
    if (in.max_depth + 2 > kGridStack) return 0;
    GridGeom3 g;
    double ext[3], mx = 0.0;
    for (int k = 0; k < 3; ++k) {
        ext[k] = (double)in.amax[k] - in.amin[k];
        mx = ext[k] > mx ? ext[k] : mx;
    }
    double cell = mx / resolution;
    if (!(cell > 0.0)) return 0;
    g.cell = (float)cell;
    g.ox = (float)(in.amin[0] - cell);
    g.oy = (float)(in.amin[1] - cell);
    g.oz = (float)(in.amin[2] - cell);
    g.nx = (int)ceil(ext[0] / cell) + 2;
    g.ny = (int)ceil(ext[1] / cell) + 2;
    g.nz = (int)ceil(ext[2] / cell) + 2;
    g.margin = 8.0f * in.tol;
    int *start = nullptr;
    const long ncells = (long)g.nx * g.ny * g.nz;
    BCHK(hipMalloc((void **)&start, sizeof(int4) * (size_t)ncells));
    k_start_cells_3d<<<grid_for(ncells), kBlock, 0, stream>>>(in.nodes, in.leaves, in.n_prims, g,
                                                                reinterpret_cast<int4 *>(start));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        (void)hipFree(start);
        set_error("start grid: %s", hipGetErrorString(e));
        return -1;
    }
    out.start = start;
    out.ox = g.ox;
    out.oy = g.oy;
    out.oz = g.oz;
    out.inv_cell = 1.0f / g.cell;
    out.nx = g.nx;
    out.ny = g.ny;
    out.nz = g.nz;
    return 0;
}

// ---------------------------------------------------------------------------
// spatial order of query points
// ---------------------------------------------------------------------------
// Morton codes of the points inside the scene's box (points outside clamp to
// its faces); the walk kernels visit points in this order, so the walks in
// flight at any time start close together and share the BVH nodes they
// touch in L2.  Results do not depend on it (walks are keyed by their
// original point and walk index).
__global__ void k_point_morton(const float *p, int n, int D, float3 lo, float3 inv, uint32_t *codes, int *vals) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float l[3] = {lo.x, lo.y, lo.z}, iv[3] = {inv.x, inv.y, inv.z};
        float u[3];
        for (int k = 0; k < 3; ++k) u[k] = k < D ? (p[(long)i * D + k] - l[k]) * iv[k] : 0.5f;
        uint32_t code;
        if (D == 2) {
            uint32_t x = (uint32_t)fminf(fmaxf(u[0] * 65536.0f, 0.0f), 65535.0f);
            uint32_t y = (uint32_t)fminf(fmaxf(u[1] * 65536.0f, 0.0f), 65535.0f);
            code = expand2(x) | (expand2(y) << 1);
        } else {
            uint32_t x = (uint32_t)fminf(fmaxf(u[0] * 1024.0f, 0.0f), 1023.0f);
            uint32_t y = (uint32_t)fminf(fmaxf(u[1] * 1024.0f, 0.0f), 1023.0f);
            uint32_t z = (uint32_t)fminf(fmaxf(u[2] * 1024.0f, 0.0f), 1023.0f);
            code = (expand3(x) << 2) | (expand3(y) << 1) | expand3(z);
        }
        codes[i] = code;
        vals[i] = i;
    }
}

int sort_points_morton(const float *points, int n, int D, const float *amin, const float *amax, int *order,
                       hipStream_t stream) {
    uint32_t *keys[2] = {nullptr, nullptr}, *hist = nullptr;
    int *vals[2] = {nullptr, nullptr};
    const int nblocks = (n + kRsTile - 1) / kRsTile;
    for (int b = 0; b < 2; ++b) {
        BCHK(hipMallocAsync((void **)&keys[b], sizeof(uint32_t) * (size_t)n, stream));
        BCHK(hipMallocAsync((void **)&vals[b], sizeof(int) * (size_t)n, stream));
    }
    BCHK(hipMallocAsync((void **)&hist, sizeof(uint32_t) * 256 * (size_t)nblocks, stream));
    float3 lo = make_float3(amin[0], amin[1], amin[2]);
    float e[3];
    for (int k = 0; k < 3; ++k) e[k] = amax[k] - amin[k];
    float3 inv = make_float3(e[0] > 0 ? 1.0f / e[0] : 0.0f, e[1] > 0 ? 1.0f / e[1] : 0.0f, e[2] > 0 ? 1.0f / e[2] : 0.0f);
    k_point_morton<<<grid_for(n), kBlock, 0, stream>>>(points, n, D, lo, inv, keys[0], vals[0]);
    // the order only steers cache locality, so the sort stops at 8 bits per
    // axis (256^2 or 256^3 cells; equal cells keep the input order): the top
    // 16 bits of the 2D code, the top 24 of the 30-bit 3D code
    const int passes = D == 2 ? 2 : 3, low = D == 2 ? 16 : 6;
    for (int pass = 0; pass < passes; ++pass) {
        int src = pass & 1, dst = src ^ 1;
        k_rs_hist<<<nblocks, kBlock, 0, stream>>>(keys[src], n, low + 8 * pass, nblocks, hist);
        k_rs_scan<<<1, 1024, 0, stream>>>(hist, 256L * nblocks);
        k_rs_scatter<<<nblocks, kBlock, 0, stream>>>(keys[src], vals[src], n, low + 8 * pass, nblocks, hist,
                                                     keys[dst], vals[dst]);
    }
    BCHK(hipMemcpyAsync(order, vals[passes & 1], sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, stream));
    for (int b = 0; b < 2; ++b) {
        (void)hipFreeAsync(keys[b], stream);
        (void)hipFreeAsync(vals[b], stream);
    }
    (void)hipFreeAsync(hist, stream);
    BCHK(hipGetLastError());
    return 0;
}

}  // namespace wos
