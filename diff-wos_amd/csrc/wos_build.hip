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

// Leaf payload in Morton order.
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
        } else {
            int i0 = idx[3 * f], i1 = idx[3 * f + 1], i2 = idx[3 * f + 2];
            leaves[3 * j] = make_float4(vtx[3 * i0], vtx[3 * i0 + 1], vtx[3 * i0 + 2], vtx[3 * i1]);
            leaves[3 * j + 1] = make_float4(vtx[3 * i1 + 1], vtx[3 * i1 + 2], vtx[3 * i2], vtx[3 * i2 + 1]);
            leaves[3 * j + 2] = make_float4(vtx[3 * i2 + 2], __int_as_float(f), 0.0f, 0.0f);
        }
    }
}

__device__ __forceinline__ void write_child_box(float4 *nodes, int D, int parent, int slot, const float *lo,
                                                const float *hi) {
    if (D == 2) {
        nodes[3 * parent + slot] = make_float4(lo[0], lo[1], hi[0], hi[1]);
    } else {
        float *f = reinterpret_cast<float *>(nodes + 4 * parent);
        int o = slot * 6;
        f[o + 0] = lo[0]; f[o + 1] = lo[1]; f[o + 2] = lo[2];
        f[o + 3] = hi[0]; f[o + 4] = hi[1]; f[o + 5] = hi[2];
    }
}

__device__ __forceinline__ void read_child_box(const float4 *nodes, int D, int node, int slot, float *lo,
                                               float *hi) {
    if (D == 2) {
        float4 b = nodes[3 * node + slot];
        lo[0] = b.x; lo[1] = b.y; hi[0] = b.z; hi[1] = b.w;
        lo[2] = hi[2] = 0.0f;
    } else {
        const float *f = reinterpret_cast<const float *>(nodes + 4 * node);
        int o = slot * 6;
        lo[0] = f[o + 0]; lo[1] = f[o + 1]; lo[2] = f[o + 2];
        hi[0] = f[o + 3]; hi[1] = f[o + 4]; hi[2] = f[o + 5];
    }
}

// Writes the children words of every node (left, right).
__global__ void k_node_children(const int *children, int n, int D, float4 *nodes) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += gridDim.x * blockDim.x) {
        int s = (D == 2) ? 3 * i + 2 : 4 * i + 3;
        nodes[s] = make_float4(__int_as_float(children[2 * i]), __int_as_float(children[2 * i + 1]), 0.0f, 0.0f);
    }
}

// Bottom-up boxes: the second child to arrive at a node forms its union and
// carries on to the parent.  Release/acquire at agent scope orders the child
// box stores before the counter update across XCDs.
__global__ void k_refit(int n, int D, const int *children, const int *par_int, const int *par_leaf,
                        const float *vtx, const int *idx, const int *sorted_prim, float4 *nodes, uint32_t *flags) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        int f = sorted_prim[j];
        float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
        float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
        for (int c = 0; c < D; ++c) {
            int v = idx[(long)f * D + c];
            for (int k = 0; k < D; ++k) {
                float x = vtx[(long)v * D + k];
                lo[k] = fminf(lo[k], x);
                hi[k] = fmaxf(hi[k], x);
            }
        }
        int child = ~j;
        int node = par_leaf[j];
        while (node >= 0) {
            int slot = (children[2 * node] == child) ? 0 : 1;
            write_child_box(nodes, D, node, slot, lo, hi);
            __threadfence();
            uint32_t old = __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (old == 0u) break;
            __threadfence();
            float l0[3], h0[3], l1[3], h1[3];
            read_child_box(nodes, D, node, 0, l0, h0);
            read_child_box(nodes, D, node, 1, l1, h1);
            for (int k = 0; k < D; ++k) {
                lo[k] = fminf(l0[k], l1[k]);
                hi[k] = fmaxf(h0[k], h1[k]);
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
__global__ void k_reorder_nodes(const float4 *src, const int *pos, int n_int, int D, float4 *dst) {
    const int W = (D == 2) ? 3 : 4;
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

// 2D: folds every leaf child into its parent node -- slot k holds the
// child's box {lo, hi} or, for a leaf, its segment {a, b - a}; the aux word
// holds {inv, id} per slot with inv = 1/|b - a|^2 for a leaf (id = primitive)
// and inv = -1 for an internal child (id = node).  A traversal step then
// needs one 48-byte fetch and no branch on the child kind.
__global__ void k_pack_nodes_2d(const float4 *src, const float4 *leaves, int n_int, float4 *dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_int; i += gridDim.x * blockDim.x) {
        float4 c = src[3 * i + 2];
        int ch[2] = {__float_as_int(c.x), __float_as_int(c.y)};
        float aux[4];
        for (int k = 0; k < 2; ++k) {
            if (ch[k] < 0) {
                int j = ~ch[k];
                dst[3 * i + k] = leaves[2 * j];
                float4 l1 = leaves[2 * j + 1];
                aux[2 * k] = l1.x;
                aux[2 * k + 1] = l1.y;
            } else {
                dst[3 * i + k] = src[3 * i + k];
                aux[2 * k] = -1.0f;
                aux[2 * k + 1] = __int_as_float(ch[k]);
            }
        }
        dst[3 * i + 2] = make_float4(aux[0], aux[1], aux[2], aux[3]);
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

// 3D angle-weighted vertex normals (include/scene.cuh:381-408).
__global__ void k_face_normals_3d(const float *vtx, const int *idx, int n, float *nrm) {
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
            for (int k = 0; k < 3; ++k) atomicAdd(&nrm[3 * ids[c] + k], ang * fn[k]);
        }
    }
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
    float4 *tmp_nodes = nullptr;

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
        k_karras<<<grid_for(n - 1), kBlock, 0, stream>>>(keys[0], n, children, par_int, par_leaf);
        k_node_children<<<grid_for(n - 1), kBlock, 0, stream>>>(children, n, D, out.nodes);
        k_refit<<<grid_for(n), kBlock, 0, stream>>>(n, D, children, par_int, par_leaf, in.vtx, in.idx, vals[0],
                                                  out.nodes, flags);
        k_depth<<<grid_for(n), kBlock, 0, stream>>>(n, par_int, par_leaf, meta + 13);
        // depth-major node order: the top levels become a contiguous prefix
        // (root stays 0), which the traversal can keep in LDS
        const int ni = n - 1;
        const int nb = (ni + kRsTile - 1) / kRsTile;
        const int W = (D == 2) ? 3 : 4;
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
        k_reorder_nodes<<<grid_for(ni), kBlock, 0, stream>>>(out.nodes, pos, ni, D, tmp_nodes);
        if (D == 2)
            k_pack_nodes_2d<<<grid_for(ni), kBlock, 0, stream>>>(tmp_nodes, out.leaves, ni, out.nodes);
        else
            BCHK(hipMemcpyAsync(out.nodes, tmp_nodes, sizeof(float4) * W * (size_t)ni, hipMemcpyDeviceToDevice,
                                stream));
        (void)hipFreeAsync(hist2, stream);
        if (out.children) BCHK(hipMemcpyAsync(out.children, children, sizeof(int) * 2 * (size_t)(n - 1),
                                              hipMemcpyDeviceToDevice, stream));
    }
    if (D == 2) {
        k_fill_u32<<<grid_for(2L * nv), kBlock, 0, stream>>>(reinterpret_cast<uint32_t *>(out.normals), 0u, 2L * nv);
        k_face_normals_2d<<<grid_for(n), kBlock, 0, stream>>>(in.vtx, in.idx, n, out.normals);
    } else {
        k_fill_u32<<<grid_for(3L * nv), kBlock, 0, stream>>>(reinterpret_cast<uint32_t *>(out.normals), 0u, 3L * nv);
        k_face_normals_3d<<<grid_for(n), kBlock, 0, stream>>>(in.vtx, in.idx, n, out.normals);
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
    return rc;
}

}  // namespace wos
