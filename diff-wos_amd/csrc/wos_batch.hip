// wos_batch.hip -- many scenes of uneven size in one persistent launch.
//
// The reference solves one scene per call (Solver.walk_reduce,
// wos/solver.py:49-65) and differentiates it per call (CustomSolver,
// wos/solver.py:90-125).  A batch concatenates the walk index spaces of all
// 2D items and of all 3D items; one persistent launch per dimension walks
// them, each lane resolving its walk to (item, point, walk) by a binary
// search over the item offsets kept in LDS.  Every item keeps its own scene,
// walk settings, seed and outputs, and its walks draw exactly the streams of
// a single-scene walk_reduce with the same seed, so batched and per-scene
// results are identical.  The reverse pass (wos_batch_adjoint) maps the
// stored exit records of every item back onto its vertex values and Fourier
// coefficients in one launch per dimension over fixed tiles of walks.  All
// buffers are owned by the batch handle, so neither pass allocates and both
// can be captured into one graph together with the caller's loss gradient.
#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/wos_amd.h"
#include "wos_build.h"
#include "wos_grad.h"

namespace wos {

constexpr int kMaxBatchItems = 1024;
constexpr int kMaxBatchBasis = 16;
constexpr int kBatchHist = 6144;    // vertex-value histogram (64-bit entries) of the reverse pass kept in LDS up to this size
constexpr long long kAdjointTile = 16384;   // walks of one item per adjoint tile
constexpr int kBatchFinish = 8;             // finished walks written together (see k_walk_persistent)
constexpr int kInitChunk = 64;              // points per queue claim of the 3D start queries

// This file's device copy of the tabulated Green's-function CDF.
static __device__ float g_green_cdf[kGreenCdfSize];

static int ensure_green_cdf_batch() {
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

struct BatchItem {
    SceneView v;
    WalkSettings w;
    const float *points;
    int n_points;
    int32_t seed;
    int K;                      // Fourier terms written to basis (0: none)
    double basis_scale;         // 2^basis_exponent of the item's walk settings
    long long walk_begin;       // offset in the dimension group's walk space
    long long point_begin;      // offset in the dimension group's point space
    float *walk_value, *rec_t, *mean;
    unsigned long long *basis;  // K x n_points fixed-point sums, then doubles (the caller's buffer)
    int *rec_prim;
    double *mean64, *var64;
    const float *grad_out;      // reverse pass: dL/d mean per point (nullable)
    float *grad_values;         // adds dL/d vertex values (nullable)
    unsigned long long *vacc;   // with grad_values: n_vertices fixed-point sums, then their scale (a double)
    float *grad_coeffs;         // adds dL/d Fourier coefficients (nullable)
};

// A tile of the reverse pass: walks [begin, end) of one item (item-local
// walk indices), so a workgroup reduces the coefficient sums of one item.
struct AdjointTile {
    int item;
    long long begin, end;
};

__device__ __forceinline__ int find_item(const long long *begins, int n, long long x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (begins[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Start queries of every point of a group's items with persistent lanes
// (the scheme of k_points_persistent3): a lane advances its query `window`
// node visits per round and takes the next point of its wave's chunk when the
// query ends.  Points of items without a usable tree (no BVH, one primitive)
// or, in 2D, with a near-boundary candidate grid are answered at once by the
// one-shot query.  The item-start table is dynamic LDS (n_items entries).
template <int D, class Item>
__global__ __launch_bounds__(kBlock) void k_batch_init_persistent(const Item *items, int n_items,
                                                                   long long n_total_pts, unsigned long long *counter,
                                                                   int chunk, int window, int *iprim, float *isd,
                                                                   int2 *spill, long stride) {
    extern __shared__ long long s_init_begin[];
    long long *s_begin = s_init_begin;
    for (int i = threadIdx.x; i < n_items; i += kBlock) s_begin[i] = items[i].point_begin;
    __syncthreads();
    TravStack st{spill, stride};
    const long lane_g = (long)blockIdx.x * kBlock + threadIdx.x;
    const int lane = __lane_id();
    long long g = -1;
    int it = 0;
    bool busy = false;
    float q3[3] = {0.0f, 0.0f, 0.0f};
    Query q;
    q.node = q.sp = 0;
    q.best = q.thr = 0.0f;
    q.best_id = -1;
    q.nbox = q.nprim = 0;
    SceneView tv;            // the traversal's view: the item's nodes (no LDS copy)
    tv.nodes = nullptr;
    tv.leaves = nullptr;
    tv.pleaves = nullptr;
    tv.n_top = 0;
    float tol = 0.0f;
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
                if (bb >= n_total_pts) {
                    exhausted = true;
                } else {
                    fresh = true;
                    base2 = bb;
                    end2 = bb + chunk < n_total_pts ? bb + chunk : n_total_pts;
                }
            }
            if (!busy) {
                const int r = rank_in_mask(idle);
                long long j = -1;
                if (r < avail) j = c_next + r;
                else if (fresh && (r - avail) < end2 - base2) j = base2 + (r - avail);
                if (j >= 0) {
                    g = j;
                    it = find_item(s_begin, n_items, g);
                    const SceneView &s = items[it].v;
                    load_point<D>(items[it].points, g - s_begin[it], q3);
                    if (!s.use_bvh || s.n_prims == 1 || (D == 2 && s.grid_start)) {
                        float d2;
                        const int id = closest_prim<D>(s, q3, __builtin_inff(), d2, st, threadIdx.x, lane_g);
                        iprim[g] = id;
                        isd[g] = id < 0 ? __builtin_inff() : signed_dist<D>(s, id, q3);
                    } else {
                        tv.nodes = s.nodes;
                        tol = s.tol;
                        query_begin_at<D>(q, s, q3, __builtin_inff(), 0x7fffffff, __builtin_inff(), st, threadIdx.x,
                                          lane_g);
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
            if (busy && query_step<D, false>(tv, q3, q, st, threadIdx.x, lane_g, tol)) {
                const int id = q.best_id == 0x7fffffff ? -1 : q.best_id;
                iprim[g] = id;
                isd[g] = id < 0 ? __builtin_inff() : signed_dist<D>(items[it].v, id, q3);
                busy = false;
            }
            if (__ballot(busy) == 0) break;
        }
    }
}

template <int D, int KB>
__device__ __forceinline__ void batch_write(const BatchItem &b, long long k, bool valid, int prim, const float *p,
                                            float contrib, const float *bas) {
    float v = 0.0f, t0 = 0.0f, t1 = 0.0f;
    int pr = -1;
    if (valid) {
        Rec r;
        closest_record<D>(b.v, prim, p, r);
        v = dirichlet_at<D>(b.v, prim, r.t, r.uv[0], r.uv[1]) + contrib;
        pr = prim;
        t0 = (D == 2) ? r.t : r.uv[0];
        t1 = r.uv[1];
    }
    if (b.walk_value) b.walk_value[k] = v;
    if (b.rec_prim) b.rec_prim[k] = pr;
    if (b.rec_t) {
        if (D == 2) b.rec_t[k] = t0;
        else { b.rec_t[2 * k] = t0; b.rec_t[2 * k + 1] = t1; }
    }
    if constexpr (KB > 0) {
        if (valid && b.basis)
            add_basis<KB>(b.basis + point_of_walk(k, b.w.nwalks), b.n_points, b.K, b.basis_scale, bas);
    }
}

// Per item (blockIdx.y): zeroes the fixed-point basis accumulators before the
// walks (finish = 0), or turns them into doubles after them (finish = 1).
__global__ void k_batch_basis(const BatchItem *items, int finish) {
    const BatchItem &b = items[blockIdx.y];
    if (!b.basis || b.K <= 0) return;
    const long n = (long)b.K * b.n_points;
    const double inv = 1.0 / b.basis_scale;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
        if (finish) b.basis[j] = (unsigned long long)__double_as_longlong((double)(long long)b.basis[j] * inv);
        else b.basis[j] = 0ull;
    }
}

// The persistent walk of the single-scene kernel (k_walk_persistent) over
// the concatenated walk space of a dimension group: a lane resolves a new
// walk to its item by binary search, keeps the item's node pointer in
// registers for its resumable queries and reads the rest of the item's
// scene view from the item table (cached; items are few and shared).
// Finished walks wait until kBatchFinish lanes of the wave are done or idle
// and then write their outputs together.
// KB: Fourier basis capacity (the group's largest K rounded up to 8 or 16; 0
// in 3D, which has no source).  With K <= 8 the 2D kernel is held to 64 VGPRs
// (8 waves per SIMD, as the single-scene walk: 8 % faster on the clover), with
// 16 terms to 80 (6 waves); the 3D kernel measured 11 % slower at 72 VGPRs (7
// waves) than unconstrained.
template <int D, int KB>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(D == 2 ? (KB <= 8 ? 8 : 6) : 1))) void
k_batch_walk(const BatchItem *__restrict__ items, int n_items, long long total, unsigned long long *counter, int chunk,
             int window, const int *iprim, const float *isd, int2 *spill, long stride) {
    // the items' first walks, n_items entries of dynamic LDS (a static
    // kMaxBatchItems table would hold the kernel to 6 waves per SIMD)
    extern __shared__ long long s_walk_begin[];
    long long *s_begin = s_walk_begin;
    for (int i = threadIdx.x; i < n_items; i += kBlock) s_begin[i] = items[i].walk_begin;
    __syncthreads();
    TravStack st{spill, stride};
    const long lane_g = (long)blockIdx.x * kBlock + threadIdx.x;
    const int lane = __lane_id();

    int item = 0;
    long long k = -1;      // walk index within the item
    int mode = kIdle;
    float p[3] = {0.0f, 0.0f, 0.0f};
    float d = 0.0f, contrib = 0.0f;
    int step = 0, prim = -1;
    bool done_valid = false;
    Pcg32 rng;
    rng.state = rng.inc = 0;
    Query q;
    q.node = q.sp = 0;
    q.best = q.thr = 0.0f;
    q.best_id = -1;
    q.nbox = q.nprim = 0;
    SceneView tv;            // what a traversal step reads: the item's nodes and leaf triangles (no LDS copy)
    tv.nodes = nullptr;
    tv.leaves = nullptr;
    tv.n_top = 0;
    float tol = 0.0f;
    constexpr int NB = KB > 0 ? KB : 1;
    float bas[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) bas[j] = 0.0f;
    long long c_next = 0, c_end = 0;
    bool exhausted = false;

    for (;;) {
        // 1. completed queries: stop, give up, or take the next step
        if (mode == kReady) {
            const BatchItem &b = items[item];
            prim = q.best_id;
            d = sqrtf(q.best);
            done_valid = d < b.w.epsilon && step < b.w.nsteps;
            const bool done_invalid = !done_valid && step >= b.w.nsteps;
            mode = (done_valid || done_invalid) ? kDone : kStart;
        }
        // 2. write finished walks once kBatchFinish lanes are done or idle
        //    (or nothing else is left to run), then refill idle lanes
        if (__popcll(__ballot(mode == kDone || mode == kIdle)) >= kBatchFinish || exhausted ||
            __ballot(mode == kStart || mode == kQuery) == 0) {
            if (mode == kDone) {
                // (the item at a wave-uniform index as in the refill below: 1 % faster)
                const int iu = __builtin_amdgcn_readfirstlane(item);
                if (D == 2 && __ballot(item != iu) == 0)
                    batch_write<D, KB>(items[iu], k, done_valid, prim, p, contrib, bas);
                else
                    batch_write<D, KB>(items[item], k, done_valid, prim, p, contrib, bas);
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
                unsigned long long bq = 0;
                if (lane == 0) bq = atomicAdd(counter, (unsigned long long)chunk);
                uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)bq);
                uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(bq >> 32));
                long long bb = (long long)(((uint64_t)hi << 32) | lo);
                if (bb >= total) {
                    exhausted = true;
                } else {
                    fresh = true;
                    base2 = bb;
                    end2 = bb + chunk < total ? bb + chunk : total;
                }
            }
            if (mode == kIdle) {
                int r = rank_in_mask(idle);
                long long gk = -1;
                if (r < avail) gk = c_next + r;
                else if (fresh && (r - avail) < end2 - base2) gk = base2 + (r - avail);
                if (gk >= 0) {
                    item = find_item(s_begin, n_items, gk);
                    // in 2D the item's fields by scalar loads when every refilled lane of
                    // the wave takes a walk of the same item (chunks lie inside items):
                    // one-item clover batch 16.4 -> 15.3 ms
                    auto refill = [&](const BatchItem &b) {
                        long long kk = gk - s_begin[item];
                        long long i = point_of_walk(kk, b.w.nwalks);
                        long long gp = b.point_begin + i;
                        int ip = iprim[gp];
                        float sd = isd[gp];
                        load_point<D>(b.points, i, p);
                        contrib = 0.0f;
#pragma unroll
                        for (int j = 0; j < NB; ++j) bas[j] = 0.0f;
                        // include/wos.cuh:100-106, then the first loop test
                        bool outside = ip < 0 || (sd >= 0.0f && !b.w.double_sided);
                        float d0 = fabsf(sd);
                        if (b.w.prevent_fd_artifacts) d0 *= 0.5f;
                        if (outside || b.w.nsteps <= 0) {
                            batch_write<D, KB>(b, kk, false, ip, p, 0.0f, bas);
                        } else if (d0 < b.w.epsilon) {
                            batch_write<D, KB>(b, kk, true, ip, p, 0.0f, bas);
                        } else {
                            k = kk;
                            rng = walk_stream(b.seed, i, kk);
                            prim = ip;
                            d = d0;
                            step = 0;
                            tv.nodes = b.v.nodes;
                            tv.leaves = b.v.leaves;
                            tol = b.v.tol;
                            mode = kStart;
                        }
                    };
                    const int iu = __builtin_amdgcn_readfirstlane(item);
                    if (D == 2 && __ballot(item != iu) == 0) refill(items[iu]);
                    else refill(items[item]);
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

        // 3. a new step: source sample, jump to the sphere, begin the query.
        //    In 2D, when every stepping lane of the wave walks the same item
        //    (the common case: chunks lie inside items), the item is read at a
        //    wave-uniform index, so its fields come in by scalar loads (the
        //    64-scene batch's 2D kernel 7.5 % faster).
        if (mode == kStart) {
            auto step_phase = [&](const BatchItem &b) {
                const SceneView &s = b.v;
                if (D == 2 && s.source_type != 0) {
                    const int K = b.K;
                    const float u0 = rng.next_1d(), u1 = rng.next_1d();
                    float dyx, dyy, G, w;
                    if (interior_sample(u0, u1, d, b.w.use_is != 0, g_green_cdf, dyx, dyy, G, w)) {
                        float yx = p[0] + dyx, yy = p[1] + dyy;
                        if (KB > 0 && K > 0) {
                            float f = 0.0f;
#pragma unroll
                            for (int j = 0; j < NB; ++j) {
                                if (j < K) {
                                    const const_f32 *c = as_constant(s.src) + 4 * j;
                                    float bq = sin_cw(fmaf(c[0], yx, fmaf(c[1], yy, c[2])));
                                    f = fmaf(c[3], bq, f);
                                    bas[j] = fmaf(bq * G, w, bas[j]);
                                }
                            }
                            contrib = fmaf(f * G, w, contrib);
                        } else {
                            contrib = fmaf(source_at(s, yx, yy) * G, w, contrib);
                        }
                    }
                }
                move_point<D>(p, d, rng);
                ++step;
                if (!s.use_bvh) {
                    float d2;
                    int id = closest_brute<D>(s, p, d2);
                    q.best = d2;
                    q.best_id = id;
                    mode = kReady;
                } else {
                    const float seed_d2 = prim_d2<D>(s, prim, p);
                    if (s.n_prims == 1) {
                        q.best = seed_d2;
                        q.best_id = prim;
                        mode = kReady;
                    } else {
                        query_begin_from<D>(q, start_nodes_of<D>(s, p), bound_d2(s, 2.0f * d), prim, seed_d2, tol, st,
                                            threadIdx.x, lane_g);
                        mode = kQuery;
                    }
                }
            };
            if constexpr (D == 2) {
                const int iu = __builtin_amdgcn_readfirstlane(item);
                if (__ballot(item != iu) == 0) step_phase(items[iu]);
                else step_phase(items[item]);
            } else {
                step_phase(items[item]);   // 3D: the uniform copy measured 5 % slower
            }
        }
        // 4. advance every pending query by up to `window` node visits
        for (int it = 0; it < window; ++it) {
            if (mode == kQuery) {
                if (query_step<D, false>(tv, p, q, st, threadIdx.x, lane_g, tol)) mode = kReady;
            }
            if (__ballot(mode == kQuery) == 0) break;
        }
    }
}

// One wave per point of the group: FP64 mean / variance into each item's outputs.
__global__ __launch_bounds__(kBlock) void k_batch_reduce(const BatchItem *items, int n_items, long long n_pts) {
    __shared__ long long s_begin[kMaxBatchItems];
    for (int i = threadIdx.x; i < n_items; i += kBlock) s_begin[i] = items[i].point_begin;
    __syncthreads();
    const int lane = __lane_id();
    long long wave = ((long long)blockIdx.x * kBlock + threadIdx.x) >> 6;
    long long nwaves = ((long long)gridDim.x * kBlock) >> 6;
    for (long long g = wave; g < n_pts; g += nwaves) {
        int it = find_item(s_begin, n_items, g);
        const BatchItem &b = items[it];
        long long i = g - s_begin[it];
        const int W = b.w.nwalks;
        double sm = 0.0, s2 = 0.0;
        for (int w = lane; w < W; w += 64) {
            double v = b.walk_value[i * W + w];
            sm += v;
            s2 += v * v;
        }
        for (int off = 32; off > 0; off >>= 1) {
            sm += __shfl_xor(sm, off);
            s2 += __shfl_xor(s2, off);
        }
        if (lane == 0) {
            double m = sm / W;
            if (b.mean) b.mean[i] = (float)m;
            if (b.mean64) b.mean64[i] = m;
            if (b.var64) b.var64[i] = W > 1 ? (s2 - sm * m) / (W - 1) : 0.0;
        }
    }
}

// Per item with grad_values (blockIdx.x): zeroes its fixed-point value sums
// and sets their scale from its grad_out (value_scale_block).
__global__ __launch_bounds__(kBlock) void k_batch_value_prep(const BatchItem *items) {
    const BatchItem &b = items[blockIdx.x];
    if (!b.grad_values || b.n_points <= 0) return;
    for (int v = threadIdx.x; v < b.v.n_vertices; v += kBlock) b.vacc[v] = 0ull;
    value_scale_block(b.grad_out, b.n_points, reinterpret_cast<double *>(b.vacc + b.v.n_vertices));
}

// Per item (blockIdx.y): fixed-point value sums -> grad_values (added into).
__global__ void k_batch_value_finish(const BatchItem *items) {
    const BatchItem &b = items[blockIdx.y];
    if (!b.grad_values || b.n_points <= 0) return;
    const double inv = 1.0 / reinterpret_cast<const double *>(b.vacc + b.v.n_vertices)[0];
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < b.v.n_vertices; v += gridDim.x * blockDim.x)
        if (b.vacc[v] != 0ull) b.grad_values[v] += (float)((double)(long long)b.vacc[v] * inv);
}

// Reverse pass over the per-walk records (Solver.walk_reduce's adjoint,
// wos/solver.py:90-125): each record adds g_i / W times the interpolation
// weights of its exit primitive to the vertex values (include/scene.cuh:
// 251-264 transposed), rounded to the item's fixed-point unit and summed as
// 64-bit integers (bit-reproducible).  A workgroup takes a contiguous run of
// tiles, so it mostly stays on one item: the sums go to a histogram in LDS
// (items with at most hist_size vertices; larger meshes add to the item's
// sums directly) and reach memory once per item the run touches.
template <int D>
__global__ __launch_bounds__(kBlock) void k_batch_adjoint(const BatchItem *items, const AdjointTile *tiles,
                                                          int n_tiles, int hist_size) {
    extern __shared__ unsigned long long h[];
    for (int v = threadIdx.x; v < hist_size; v += kBlock) h[v] = 0ull;
    const int per = (n_tiles + (int)gridDim.x - 1) / (int)gridDim.x;
    const int t0 = (int)blockIdx.x * per, t1 = min(n_tiles, t0 + per);
    int cur = -1;
    for (int t = t0; t <= t1; ++t) {
        const int item = t < t1 ? tiles[t].item : -1;
        if (item != cur) {
            // flush the previous item's sums
            if (cur >= 0) {
                const BatchItem &b = items[cur];
                __syncthreads();
                if (b.grad_values && b.v.n_vertices <= hist_size) {
                    for (int v = threadIdx.x; v < b.v.n_vertices; v += kBlock) {
                        const unsigned long long x = h[v];
                        if (x != 0ull) atomicAdd(&b.vacc[v], x);
                        h[v] = 0ull;
                    }
                }
            }
            __syncthreads();
            cur = item;
        }
        if (t == t1) break;
        const AdjointTile tl = tiles[t];
        const BatchItem &b = items[tl.item];
        const int W = b.w.nwalks;
        const bool lds = b.v.n_vertices <= hist_size;
        if (!b.grad_values) continue;
        unsigned long long *dst = lds ? h : b.vacc;
        const double sc = reinterpret_cast<const double *>(b.vacc + b.v.n_vertices)[0];
        for (long long k = tl.begin + threadIdx.x; k < tl.end; k += kBlock) {
            const int f = b.rec_prim[k];
            if (f < 0) continue;
            const float g = b.grad_out[point_of_walk(k, W)] / (float)W;
            const int *ix = b.v.idx;
            if constexpr (D == 2) {
                const float tt = b.rec_t[k];
                const int i0 = ix[2 * f], i1 = ix[2 * f + 1];
                atomicAdd(&dst[i0], value_term(g * (1.0f - tt), sc));
                atomicAdd(&dst[i1], value_term(g * tt, sc));
            } else {
                const float u = b.rec_t[2 * k], v = b.rec_t[2 * k + 1];
                const int i0 = ix[3 * f], i1 = ix[3 * f + 1], i2 = ix[3 * f + 2];
                atomicAdd(&dst[i0], value_term(g * (1.0f - u - v), sc));
                atomicAdd(&dst[i1], value_term(g * u, sc));
                atomicAdd(&dst[i2], value_term(g * v, sc));
            }
        }
    }
}

// Coefficient part of the reverse pass: dL/d coeff_q of item blockIdx.y (q =
// blockIdx.x) = sum_i g_i / W * basis[q][i] over the item's per-point basis
// sums, in FP64 with a fixed-order tree, so the result is bit-reproducible.
__global__ __launch_bounds__(kBlock) void k_batch_adjoint_coeffs(const BatchItem *items) {
    __shared__ double part[kBlock];
    const BatchItem &b = items[blockIdx.y];
    const int q = blockIdx.x;
    if (!b.grad_coeffs || !b.basis || q >= b.K || b.n_points <= 0) return;
    const double *bb = reinterpret_cast<const double *>(b.basis) + (long)q * b.n_points;
    double acc = 0.0;
    for (long i = threadIdx.x; i < b.n_points; i += kBlock) acc = fma((double)b.grad_out[i], bb[i], acc);
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) b.grad_coeffs[q] += (float)(part[0] / b.w.nwalks);
}

__global__ void k_batch_zero(unsigned long long *c, int n) {
    if (threadIdx.x < n) c[threadIdx.x] = 0ull;
}

// A shape-gradient item: CustomSolver backward with OursCUDA / Ours3D
// (wos/solver.py:109-114, wos/wos_grad.py:130-151, wos/wos_grad_3d.py:
// 117-137) on one scene, its own sample settings and the batch item's seed.
struct GradItem {
    SceneView v;
    GradSettings g;
    const float *points;
    int n_points;
    int32_t seed;
    long long walk_begin;       // offset in the group's sample space
    long long point_begin;      // offset in the group's point space (start records)
    const float *grad_out;      // dL/d mean per point
    float *grad_vertices;       // adds dL/d vertex positions (n_vertices * dim)
    float4 *vstage;             // the group's staged gradient terms (dim - 1 per sample, at walk_begin + sample)
};

// Sample source of the batched shape-gradient launch (grad_samples in
// wos_grad.h): a claimed slot of the group's sample space is resolved to its
// item by binary search over the item offsets in LDS; the item's scene and
// settings are read from the item table (cached; items are few and shared),
// except the node pointer and pruning margin of the traversal, which the
// lane keeps in registers.
struct GradBatch {
    const GradItem *items;
    int n_items;
    const long long *begins;    // LDS: walk_begin of each item
    long long n_samples;
    unsigned long long *queue;
    int claim, visits;
    const int *iprim;
    const float *isd;
    __device__ __forceinline__ long long total() const { return n_samples; }
    __device__ __forceinline__ unsigned long long *counter() const { return queue; }
    __device__ __forceinline__ int chunk() const { return claim; }
    __device__ __forceinline__ int window() const { return visits; }
    __device__ __forceinline__ const SceneView &scene(int it) const { return items[it].v; }
    __device__ __forceinline__ const GradSettings &settings(int it) const { return items[it].g; }
    __device__ __forceinline__ const SceneView &trav(int, const SceneView &tv) const { return tv; }
    __device__ __forceinline__ void bind(int it, SceneView &tv) const {
        tv.nodes = items[it].v.nodes;
        tv.tol = items[it].v.tol;
        tv.n_top = 0;
    }
    __device__ __forceinline__ int take(long long wslot, long long &i, long long &kk) const {
        const int it = find_item(begins, n_items, wslot);
        kk = wslot - begins[it];
        i = point_of_walk(kk, items[it].g.nwalks);
        return it;
    }
    __device__ __forceinline__ int start_prim(int it, long long i) const { return iprim[items[it].point_begin + i]; }
    __device__ __forceinline__ float start_sd(int it, long long i) const { return isd[items[it].point_begin + i]; }
    template <int D>
    __device__ __forceinline__ void load(int it, long long i, float *p) const { load_point<D>(items[it].points, i, p); }
    __device__ __forceinline__ int32_t seed(int it) const { return items[it].seed; }
    template <int D>
    __device__ __forceinline__ void skip(int it, long long kk) const {
        const float zero[3] = {0.0f, 0.0f, 0.0f};
        stage_vertex_grad<D>(items[it].vstage, items[it].walk_begin + kk, false, 0, 0.0f, 0.0f, 0.0f, zero);
    }
    template <int D>
    __device__ __forceinline__ void finish(int it, long long k, bool ok, int rprim, float rt0, float rt1,
                                           const float *gu) const {
        const GradItem &b = items[it];
        const int W = b.g.nwalks;
        stage_vertex_grad<D>(b.vstage, b.walk_begin + k, ok, rprim, rt0, rt1, b.grad_out[point_of_walk(k, W)] / (float)W,
                             gu);
    }
};

// Shape-gradient samples of every gradient item of a dimension group in one
// persistent launch; registers held as in the single-scene kernel (2D 64, 3D
// 80), the item-start table in dynamic LDS (n_items entries).
template <int D>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(D == 2 ? 8 : 6))) void k_batch_grad(
    const GradItem *items, int n_items, long long n_samples, unsigned long long *queue, int claim, int visits,
    const int *iprim, const float *isd, int2 *spill, long stride) {
    extern __shared__ long long s_grad_begin[];
    long long *s_begin = s_grad_begin;
    for (int i = threadIdx.x; i < n_items; i += kBlock) s_begin[i] = items[i].walk_begin;
    __syncthreads();
    const TravStack st{spill, stride};
    const GradBatch src{items, n_items, s_begin, n_samples, queue, claim, visits, iprim, isd};
    grad_samples<D, false>(src, st, threadIdx.x, (long)blockIdx.x * kBlock + threadIdx.x);
}

struct BatchGroup {
    BatchItem *items = nullptr;   // device copy
    int n_items = 0;
    long long walks = 0, points = 0;
    int *iprim = nullptr;
    float *isd = nullptr;
    int max_depth = 0;
    AdjointTile *tiles = nullptr; // device copy (null: no item has a reverse pass)
    int n_tiles = 0;
    int hist = 0;                 // LDS histogram entries of the reverse pass
    int max_k = 0;                // largest Fourier term count of the group's items
    // shape-gradient items (wos_batch_shape_gradient)
    GradItem *gitems = nullptr;   // device copy
    int n_gitems = 0;
    long long gsamples = 0, gpoints = 0;
    int *giprim = nullptr;
    float *gisd = nullptr;
    float4 *vstage = nullptr;     // staged gradient terms of every sample
    VgradRange *vranges = nullptr; // per item: its samples, vertex sums and scale
    double *vpart = nullptr;      // slice sums (kVgradSlices per item)
    int vwin = 1;                 // LDS windows of the vertex scatter (the largest item's)
};

}  // namespace wos

using namespace wos;

struct wos_batch {
    BatchGroup g[2];          // [0] 2D items, [1] 3D items
    // queue heads: walks (2D, 3D), shape-gradient samples (2D, 3D), start
    // queries of the walks and of the shape-gradient samples
    unsigned long long *counters = nullptr;
    int2 *spill = nullptr;    // traversal spill owned by the batch
    long spill_lanes = 0;
    int device = 0;
    // each item's scene and the scene's generation when the item table was
    // built: a setter that frees or re-types what the table points at
    // (source, grids, use_bvh) makes the next run fail instead of reading it
    std::vector<std::pair<const wos_scene *, unsigned long long>> scenes;
    std::vector<unsigned long long *> value_sums;   // per item with grad_values: its fixed-point sums
};

// defined in wos_capi.hip
namespace wos {
SceneView scene_view(const wos_scene *s);
unsigned long long scene_generation(const wos_scene *s);
}

#define BREQ(cond, ...)              \
    do {                             \
        if (!(cond)) {               \
            set_error(__VA_ARGS__);  \
            return -1;               \
        }                            \
    } while (0)

static void free_batch(wos_batch *b) {
    if (!b) return;
    for (auto &g : b->g) {
        (void)hipFree(g.items);
        (void)hipFree(g.iprim);
        (void)hipFree(g.isd);
        (void)hipFree(g.tiles);
        (void)hipFree(g.gitems);
        (void)hipFree(g.giprim);
        (void)hipFree(g.gisd);
        (void)hipFree(g.vstage);
        (void)hipFree(g.vranges);
        (void)hipFree(g.vpart);
    }
    (void)hipFree(b->counters);
    (void)hipFree(b->spill);
    for (unsigned long long *p : b->value_sums) (void)hipFree(p);
    delete b;
}

static int occupancy_grid_b() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (cus <= 0) cus = 256;
    }
    return cus * (g_tuning.wgs_per_cu > 0 ? g_tuning.wgs_per_cu : 8);
}

static int check_scenes(const wos_batch *B) {
    for (size_t j = 0; j < B->scenes.size(); ++j)
        BREQ(scene_generation(B->scenes[j].first) == B->scenes[j].second,
             "the scene of batch item %zu changed after the batch was created (source, grid or use_bvh); "
             "create the batch again", j);
    return 0;
}

extern "C" int wos_batch_create(const wos_batch_item *items, int n_items, wos_batch **out) {
    BREQ(out && items && n_items > 0, "items and out are required");
    *out = nullptr;
    std::vector<BatchItem> grp[2];
    std::vector<GradItem> ggrp[2];
    std::vector<std::pair<const wos_scene *, unsigned long long>> scenes;
    bool any_is = false;
    for (int j = 0; j < n_items; ++j) {
        const wos_batch_item &it = items[j];
        BREQ(it.scene, "item %d has no scene", j);
        BREQ(it.params.nwalks >= 1 && it.params.nsteps >= 0, "item %d: bad walk settings", j);
        BREQ(it.n_points >= 0 && (it.n_points == 0 || it.points), "item %d: bad points", j);
        BREQ(it.n_points == 0 || it.walk_value, "item %d: walk_value is required", j);
        BREQ((long long)it.n_points * it.params.nwalks < (1LL << 32), "item %d: at most 2^32 - 1 walks", j);
        SceneView v = scene_view(it.scene);
        v.n_top = 0;
        BatchItem b;
        b.v = v;
        b.w = WalkSettings{it.params.nwalks, it.params.nsteps, it.params.epsilon, it.params.double_sided,
                           it.params.prevent_fd_artifacts, it.params.use_IS_for_greens ? 1 : 0};
        any_is = any_is || (b.w.use_is && v.source_type != 0);
        b.points = it.points;
        b.n_points = it.n_points;
        b.seed = it.seed;
        b.K = 0;
        if (it.basis) {
            BREQ(v.dim == 2 && v.source_type == 3, "item %d: basis needs a 2D Fourier source", j);
            b.K = v.n_src / 4;
            BREQ(b.K <= kMaxBatchBasis, "item %d: at most %d Fourier terms", j, kMaxBatchBasis);
        }
        b.walk_value = it.walk_value;
        b.rec_prim = it.rec_prim;
        b.rec_t = it.rec_t;
        b.basis = reinterpret_cast<unsigned long long *>(it.basis);
        b.basis_scale = ldexp(1.0, basis_exponent(b.w.nwalks, b.w.nsteps));
        b.mean = it.mean;
        b.mean64 = it.mean64;
        b.var64 = it.var64;
        b.grad_out = it.grad_out;
        b.grad_values = it.grad_values;
        b.grad_coeffs = it.grad_coeffs;
        if (it.n_points > 0 && (it.grad_values || it.grad_coeffs)) {
            BREQ(it.grad_out, "item %d: the reverse pass needs grad_out", j);
            BREQ(it.rec_prim && it.rec_t, "item %d: the reverse pass needs the exit records (rec_prim, rec_t)", j);
            BREQ(!it.grad_coeffs || it.basis, "item %d: coefficient gradients need the basis output", j);
        }
        if (it.n_points > 0 && it.grad_vertices) {
            const wos_grad_params &gp = it.grad_params;
            BREQ(it.grad_out, "item %d: the shape gradient needs grad_out", j);
            BREQ(v.use_bvh, "item %d: the shape gradient needs the scene's BVH (use_bvh)", j);
            BREQ(gp.nwalks >= 1 && gp.nsteps >= 0 && gp.epsilon > 0.0f && gp.epsilon2 > 0.0f && gp.clamping >= 0.0f,
                 "item %d: bad shape-gradient settings", j);
            BREQ((long long)it.n_points * gp.nwalks < (1LL << 32), "item %d: at most 2^32 - 1 samples", j);
            GradItem gi;
            gi.v = v;
            gi.g = GradSettings{gp.nwalks, gp.nsteps, gp.epsilon, gp.epsilon2, gp.clamping, gp.double_sided,
                                gp.control_variates};
            gi.points = it.points;
            gi.n_points = it.n_points;
            gi.seed = it.seed;
            gi.grad_out = it.grad_out;
            gi.grad_vertices = it.grad_vertices;
            ggrp[v.dim == 2 ? 0 : 1].push_back(gi);
        }
        grp[v.dim == 2 ? 0 : 1].push_back(b);
        scenes.emplace_back(it.scene, scene_generation(it.scene));
    }
    for (int d = 0; d < 2; ++d)
        BREQ((int)grp[d].size() <= kMaxBatchItems, "at most %d items per dimension", kMaxBatchItems);
    if (any_is && ensure_green_cdf_batch() != 0) return -1;
    wos_batch *B = new wos_batch();
    B->scenes = std::move(scenes);
    for (auto &gv : grp)
        for (BatchItem &b : gv) {
            b.vacc = nullptr;
            if (b.n_points > 0 && b.grad_values) {
                unsigned long long *p = nullptr;
                if (hipMalloc((void **)&p, sizeof(unsigned long long) * ((size_t)b.v.n_vertices + 1)) != hipSuccess) {
                    (void)hipGetLastError();
                    free_batch(B);
                    set_error("cannot allocate the batch value-adjoint sums");
                    return -1;
                }
                B->value_sums.push_back(p);
                b.vacc = p;
            }
        }
    (void)hipGetDevice(&B->device);
    for (int d = 0; d < 2; ++d) {
        auto &v = grp[d];
        long long wb = 0, pb = 0;
        int depth = 0;
        std::vector<AdjointTile> tiles;
        for (int j = 0; j < (int)v.size(); ++j) {
            BatchItem &b = v[j];
            b.walk_begin = wb;
            b.point_begin = pb;
            const long long nw = (long long)b.n_points * b.w.nwalks;
            wb += nw;
            pb += b.n_points;
            depth = std::max(depth, b.v.use_bvh ? b.v.max_depth : 0);
            if (b.n_points > 0 && b.grad_values)
                for (long long t = 0; t < nw; t += kAdjointTile) tiles.push_back(AdjointTile{j, t, std::min(nw, t + kAdjointTile)});
            if (b.n_points > 0 && b.grad_values && b.v.n_vertices <= kBatchHist)
                B->g[d].hist = std::max(B->g[d].hist, b.v.n_vertices);
        }
        BatchGroup &g = B->g[d];
        g.n_items = (int)v.size();
        for (const BatchItem &b : v) g.max_k = std::max(g.max_k, b.K);
        g.walks = wb;
        g.points = pb;
        g.max_depth = depth;
        g.n_tiles = (int)tiles.size();
        {
            auto &gv = ggrp[d];
            long long sb = 0, gb = 0;
            for (GradItem &gi : gv) {
                gi.walk_begin = sb;
                gi.point_begin = gb;
                sb += (long long)gi.n_points * gi.g.nwalks;
                gb += gi.n_points;
            }
            g.n_gitems = (int)gv.size();
            g.gsamples = sb;
            g.gpoints = gb;
            if (g.n_gitems) {
                // staged terms of every sample, and per item its vertex sums
                std::vector<VgradRange> rv(gv.size());
                bool ok = hipMalloc((void **)&g.vstage, sizeof(float4) * (d == 0 ? 1 : 2) * (size_t)std::max(sb, 1LL)) ==
                              hipSuccess &&
                          hipMalloc((void **)&g.vranges, sizeof(VgradRange) * gv.size()) == hipSuccess &&
                          hipMalloc((void **)&g.vpart, sizeof(double) * kVgradSlices * gv.size()) == hipSuccess;
                for (size_t j = 0; ok && j < gv.size(); ++j) {
                    GradItem &gi = gv[j];
                    gi.vstage = g.vstage;
                    unsigned long long *acc = nullptr;
                    ok = hipMalloc((void **)&acc, sizeof(unsigned long long) * (size_t)gi.v.n_vertices * gi.v.dim) ==
                         hipSuccess;
                    if (ok) B->value_sums.push_back(acc);
                    g.vwin = std::max(g.vwin, vgrad_windows((long)gi.v.n_vertices * gi.v.dim));
                    rv[j] = VgradRange{gi.walk_begin, (long long)gi.n_points * gi.g.nwalks, gi.v.n_vertices, gi.v.idx,
                                       acc, gi.grad_vertices, 1.0};
                }
                ok = ok &&
                     hipMalloc((void **)&g.gitems, sizeof(GradItem) * gv.size()) == hipSuccess &&
                     hipMalloc((void **)&g.giprim, sizeof(int) * (size_t)std::max(gb, 1LL)) == hipSuccess &&
                     hipMalloc((void **)&g.gisd, sizeof(float) * (size_t)std::max(gb, 1LL)) == hipSuccess &&
                     hipMemcpy(g.gitems, gv.data(), sizeof(GradItem) * gv.size(), hipMemcpyHostToDevice) ==
                         hipSuccess &&
                     hipMemcpy(g.vranges, rv.data(), sizeof(VgradRange) * rv.size(), hipMemcpyHostToDevice) ==
                         hipSuccess;
                if (!ok) {
                    (void)hipGetLastError();
                    free_batch(B);
                    set_error("cannot allocate the batch shape-gradient tables");
                    return -1;
                }
            }
        }
        if (!g.n_items) continue;
        if (hipMalloc((void **)&g.items, sizeof(BatchItem) * v.size()) != hipSuccess ||
            hipMalloc((void **)&g.iprim, sizeof(int) * (size_t)std::max(pb, 1LL)) != hipSuccess ||
            hipMalloc((void **)&g.isd, sizeof(float) * (size_t)std::max(pb, 1LL)) != hipSuccess ||
            hipMemcpy(g.items, v.data(), sizeof(BatchItem) * v.size(), hipMemcpyHostToDevice) != hipSuccess ||
            (g.n_tiles > 0 &&
             (hipMalloc((void **)&g.tiles, sizeof(AdjointTile) * tiles.size()) != hipSuccess ||
              hipMemcpy(g.tiles, tiles.data(), sizeof(AdjointTile) * tiles.size(), hipMemcpyHostToDevice) !=
                  hipSuccess))) {
            (void)hipGetLastError();
            free_batch(B);
            set_error("cannot allocate the batch tables");
            return -1;
        }
    }
    // queue heads: primal walks of each group, then shape-gradient samples
    if (hipMalloc((void **)&B->counters, 6 * sizeof(unsigned long long)) != hipSuccess) {
        (void)hipGetLastError();
        free_batch(B);
        set_error("cannot allocate the batch counters");
        return -1;
    }
    int depth = std::max(B->g[0].max_depth, B->g[1].max_depth);
    long deep = (long)depth + 2 - kLdsStack;
    B->spill_lanes = (long)occupancy_grid_b() * kBlock;
    if (deep > 0 && hipMalloc((void **)&B->spill, sizeof(int2) * (size_t)deep * (size_t)B->spill_lanes) != hipSuccess) {
        (void)hipGetLastError();
        free_batch(B);
        set_error("cannot allocate the batch traversal spill");
        return -1;
    }
    *out = B;
    return 0;
}

extern "C" int wos_batch_destroy(wos_batch *b) {
    (void)hipDeviceSynchronize();
    free_batch(b);
    return 0;
}

extern "C" int wos_batch_run(wos_batch *B, void *stream) {
    BREQ(B, "batch is NULL");
    if (check_scenes(B)) return -1;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int cap = (int)(B->spill_lanes / kBlock);
    k_batch_zero<<<1, 64, 0, st>>>(B->counters, 2);
    for (int d = 0; d < 2; ++d) {
        BatchGroup &g = B->g[d];
        if (!g.n_items || !g.points) continue;
        long stride = B->spill_lanes;
        int2 *spill = B->spill;
        long gpts = (g.points + kBlock - 1) / kBlock;
        int gi = (int)std::min<long>(std::max<long>(gpts, 1), cap);
        long gw = (g.walks / 64 + 4) / 4;
        int gwk = (int)std::min<long>(std::max<long>(gw, 1), cap);
        int chunk = g_tuning.chunk > 0 ? g_tuning.chunk : 256;
        int window = window_for(d == 0 ? 2 : 3);
        const size_t hb = sizeof(long long) * (size_t)g.n_items;   // item table of k_batch_walk
        k_batch_basis<<<dim3(256, g.n_items), kBlock, 0, st>>>(g.items, 0);
        if (d == 0) {
            k_batch_zero<<<1, 64, 0, st>>>(B->counters + 4, 1);
            k_batch_init_persistent<2, BatchItem><<<gi, kBlock, hb, st>>>(g.items, g.n_items, g.points, B->counters + 4,
                                                                          kInitChunk, window, g.iprim, g.isd, spill,
                                                                          stride);
            if (g.max_k <= 8)
                k_batch_walk<2, 8><<<gwk, kBlock, hb, st>>>(g.items, g.n_items, g.walks, B->counters + d, chunk,
                                                           window, g.iprim, g.isd, spill, stride);
            else
                k_batch_walk<2, kMaxBatchBasis><<<gwk, kBlock, hb, st>>>(g.items, g.n_items, g.walks, B->counters + d,
                                                                        chunk, window, g.iprim, g.isd, spill, stride);
        } else {
            k_batch_zero<<<1, 64, 0, st>>>(B->counters + 4, 1);
            k_batch_init_persistent<3, BatchItem><<<gi, kBlock, hb, st>>>(g.items, g.n_items, g.points, B->counters + 4,
                                                                          kInitChunk, window, g.iprim, g.isd, spill,
                                                                          stride);
            k_batch_walk<3, 0><<<gwk, kBlock, hb, st>>>(g.items, g.n_items, g.walks, B->counters + d, chunk, window,
                                                       g.iprim, g.isd, spill, stride);
        }
        k_batch_basis<<<dim3(256, g.n_items), kBlock, 0, st>>>(g.items, 1);
        long gr = std::min<long>((g.points * 64 + kBlock - 1) / kBlock, 8192);
        k_batch_reduce<<<(int)std::max<long>(gr, 1), kBlock, 0, st>>>(g.items, g.n_items, g.points);
    }
    hipError_t e = hipGetLastError();
    BREQ(e == hipSuccess, "batch launch: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int wos_batch_shape_gradient(wos_batch *B, void *stream) {
    BREQ(B, "batch is NULL");
    if (check_scenes(B)) return -1;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int cap = (int)(B->spill_lanes / kBlock);
    k_batch_zero<<<1, 64, 0, st>>>(B->counters + 2, 2);
    for (int d = 0; d < 2; ++d) {
        BatchGroup &g = B->g[d];
        if (!g.n_gitems || !g.gsamples) continue;
        const long stride = B->spill_lanes;
        const int gi = (int)std::min<long>(std::max<long>((g.gpoints + kBlock - 1) / kBlock, 1), cap);
        const int gw = (int)std::min<long>(std::max<long>((g.gsamples / 64 + 4) / 4, 1), cap);
        const int chunk = g_tuning.chunk > 0 ? g_tuning.chunk : 128;
        const int window = window_for(d == 0 ? 2 : 3);
        const size_t hg = sizeof(long long) * (size_t)g.n_gitems;   // item table of k_batch_grad
        if (d == 0) {
            k_batch_zero<<<1, 64, 0, st>>>(B->counters + 5, 1);
            k_batch_init_persistent<2, GradItem><<<gi, kBlock, hg, st>>>(g.gitems, g.n_gitems, g.gpoints,
                                                                         B->counters + 5, kInitChunk, window, g.giprim,
                                                                         g.gisd, B->spill, stride);
            k_batch_grad<2><<<gw, kBlock, hg, st>>>(g.gitems, g.n_gitems, g.gsamples, B->counters + 2 + d, chunk,
                                                    window, g.giprim, g.gisd, B->spill, stride);
            k_vgrad_bound<2><<<dim3(kVgradSlices, g.n_gitems), kBlock, 0, st>>>(g.vstage, g.vranges, g.vpart);
            k_vgrad_scale<<<g.n_gitems, 64, 0, st>>>(g.vranges, g.vpart);
            k_vgrad_scatter<2><<<dim3(64, g.n_gitems, g.vwin), kBlock, sizeof(unsigned long long) * kVgradLds, st>>>(
                g.vstage, g.vranges, kVgradLds, nullptr);
            k_vgrad_finish<2><<<dim3(64, g.n_gitems), kBlock, 0, st>>>(g.vranges);
        } else {
            k_batch_zero<<<1, 64, 0, st>>>(B->counters + 5, 1);
            k_batch_init_persistent<3, GradItem><<<gi, kBlock, hg, st>>>(g.gitems, g.n_gitems, g.gpoints,
                                                                         B->counters + 5, kInitChunk, window, g.giprim,
                                                                         g.gisd, B->spill, stride);
            k_batch_grad<3><<<gw, kBlock, hg, st>>>(g.gitems, g.n_gitems, g.gsamples, B->counters + 2 + d, chunk,
                                                    window, g.giprim, g.gisd, B->spill, stride);
            k_vgrad_bound<3><<<dim3(kVgradSlices, g.n_gitems), kBlock, 0, st>>>(g.vstage, g.vranges, g.vpart);
            k_vgrad_scale<<<g.n_gitems, 64, 0, st>>>(g.vranges, g.vpart);
            k_vgrad_scatter<3><<<dim3(64, g.n_gitems, g.vwin), kBlock, sizeof(unsigned long long) * kVgradLds, st>>>(
                g.vstage, g.vranges, kVgradLds, nullptr);
            k_vgrad_finish<3><<<dim3(64, g.n_gitems), kBlock, 0, st>>>(g.vranges);
        }
    }
    hipError_t e = hipGetLastError();
    BREQ(e == hipSuccess, "batch shape-gradient launch: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int wos_batch_adjoint(wos_batch *B, void *stream) {
    BREQ(B, "batch is NULL");
    if (check_scenes(B)) return -1;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int cap = (int)(B->spill_lanes / kBlock);
    for (int d = 0; d < 2; ++d) {
        BatchGroup &g = B->g[d];
        if (d == 0 && g.n_items) k_batch_adjoint_coeffs<<<dim3(kMaxBatchBasis, g.n_items), kBlock, 0, st>>>(g.items);
        if (!g.n_tiles) continue;
        const int grid = std::max(1, std::min(g.n_tiles, cap));
        const size_t hb = sizeof(unsigned long long) * (size_t)g.hist;
        k_batch_value_prep<<<g.n_items, kBlock, 0, st>>>(g.items);
        if (d == 0) k_batch_adjoint<2><<<grid, kBlock, hb, st>>>(g.items, g.tiles, g.n_tiles, g.hist);
        else k_batch_adjoint<3><<<grid, kBlock, hb, st>>>(g.items, g.tiles, g.n_tiles, g.hist);
        k_batch_value_finish<<<dim3(64, g.n_items), kBlock, 0, st>>>(g.items);
    }
    hipError_t e = hipGetLastError();
    BREQ(e == hipSuccess, "batch adjoint launch: %s", hipGetErrorString(e));
    return 0;
}
