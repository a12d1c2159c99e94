"""Benchmark: primal + gradient walks/s of the 2D Poisson clover workload.

Workload (one step, the default): 126,820 query points on a 512x512 grid
inside the clover r = 1 + 0.3 cos 4t (4,096 segments), 128 walks per point
(16.2M walks), eps = 1e-4, 256 steps, source sin3x cos2y as 8 Fourier terms
sampled uniformly in each ball (the interior sampling the reference's CUDA
walk executes), Dirichlet data g = x^2 - y^2.  A step solves u, forms L =
mean((u - u_target)^2) against a fixed target solved at 1.1 g, and
back-propagates to the 4,096 vertex values and 8 source coefficients.
Inputs are resident in HBM before timing.

Also reported: the same step through the reference-style path ("plain": one
lane per walk, brute-force closest point, per-walk records, walks re-run in
the backward pass) with the RMSE of both paths on the same points, the
importance-sampled variant, the roofline of the dominant kernel and the CPU
baseline (the oracle with its CPU LBVH on the box's 16-core share).

Other workloads (--workload): circle / sphere (primal Laplace, 1,024-segment
circle and 81,920-triangle displaced icosphere with 1M points), shape2d /
shape3d (shape gradients), batch (64 uneven scenes, primal + adjoint in one
captured graph, against the plain path at equal RMSE).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-plain] [--no-cpu]
       [--no-counts]

The HBM traffic of a walk kernel comes from the FETCH_SIZE / WRITE_SIZE
counter runs summarised under profiles/ (rocprofv3 cannot run inside this
process); it is null when the summary is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "diff-wos_amd"))

PEAK_FP32 = 157.3e12      # MI355X FP32 vector, FLOP/s
PEAK_HBM = 8.0e12         # bytes/s
CPU_THREADS = 16          # a one-GPU job's share of the host's cores
METRIC = "primal+grad walks/s at 1 GPU, batched vs plain, equal RMSE; % VALU/HBM peak"

# algorithmic flop model (FMA = 2): per bound test, per primitive test, per
# walk step (direction, move, bound), per source sample and Fourier term,
# per walk.  2D: segment 20, box 11; 3D: triangle 85, box 17.
C_BOX = {2: 11, 3: 17}
C_PRIM = {2: 20, 3: 85}
C_STEP, C_SRC, C_TERM, C_WALK = 30, 25, 34, 15


def measured_traffic(name):
    """HBM bytes per walk launch recorded by the counter runs, or None."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def emit(line):
    print(json.dumps(line), flush=True)


def synced_ms(torch, fn, steps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        fn(i)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def roofline(flops, kernel_ms, kernel, traffic, extra=None):
    s = kernel_ms * 1e-3
    achieved = flops / s
    r = {"bound": "valu", "achieved": round(achieved / 1e12, 3), "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
         "frac": round(achieved / PEAK_FP32, 4), "traffic": traffic, "kernel": kernel,
         "flops_per_launch": float(flops), "launch_ms": round(kernel_ms, 3)}
    if traffic:
        r["hbm_GBs"] = round(traffic / s / 1e9, 1)
        r["hbm_frac"] = round(traffic / s / PEAK_HBM, 4)
    if extra:
        r.update(extra)
    return r


def walk_flops(counts, dim, walks, K=0, src=False):
    per_step = C_STEP + ((C_SRC + C_TERM * K) if src else 0)
    return (counts["box_tests"] * C_BOX[dim] + counts["prim_tests"] * C_PRIM[dim] + counts["steps"] * per_step
            + walks * C_WALK)


def cpu_leg(run_chunk, n_points, W, seconds, what):
    """Times run_chunk(index array) over an evenly strided sample of the
    points, growing the sample until about `seconds` of CPU work."""
    done, tc = 0, 0.0
    stride = max(1, n_points // 4096)
    order = np.concatenate([np.arange(o, n_points, stride) for o in range(stride)])
    blk = 64
    while tc < seconds and done < n_points:
        idx = order[done: done + blk]
        ta = time.perf_counter()
        run_chunk(idx)
        tc += time.perf_counter() - ta
        done += idx.size
        blk = min(blk * 2, 4096)
    return {"value": done * W / tc, "unit": "walks/s", "cores": CPU_THREADS, "kind": "port",
            "sample": f"{done} of {n_points} points, evenly strided (stride {stride}) x {W} walks, {what}, "
                      f"oracle with its CPU LBVH, {tc:.1f} s"}


def plain_step(torch, scene_bf, cwos, pts, target, seed, W):
    """Reference data flow: PCG32 per walk (Dr.Jit PCG32 + TEA), one lane per
    walk with the brute-force scan, records to HBM, dirichlet + scatter mean;
    backward re-runs the walks from the same seed and scatters the
    interpolation weights (2D: 1 - t, t; 3D: 1 - u - v, u, v)."""
    from wos.fwd import PCG32
    from wos.utils import sample_tea_32
    P = pts.shape[0]
    idx = torch.arange(P * W, device=pts.device) // W
    prep = pts.repeat_interleave(W, dim=0)

    def walk():
        v0, v1 = sample_tea_32(seed, idx)
        smp = PCG32(P * W, v0.to(torch.int64), v1.to(torch.int64))
        return cwos.single_walk(prep, scene_bf, smp)

    rec = walk()
    val = torch.where(rec.valid, scene_bf.dirichlet(rec) + rec.contrib, torch.zeros_like(rec.contrib))
    u = torch.zeros(P, device=pts.device).index_add_(0, idx, val) / W
    g = (2.0 / P) * (u - target)
    rec = walk()   # backward: recompute the paths
    gw = torch.where(rec.valid, g[idx] / W, torch.zeros_like(rec.contrib))
    f = scene_bf.c_indices[rec.prim_id.clamp(min=0).long()].long()
    gv = torch.zeros(scene_bf.c_values.numel(), device=pts.device)
    if scene_bf.dim == 2:
        t = rec.t
        gv.index_add_(0, f[:, 0], gw * (1 - t)).index_add_(0, f[:, 1], gw * t)
    else:
        uv = rec.uv
        gv.index_add_(0, f[:, 0], gw * (1 - uv[:, 0] - uv[:, 1])).index_add_(0, f[:, 1], gw * uv[:, 0])
        gv.index_add_(0, f[:, 2], gw * uv[:, 1])
    return u, gv


def fd_values_and_coefficients(step, wos_ext, t, cwos, V, F, G, SP, pts, target):
    """The value / coefficient adjoint against central differences (h = 1e-3, the walks of the adjoint's
    seed on both sides): a random direction in the vertex values and one in the source coefficients."""
    gvals, gcoef = step(0, timed=False)
    rng = np.random.default_rng(3)
    dvals = t(rng.normal(size=G.shape[0]).astype(np.float32))
    dcoef = t(rng.normal(size=SP.shape[0]).astype(np.float32))
    h = 1e-3
    t64 = target.double()
    loss = lambda sc: float(((cwos.walk_reduce(pts, sc, seed=100, stats=True)["mean64"] - t64) ** 2).mean())
    fd = {}
    for name, dirn, grad in (("values", dvals, gvals), ("source_coefficients", dcoef, gcoef)):
        diff = 0.0
        for sgn in (1, -1):
            if name == "values":
                sc = wos_ext.Scene(V, F, G + sgn * h * dirn, use_bvh=True, source_type=3, source_params=SP)
            else:
                sp2 = SP.clone()
                sp2[:, 3] += sgn * h * dirn
                sc = wos_ext.Scene(V, F, G, use_bvh=True, source_type=3, source_params=sp2)
            diff += sgn * loss(sc)
        cd = diff / (2 * h)
        a = float((grad.double() * dirn.double()).sum())
        fd[name] = {"adjoint": a, "central_difference": cd, "relative_error": abs(a - cd) / abs(cd)}
    log(f"fd check (h = 1e-3): values {fd['values']['relative_error']:.2e}, coefficients "
        f"{fd['source_coefficients']['relative_error']:.2e}")
    return fd


# ---------------------------------------------------------------------------
# default: the 2D Poisson clover, primal + value / coefficient adjoint
# ---------------------------------------------------------------------------
def clover_workload(args, torch, wos_ext, C):
    c = C.config2()
    dv = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dv)
    V, F, G, pts, SP = t(c["vertices"]), t(c["indices"]), t(c["values"]), t(c["points"]), t(c["source_params"])
    P, W = pts.shape[0], c["nwalks"]
    log(f"workload: {P} points x {W} walks, 4096 segments, eps 1e-4")
    scene = wos_ext.Scene(V, F, G, use_bvh=True, source_type=3, source_params=SP)
    cwos = wos_ext.WoS(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"], use_IS_for_greens=False)
    tscene = wos_ext.Scene(V, F, G * 1.1, use_bvh=True, source_type=3, source_params=SP)
    target = cwos.walk_reduce(pts, tscene, seed=12345)["mean"]
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * max(args.steps, 1))]

    def step(i, walker=cwos, timed=True):
        if timed:
            ev[2 * i].record(stream)
        out = walker.walk_reduce(pts, scene, seed=100 + i, records=True, basis=True)
        if timed:
            ev[2 * i + 1].record(stream)
        u = out["mean"]
        g = (2.0 / P) * (u - target)
        return wos_ext.adjoint_values(scene, P, W, out["rec_prim"], out["rec_t"], g, basis=out["basis"],
                                      want_values=True, want_coeffs=True)

    for i in range(args.warmup):
        step(i, timed=False)
    ms_per_step = synced_ms(torch, step, args.steps)
    kernel_ms = float(np.mean([ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(args.steps)]))
    walks = P * W
    value = walks / (ms_per_step * 1e-3)
    log(f"fused: {ms_per_step:.2f} ms/step, {value:.3e} walks/s; walk launch {kernel_ms:.2f} ms")

    # the adjoint against central differences (h = 1e-3, the walks of the adjoint's seed on both sides):
    # a random direction in the 4,096 vertex values and one in the 8 source coefficients
    fd = None
    if not args.no_fd:
        fd = fd_values_and_coefficients(step, wos_ext, t, cwos, V, F, G, SP, pts, target)

    rf = None
    if not args.no_counts:
        counts = cwos.walk_counts(pts, scene, seed=100)
        K = c["source_params"].shape[0]
        flops = walk_flops(counts, 2, walks, K=K, src=True)
        rf = roofline(flops, kernel_ms, "k_walk_persistent<2,true,8,false>", measured_traffic("walk_clover_hbm.md"),
                      {"per_query": {"box_tests": counts["box_tests"] / counts["queries"],
                                     "prim_tests": counts["prim_tests"] / counts["queries"]},
                       "mean_steps": counts["steps"] / walks})

    # accuracy: the manufactured twin (exact u known), both interior samplers
    ct = C.config2(twin=True)
    tw = wos_ext.Scene(V, F, t(ct["values"]), use_bvh=True, source_type=3, source_params=SP)
    o = cwos.walk_reduce(pts, tw, seed=7, stats=True)
    m_fused = o["mean64"].cpu().numpy()
    rmse = float(np.sqrt(np.mean((m_fused - ct["exact"]) ** 2)))
    is_variant = None
    if not args.no_variants:
        cis = wos_ext.WoS(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"], use_IS_for_greens=True)
        step(0, walker=cis, timed=False)
        ms_is = synced_ms(torch, lambda i: step(i, walker=cis, timed=False), max(2, args.steps // 2))
        o_is = cis.walk_reduce(pts, tw, seed=7, stats=True)
        rmse_is = float(np.sqrt(np.mean((o_is["mean64"].cpu().numpy() - ct["exact"]) ** 2)))
        is_variant = {"ms_per_step": ms_is, "walks_per_s": walks / (ms_is * 1e-3), "rmse_twin": rmse_is,
                      "interior_sampling": "tabulated Green's-function CDF"}
        log(f"importance-sampled interior: {ms_is:.2f} ms/step, rmse {rmse_is:.4f} (uniform {rmse:.4f})")

    plain = None
    if not args.no_plain:
        scene_bf = wos_ext.Scene(V, F, G, use_bvh=False, source_type=3, source_params=SP)
        sub = slice(None, None, 8)   # every 8th grid point keeps the brute-force run short
        psub, tsub = pts[sub].contiguous(), target[sub].contiguous()
        plain_step(torch, scene_bf, cwos, psub, tsub, 0, W)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        plain_step(torch, scene_bf, cwos, psub, tsub, 1, W)
        torch.cuda.synchronize()
        tp = time.perf_counter() - ta
        # the same data flow over this library's LBVH instead of the reference's brute-force scan
        plain_step(torch, scene, cwos, psub, tsub, 2, W)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        plain_step(torch, scene, cwos, psub, tsub, 3, W)
        torch.cuda.synchronize()
        tpb = time.perf_counter() - ta
        # equal RMSE: both paths on the twin, same points, same walks per point
        tw_bf = wos_ext.Scene(V, F, t(ct["values"]), use_bvh=False, source_type=3, source_params=SP)
        u_plain, _ = plain_step(torch, tw_bf, cwos, psub, tsub, 7, W)
        ex = ct["exact"][sub]
        rmse_plain = float(np.sqrt(np.mean((u_plain.double().cpu().numpy() - ex) ** 2)))
        rmse_fused_sub = float(np.sqrt(np.mean((m_fused[sub] - ex) ** 2)))
        plain = {"walks_per_s": psub.shape[0] * W / tp, "ms_per_step_sample": tp * 1e3,
                 "sample": f"{psub.shape[0]} points (every 8th) x {W} walks, one lane per walk, walks re-run in "
                           f"backward; brute-force scan (walks_per_s) and LBVH (walks_per_s_lbvh)",
                 "speedup_fused_vs_plain": value / (psub.shape[0] * W / tp),
                 "walks_per_s_lbvh": psub.shape[0] * W / tpb,
                 "speedup_fused_vs_plain_lbvh": value / (psub.shape[0] * W / tpb),
                 "rmse_twin_plain": rmse_plain, "rmse_twin_fused_same_points": rmse_fused_sub}
        log(f"plain: {plain['walks_per_s']:.3e} walks/s (LBVH {plain['walks_per_s_lbvh']:.3e}), rmse "
            f"{rmse_plain:.4f} vs fused {rmse_fused_sub:.4f}")

    cpu = None
    if not args.no_cpu:
        import oracle as O
        ref = O.Scene(c["vertices"], c["indices"], c["values"], 3, c["source_params"], use_bvh=True)
        pn = c["points"]
        gt = target.cpu().numpy().astype(np.float64)

        def chunk(ix):
            q = np.ascontiguousarray(pn[ix])
            r = ref.walk_reduce(q, W, c["nsteps"], c["epsilon"], seed=0, threads=CPU_THREADS, records=True)
            ref.adjoint_values(q.shape[0], W, r["rec_prim"], r["rec_t"], (2.0 / P) * (r["mean"] - gt[ix]))

        cpu = cpu_leg(chunk, P, W, args.cpu_seconds, "primal + value adjoint")
        log(f"cpu: {cpu['value']:.3e} walks/s on {CPU_THREADS} threads")

    emit({"metric": METRIC,
          "value": value, "unit": "walks/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
          "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
          "dtype": "f32", "data": "synthetic",
          "config": {"workload": "2d-poisson-clover-4096seg-512grid-128walks-eps1e-4-fourier8",
                     "points": P, "walks_per_point": W, "segments": 4096, "interior_sampling": "uniform",
                     "parallelism": "single-gpu"},
          "rmse_twin": rmse, "fd_check": fd, "roofline": rf, "cpu_baseline": cpu, "plain": plain,
          "importance_sampled": is_variant})


# ---------------------------------------------------------------------------
# primal Laplace: circle (config 1) and the displaced icosphere (config 3)
# ---------------------------------------------------------------------------
def primal_workload(args, torch, wos_ext, C):
    c = C.config1() if args.workload == "circle" else C.config3()
    dim = c["vertices"].shape[1]
    dv = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dv)
    V, F, G, pts = t(c["vertices"]), t(c["indices"]), t(c["values"]), t(c["points"])
    Scene, WoS = (wos_ext.Scene, wos_ext.WoS) if dim == 2 else (wos_ext.Scene3D, wos_ext.WoS3D)
    Scene(V, F, G, use_bvh=True)   # warm the build kernels
    builds = []
    for _ in range(3):
        torch.cuda.synchronize()
        tb = time.perf_counter()
        scene = Scene(V, F, G, use_bvh=True)
        torch.cuda.synchronize()
        builds.append((time.perf_counter() - tb) * 1e3)
    build_ms = float(np.median(builds))
    P, W = pts.shape[0], c["nwalks"]
    cw = WoS(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"])
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * max(args.steps, 1))]
    outs = {}

    def step(i):
        ev[2 * i].record(stream)
        outs["o"] = cw.walk_reduce(pts, scene, seed=100 + i, stats=True)
        ev[2 * i + 1].record(stream)

    for i in range(args.warmup):
        cw.walk_reduce(pts, scene, seed=i)
    ms = synced_ms(torch, step, args.steps)
    kernel_ms = float(np.mean([ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(args.steps)]))
    m, var = outs["o"]["mean64"].cpu().numpy(), outs["o"]["var64"].cpu().numpy()
    rmse = float(np.sqrt(np.mean((m - c["exact"]) ** 2)))
    mc_bound = float(np.sqrt(np.mean(var / W)))
    # closest-point queries/s (LBVH traversal alone, unbounded queries from the points)
    scene.closest_point_preliminary(pts)
    qms = synced_ms(torch, lambda i: scene.closest_point_preliminary(pts), 3)
    qps = P / (qms * 1e-3)
    walks = P * W
    rf = None
    cnt = None
    name = "walk_circle_hbm.md" if dim == 2 else "walk_sphere_hbm.md"
    if not args.no_counts:
        cnt = cw.walk_counts(pts, scene, seed=100, basis=False)
        rf = roofline(walk_flops(cnt, dim, walks), kernel_ms, f"k_walk_persistent<{dim},false,0,false>",
                      measured_traffic(name),
                      {"per_query": {"box_tests": cnt["box_tests"] / cnt["queries"],
                                     "prim_tests": cnt["prim_tests"] / cnt["queries"]},
                       "mean_steps": cnt["steps"] / walks})
    cpu = None
    if not args.no_cpu:
        import oracle as O
        ref = O.Scene(c["vertices"], c["indices"], c["values"], use_bvh=True)
        pn = c["points"]
        cpu = cpu_leg(lambda ix: ref.walk_reduce(np.ascontiguousarray(pn[ix]), W, c["nsteps"], c["epsilon"], seed=0,
                                                 threads=CPU_THREADS), P, W, args.cpu_seconds, "primal")
    log(f"{args.workload}: {ms:.2f} ms/solve, {walks / ms * 1e3:.3e} walks/s, rmse {rmse:.4f} (Monte Carlo bound "
        f"{mc_bound:.4f}), build {build_ms:.2f} ms, {qps:.3e} queries/s")
    emit({"metric": "walks/s (primal)", "value": walks / ms * 1e3, "unit": "walks/s", "n_gpus": 1,
          "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
          "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
          "config": {"workload": ("2d-laplace-circle-1024seg-11428pts-64walks" if dim == 2 else
                                  "3d-laplace-icosphere6-81920tris-1Mpts-32walks"),
                     "points": P, "walks_per_point": W, "primitives": int(c["indices"].shape[0])},
          "rmse_exact": rmse, "monte_carlo_bound": mc_bound, "lbvh_build_ms": build_ms,
          "closest_point_queries_per_s": qps, "counts": cnt, "roofline": rf, "cpu_baseline": cpu})


# ---------------------------------------------------------------------------
# shape gradients: the clover (2D) and the 200k-triangle sphere (config 4)
# ---------------------------------------------------------------------------
def shape_workload(args, torch, wos_ext, C):
    dim = 2 if args.workload == "shape2d" else 3
    dv = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dv)
    if dim == 2:
        c = C.config2()
        V, F, G, pts = t(c["vertices"]), t(c["indices"]), t(c["values"]), t(c["points"])
        P, W = pts.shape[0], 16
        scene = wos_ext.Scene(V, F, G, use_bvh=True)
        cw = wos_ext.WoS(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"])
        kw = dict(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"], epsilon2=c["epsilon"], clamping=0.1)
        target = cw.walk_reduce(pts, wos_ext.Scene(V, F, G * 1.1, use_bvh=True), seed=12345)["mean"]
        wl = "2d-shape-gradient-clover-4096seg-126820pts-16walks"
    else:
        c = C.config4()
        V, F, G, pts = t(c["vertices"]), t(c["indices"]), t(c["values"]), t(c["points"])
        target = t(c["target"].astype(np.float32))
        P, W = pts.shape[0], c["nwalks"]
        scene = wos_ext.Scene3D(V, F, G, use_bvh=True)
        cw = wos_ext.WoS3D(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"])
        kw = dict(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"], epsilon2=c["epsilon2"], clamping=c["clamping"])
        wl = "3d-shape-gradient-200k-tris-262144pts-16walks"
    stream = torch.cuda.current_stream()
    evs = []

    def step(i):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        ev[0].record(stream)
        u = cw.walk_reduce(pts, scene, seed=100 + i)["mean"]
        g = (2.0 / P) * (u - target)
        ev[1].record(stream)
        gv, _ = wos_ext.shape_gradient(scene, pts, 100 + i, g, **kw)
        ev[2].record(stream)
        evs.append(ev)
        return gv

    for i in range(args.warmup):
        step(i)
    evs.clear()
    ms = synced_ms(torch, step, args.steps)
    fwd = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    bwd = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    rf = None
    if not args.no_counts:
        cnt = wos_ext.shape_gradient_counts(scene, pts, 100, **kw)
        flops = (cnt["box_tests"] * C_BOX[dim] + cnt["prim_tests"] * C_PRIM[dim] + cnt["steps"] * C_STEP
                 + P * W * C_WALK * 3)
        rf = roofline(flops, bwd, f"k_grad_persistent<{dim}>", None,
                      {"per_query": {"box_tests": cnt["box_tests"] / max(cnt["queries"], 1),
                                     "prim_tests": cnt["prim_tests"] / max(cnt["queries"], 1)},
                       "steps_per_sample": cnt["steps"] / (P * W)})
    fd = None
    if not args.no_fd:
        # the adjoint of a smooth anisotropic stretch (x += 0.5 x, last axis -= 0.3 z) against central
        # differences of the loss; forward walks for u use other seeds than the gradient samples, the
        # differences use common seeds on both sides
        Sc = wos_ext.Scene if dim == 2 else wos_ext.Scene3D
        Wc = wos_ext.WoS if dim == 2 else wos_ext.WoS3D
        d = torch.zeros_like(V)
        d[:, 0] = 0.5 * V[:, 0]
        d[:, dim - 1] = -0.3 * V[:, dim - 1]
        n_ad, n_fd, h, W_fd = (8, 4, 1e-2, 128) if dim == 2 else (16, 3, 1e-2, 128)
        # points that stay inside under both perturbations (the clover's grid reaches the boundary)
        inner = scene.sdf_native(pts) < -0.02
        pf, tf = pts[inner].contiguous(), target[inner].contiguous()
        Pf = pf.shape[0]
        ad = []
        for sd in range(n_ad):
            u = cw.walk_reduce(pf, scene, seed=300 + sd)["mean"]
            gv, _ = wos_ext.shape_gradient(scene, pf, 500 + sd, (2.0 / Pf) * (u - tf), **kw)
            ad.append(float((gv.double() * d.double()).sum()))
        cfd = Wc(nwalks=W_fd, nsteps=kw["nsteps"], epsilon=kw["epsilon"])
        t64 = tf.double()
        fds = []
        for sd in range(n_fd):
            up = cfd.walk_reduce(pf, Sc(V + h * d, F, G, use_bvh=True), seed=900 + sd, stats=True)["mean64"]
            dn = cfd.walk_reduce(pf, Sc(V - h * d, F, G, use_bvh=True), seed=900 + sd, stats=True)["mean64"]
            fds.append(float((((up - t64) ** 2).mean() - ((dn - t64) ** 2).mean()) / (2 * h)))
        a, sa = float(np.mean(ad)), float(np.std(ad, ddof=1) / np.sqrt(n_ad))
        b, sb = float(np.mean(fds)), float(np.std(fds, ddof=1) / np.sqrt(n_fd))
        fd = {"direction": "stretch x by 0.5 x, last axis by -0.3 (coordinate)", "adjoint": a, "adjoint_se": sa,
              "central_difference": b, "central_difference_se": sb, "relative_error": abs(a - b) / abs(b),
              "sample": f"{Pf} of {P} points (at least 0.02 inside), adjoint over {n_ad} seeds x {W} walks, "
                        f"differences h = {h} over {n_fd} seeds x {W_fd} walks"}
        log(f"fd check: adjoint {a:+.5e} +- {sa:.1e}, central difference {b:+.5e} +- {sb:.1e}, relative error "
            f"{fd['relative_error']:.4f}")

    cpu = None
    if not args.no_cpu:
        import oracle as O
        ref = O.Scene(c["vertices"], c["indices"], c["values"], use_bvh=True)
        pn = c["points"]
        gt = target.cpu().numpy().astype(np.float64)

        def chunk(ix):
            q = np.ascontiguousarray(pn[ix])
            r = ref.walk_reduce(q, W, kw["nsteps"], kw["epsilon"], seed=0, threads=CPU_THREADS)
            ref.shape_gradient(q, W, kw["nsteps"], kw["epsilon"], kw["epsilon2"], kw["clamping"], seed=0,
                               grad_out=(2.0 / P) * (r["mean"] - gt[ix]), threads=CPU_THREADS)

        cpu = cpu_leg(chunk, P, W, args.cpu_seconds, "primal walk + shape-gradient sample")
    log(f"{args.workload}: {ms:.2f} ms/step (forward {fwd:.2f}, gradient {bwd:.2f}), {P * W / ms * 1e3:.3e} walks/s")
    emit({"metric": f"primal+grad walks/s ({dim}D shape gradient)", "value": P * W / ms * 1e3, "unit": "walks/s",
          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
          "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
          "config": {"workload": wl, "points": P, "walks_per_point": W, "primitives": int(c["indices"].shape[0])},
          "forward_ms": fwd, "gradient_ms": bwd, "fd_check": fd, "roofline": rf, "cpu_baseline": cpu})


# ---------------------------------------------------------------------------
# 64 uneven scenes (config 5): batched primal + adjoint against the plain path
# ---------------------------------------------------------------------------
def batch_workload(args, torch, wos_ext, C):
    items_np = C.config5(scale=args.batch_scale)
    dv = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dv)
    items, total = [], 0
    for it in items_np:
        src = "source_params" in it
        if it["dim"] == 2:
            sc = wos_ext.Scene(t(it["vertices"]), t(it["indices"]), t(it["values"]), use_bvh=True,
                               source_type=3 if src else 0, source_params=t(it["source_params"]) if src else None)
            cw = wos_ext.WoS(nwalks=it["nwalks"], nsteps=it["nsteps"], epsilon=it["epsilon"], use_IS_for_greens=False)
        else:
            sc = wos_ext.Scene3D(t(it["vertices"]), t(it["indices"]), t(it["values"]), use_bvh=True)
            cw = wos_ext.WoS3D(nwalks=it["nwalks"], nsteps=it["nsteps"], epsilon=it["epsilon"])
        gpar = dict(nwalks=args.batch_grad_walks, nsteps=it["nsteps"], epsilon=it["epsilon"],
                    epsilon2=5e-4 if it["dim"] == 2 else 1e-3, clamping=0.1)
        items.append(dict(scene=sc, wos=cw, points=t(it["points"]), seed=it["seed"], grad_values=True,
                          grad_coeffs=src, basis=src, grad_vertices=args.batch_grad_walks > 0, grad_params=gpar))
        total += it["points"].shape[0] * it["nwalks"]
    samples = sum(it["points"].shape[0] for it in items_np) * args.batch_grad_walks
    batch = wos_ext.Batch(items)
    tuning = batch.autotune() if args.autotune else None
    if tuning:
        log(f"autotuned: {tuning}")
    targets = [t((1.1 * it["exact"]).astype(np.float32)) for it in items_np]

    def step(_=0):
        batch.run()
        for o, tg in zip(batch.outputs, targets):
            o["grad_out"].copy_((2.0 / tg.numel()) * (o["mean"] - tg))
        batch.zero_grad()
        batch.adjoint()

    step()
    eager_ms = synced_ms(torch, step, args.steps)
    # the walk launches alone (both dimensions), HIP events on the launch stream
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps)]
    for i in range(args.steps):
        ev[2 * i].record(stream)
        batch.run()
        ev[2 * i + 1].record(stream)
    torch.cuda.synchronize()
    run_ms = float(np.mean([ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(args.steps)]))
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        step()
    graph_ms = synced_ms(torch, lambda i: g.replay(), args.steps)
    # accuracy of the batched means (FP64) against each item's exact solution
    err2, npts = 0.0, 0
    for o, it in zip(batch.outputs, items_np):
        e = o["mean64"].cpu().numpy() - it["exact"]
        err2 += float((e * e).sum())
        npts += e.size
    rmse_batched = float(np.sqrt(err2 / npts))
    log(f"batch: {total} walks; batched primal+adjoint {eager_ms:.1f} ms eager, {graph_ms:.1f} ms graph, "
        f"rmse {rmse_batched:.4f}")

    # the same step with every item's shape gradient (OursCUDA / Ours3D samples) in the graph as well
    shape = None
    if samples:
        def step_shape(_=0):
            step()
            batch.shape_gradient()

        step_shape()
        g2 = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(g2, stream=s):
            step_shape()
        graph2_ms = synced_ms(torch, lambda i: g2.replay(), args.steps)
        for i in range(args.steps):
            ev[2 * i].record(stream)
            batch.shape_gradient()
            ev[2 * i + 1].record(stream)
        torch.cuda.synchronize()
        sg_ms = float(np.mean([ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(args.steps)]))
        shape = {"samples": samples, "samples_per_point": args.batch_grad_walks,
                 "graph_replay_ms_with_shape_gradient": graph2_ms, "shape_gradient_launches_ms": sg_ms,
                 "samples_per_s": samples / (sg_ms * 1e-3),
                 "primal_walks_plus_samples_per_s": (total + samples) / (graph2_ms * 1e-3)}
        log(f"batched shape gradient: {samples} samples in {sg_ms:.1f} ms; primal + adjoint + shape gradient "
            f"{graph2_ms:.1f} ms graph")

    rf = None
    if not args.no_counts:
        flops, cnt = 0.0, {"box_tests": 0, "prim_tests": 0, "queries": 0, "steps": 0}
        for it, itn in zip(items, items_np):
            src = "source_params" in itn
            c_ = it["wos"].walk_counts(it["points"], it["scene"], seed=it["seed"], basis=src)
            K = itn["source_params"].shape[0] if src else 0
            flops += walk_flops(c_, itn["dim"], c_["walks"], K=K, src=src)
            for k in cnt:
                cnt[k] += c_[k]
        log(f"counts: {cnt['box_tests']:.3e} bound tests, {cnt['prim_tests']:.3e} primitive tests")
        rf = roofline(flops, run_ms, "k_batch_walk<2> + k_batch_walk<3> (one wos_batch_run)", None,
                      {"per_query": {"box_tests": cnt["box_tests"] / cnt["queries"],
                                     "prim_tests": cnt["prim_tests"] / cnt["queries"]},
                       "mean_steps": cnt["steps"] / total})

    plain = None
    if not args.no_plain:
        # reference data flow per item (one lane per walk, walks re-run in backward) on the same strided
        # points of every item, with the reference's brute-force scan and with this library's LBVH, and
        # the batched means on those points
        tp, tpb, pw, pwb, e_plain, e_bat, ns = 0.0, 0.0, 0, 0, 0.0, 0.0, 0
        t_full = 0.0   # the LBVH plain path's time for every item's full point set, from its per-item rates
        for j, (it, o, itn) in enumerate(zip(items, batch.outputs, items_np)):
            P = it["points"].shape[0]
            W = it["wos"].nwalks
            for bvh in (False, True):
                # the brute-force scan on a few points per item; the LBVH on enough that launch costs amortise
                stride = max(1, P // (args.plain_lbvh_points if bvh else args.plain_points))
                sel = slice(None, None, stride)
                psub = it["points"][sel].contiguous()
                tsub = targets[j][sel].contiguous()
                if itn["dim"] == 2:
                    src = "source_params" in itn
                    sc_p = wos_ext.Scene(t(itn["vertices"]), t(itn["indices"]), t(itn["values"]), use_bvh=bvh,
                                         source_type=3 if src else 0,
                                         source_params=t(itn["source_params"]) if src else None)
                else:
                    sc_p = wos_ext.Scene3D(t(itn["vertices"]), t(itn["indices"]), t(itn["values"]), use_bvh=bvh)
                if bvh:   # warm (workspace growth), then time
                    plain_step(torch, sc_p, it["wos"], psub, tsub, itn["seed"], W)
                torch.cuda.synchronize()
                ta = time.perf_counter()
                u, _ = plain_step(torch, sc_p, it["wos"], psub, tsub, itn["seed"], W)
                torch.cuda.synchronize()
                if bvh:
                    dt = time.perf_counter() - ta
                    tpb += dt
                    pwb += psub.shape[0] * W
                    t_full += dt * P / psub.shape[0]
                    ex = itn["exact"][::stride]
                    e_plain += float(((u.double().cpu().numpy() - ex) ** 2).sum())
                    e_bat += float(((o["mean64"].cpu().numpy()[::stride] - ex) ** 2).sum())
                    ns += ex.size
                else:
                    tp += time.perf_counter() - ta
                    pw += psub.shape[0] * W
            if j % 16 == 15:
                log(f"plain: {j + 1} of {len(items)} items")
        batched_rate = total / (graph_ms * 1e-3)
        plain = {"walks_per_s": pw / tp, "seconds": tp, "walks_per_s_lbvh": total / t_full,
                 "walks_per_s_lbvh_sample": pwb / tpb, "seconds_lbvh": tpb,
                 "sample": f"each item's points evenly strided, about {args.plain_points} per item for the "
                           f"brute-force scan (walks_per_s, {pw} walks) and {args.plain_lbvh_points} for the LBVH "
                           f"({pwb} walks, RMSE on these {ns} points; walks_per_s_lbvh is the whole workload's "
                           f"walks over the sum of each item's sample time scaled to its full point count); one "
                           f"lane per walk, walks re-run in backward",
                 "rmse_plain": float(np.sqrt(e_plain / ns)), "rmse_batched_same_points": float(np.sqrt(e_bat / ns)),
                 "speedup_batched_vs_plain": batched_rate / (pw / tp),
                 "speedup_batched_vs_plain_lbvh": batched_rate * t_full / total}
        log(f"plain: {plain['walks_per_s']:.3e} walks/s (LBVH {plain['walks_per_s_lbvh']:.3e}); speedup "
            f"{plain['speedup_batched_vs_plain']:.1f}x ({plain['speedup_batched_vs_plain_lbvh']:.1f}x vs LBVH); rmse "
            f"{plain['rmse_plain']:.4f} vs batched {plain['rmse_batched_same_points']:.4f}")

    cpu = None
    if not args.no_cpu:
        import oracle as O
        refs = []
        for itn in items_np:
            src = "source_params" in itn
            refs.append(O.Scene(itn["vertices"], itn["indices"], itn["values"], 3 if src else 0,
                                itn.get("source_params"), use_bvh=True))
        # every item's points thinned by the same stride (so the mix of items follows the workload's), the
        # stride halved until about cpu_seconds of work
        tc, done, stride = 0.0, 0, 1 << 14
        while tc < args.cpu_seconds and stride >= 1:
            for ref, itn, tg in zip(refs, items_np, targets):
                q = np.ascontiguousarray(itn["points"][stride // 2::stride])
                if q.shape[0] == 0:
                    continue
                W = itn["nwalks"]
                gt = tg.cpu().numpy()[stride // 2::stride].astype(np.float64)
                ta = time.perf_counter()
                r = ref.walk_reduce(q, W, itn["nsteps"], itn["epsilon"], seed=0, threads=CPU_THREADS, records=True)
                ref.adjoint_values(q.shape[0], W, r["rec_prim"], r["rec_t"], (2.0 / q.shape[0]) * (r["mean"] - gt))
                tc += time.perf_counter() - ta
                done += q.shape[0] * W
            stride //= 2
        cpu = {"value": done / tc, "unit": "walks/s", "cores": CPU_THREADS, "kind": "port",
               "sample": f"every item's points at strides 16384 down to {2 * stride} ({done} walks), primal + value "
                         f"adjoint, oracle with its CPU LBVH, {tc:.1f} s"}
        log(f"cpu: {cpu['value']:.3e} walks/s on {CPU_THREADS} threads")
    emit({"metric": "primal+grad walks/s at 1 GPU, batched vs plain (64 uneven scenes)",
          "value": total / graph_ms * 1e3, "unit": "walks/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
          "ms_per_step": graph_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
          "data": "synthetic",
          "config": {"workload": f"64-scene-batch-primal-adjoint-scale-{args.batch_scale}", "walks": total,
                     "scenes": len(items)},
          "batched_eager_ms": eager_ms, "graph_replay_ms": graph_ms, "walk_launches_ms": run_ms,
          "rmse_batched": rmse_batched, "tuning": tuning, "roofline": rf, "cpu_baseline": cpu, "plain": plain,
          "shape_gradient": shape})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--batch-grad-walks", type=int, default=4,
                    help="batch workload: OursCUDA / Ours3D samples per point of every item's shape gradient "
                         "(0: none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-fd", action="store_true", help="skip the finite-difference checks")
    ap.add_argument("--no-variants", action="store_true", help="clover: skip the importance-sampled variant")
    ap.add_argument("--no-counts", action="store_true",
                    help="skip the instrumented counting launch (keeps profiler summaries to the timed kernels)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 1 GiB device-to-device copy (3 times), the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--workload", choices=["clover", "circle", "sphere", "shape2d", "shape3d", "batch"],
                    default="clover",
                    help="clover: the headline primal+adjoint workload; circle / sphere: primal-only "
                         "Laplace workloads (1024-segment circle, 81,920-triangle displaced icosphere); "
                         "shape2d / shape3d: shape gradients (clover, 200k-triangle sphere); "
                         "batch: 64 uneven scenes, primal + adjoint")
    ap.add_argument("--autotune", action="store_true",
                    help="batch workload: choose the launch tuning first (cached per device)")
    ap.add_argument("--batch-scale", type=float, default=1.0,
                    help="query points per batched scene, as a fraction of the full workload's")
    ap.add_argument("--plain-points", type=int, default=32,
                    help="batch workload: points per item walked by the plain path with the brute-force scan")
    ap.add_argument("--plain-lbvh-points", type=int, default=8192,
                    help="batch workload: points per item walked by the plain path with the LBVH")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    import wos_ext
    from wos import configs as C

    torch.cuda.set_device(0)
    if args.calibrate_copy:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": a.numel() * 4, "bytes_written": b.numel() * 4}))
        return
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
    if args.workload in ("shape2d", "shape3d"):
        return shape_workload(args, torch, wos_ext, C)
    if args.workload == "batch":
        return batch_workload(args, torch, wos_ext, C)
    if args.workload in ("circle", "sphere"):
        return primal_workload(args, torch, wos_ext, C)
    return clover_workload(args, torch, wos_ext, C)


if __name__ == "__main__":
    main()
