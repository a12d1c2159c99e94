"""Benchmark: primal + gradient walks/s of the 2D Poisson clover workload.

Workload (one step): 126,820 query points on a 512x512 grid inside the clover
r = 1 + 0.3 cos 4t (4,096 segments), 128 walks per point (16.2M walks),
eps = 1e-4, 256 steps, source sin3x cos2y as 8 Fourier terms, Dirichlet data
g = x^2 - y^2.  A step solves u, forms L = mean((u - u_target)^2) against a
fixed target solved at 1.1 g, and back-propagates to the 4,096 vertex values
and 8 source coefficients.  Inputs are resident in HBM before timing.

Also reported: the same step through the reference-style path ("plain": one
lane per walk, brute-force closest point, per-walk records, walks re-run in
the backward pass), RMSE of both against the exact solution of the
manufactured twin, the roofline of the dominant kernel and the CPU baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-plain] [--no-cpu] [--no-counts]

The HBM traffic of the walk kernel comes from the FETCH_SIZE / WRITE_SIZE
counter runs summarised in profiles/walk_clover_hbm.md (rocprofv3 cannot run
inside this process); it is null when that file is absent.
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

# algorithmic flop model (FMA = 2): per AABB test, per segment test, per walk
# step (direction, move, bound), per source sample and Fourier term, per walk
C_BOX_2D, C_PRIM_2D = 11, 20
C_STEP, C_SRC, C_TERM, C_WALK = 30, 25, 34, 15


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "walk_clover_hbm.md")


def measured_traffic():
    """HBM bytes per walk launch recorded by the counter runs, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_inputs(torch, C):
    c = C.config2()
    dv = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dv)
    return c, dict(vertices=t(c["vertices"]), indices=t(c["indices"]), values=t(c["values"]),
                   points=t(c["points"]), source=t(c["source_params"]))


def fused_step(torch, wos_ext, scene, cwos, pts, target, seed):
    out = cwos.walk_reduce(pts, scene, seed=seed, records=True, basis=True)
    u = out["mean"]
    P = u.shape[0]
    g = (2.0 / P) * (u - target)
    loss = torch.mean((u - target) ** 2)
    gv, gc = wos_ext.adjoint_values(scene, P, cwos.nwalks, out["rec_prim"], out["rec_t"], g, basis=out["basis"],
                                    want_values=True, want_coeffs=True)
    return u, loss, gv, gc


def plain_step(torch, wos_ext, scene_bf, cwos, pts, target, seed, W):
    """Reference data flow: PCG32 per walk (Dr.Jit PCG32 + TEA), one lane per
    walk with the brute-force scan, records to HBM, dirichlet + scatter mean;
    backward re-runs the walks from the same seed and scatters g(1-t), g t."""
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
    gv.index_add_(0, f[:, 0], gw * (1 - rec.t)).index_add_(0, f[:, 1], gw * rec.t)
    return u, gv


def primal_workload(args, torch, wos_ext, C):
    """Primal-only Laplace workloads: walks/s, closest-point queries/s, RMSE
    against the exact harmonic solution, LBVH build time."""
    c = C.config1() if args.workload == "circle" else C.config3()
    dim = c["vertices"].shape[1]
    dv = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dv)
    V, F, G, pts = t(c["vertices"]), t(c["indices"]), t(c["values"]), t(c["points"])
    Scene, WoS = (wos_ext.Scene, wos_ext.WoS) if dim == 2 else (wos_ext.Scene3D, wos_ext.WoS3D)
    Scene(V, F, G, use_bvh=True)   # warm the build kernels
    torch.cuda.synchronize()
    tb = time.perf_counter()
    scene = Scene(V, F, G, use_bvh=True)
    build_ms = (time.perf_counter() - tb) * 1e3
    P, W = pts.shape[0], c["nwalks"]
    cw = WoS(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"])
    for i in range(args.warmup):
        cw.walk_reduce(pts, scene, seed=i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = cw.walk_reduce(pts, scene, seed=100 + i, stats=(i == args.steps - 1))
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    rmse = float(np.sqrt(np.mean((out["mean64"].cpu().numpy() - c["exact"]) ** 2)))
    # closest-point queries/s (LBVH traversal alone)
    scene.closest_point_preliminary(pts)
    torch.cuda.synchronize()
    tq = time.perf_counter()
    for _ in range(3):
        scene.closest_point_preliminary(pts)
    torch.cuda.synchronize()
    qps = 3 * P / (time.perf_counter() - tq)
    cnt = None if args.no_counts else cw.walk_counts(pts, scene, seed=100, basis=False)
    log(f"{args.workload}: {ms:.2f} ms/solve, {P * W / ms * 1e3:.3e} walks/s, rmse {rmse:.4f}, "
        f"build {build_ms:.2f} ms, {qps:.3e} queries/s, counts {cnt}")
    print(json.dumps({"metric": "walks/s (primal)", "value": P * W / ms * 1e3, "unit": "walks/s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                      "config": {"workload": args.workload, "points": P, "walks_per_point": W,
                                 "primitives": int(c["indices"].shape[0])},
                      "rmse_exact": rmse, "lbvh_build_ms": build_ms, "closest_point_queries_per_s": qps,
                      "counts": cnt}))


def shape2d_workload(args, torch, wos_ext, C):
    """2D shape gradient on the clover (4,096 segments, 126,820 points, 16
    walks, Laplace with g = x^2 - y^2): forward walk_reduce, loss L = mean
    (u - u*)^2 against the solution u* for 1.1 g, OursCUDA backward onto all
    vertex positions.  One "primal+grad walk" is a forward walk plus a
    gradient sample (primary walk, LIB, two secondary walks)."""
    c = C.config2()
    dv = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dv)
    V, F, G, pts = t(c["vertices"]), t(c["indices"]), t(c["values"]), t(c["points"])
    P, W = pts.shape[0], 16
    scene = wos_ext.Scene(V, F, G, use_bvh=True)
    cw = wos_ext.WoS(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"])
    kw = dict(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"], epsilon2=c["epsilon"], clamping=0.1)
    target = cw.walk_reduce(pts, wos_ext.Scene(V, F, G * 1.1, use_bvh=True), seed=12345)["mean"]
    torch.cuda.synchronize()

    def step(seed):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        ev[0].record()
        u = cw.walk_reduce(pts, scene, seed=seed)["mean"]
        g = (2.0 / P) * (u - target)
        ev[1].record()
        gv, _ = wos_ext.shape_gradient(scene, pts, seed, g, **kw)
        ev[2].record()
        return ev, gv

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs = [step(100 + i)[0] for i in range(args.steps)]
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    fwd = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    bwd = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    log(f"shape2d: {ms:.2f} ms/step (forward {fwd:.2f}, gradient {bwd:.2f}), {P * W / ms * 1e3:.3e} walks/s")
    print(json.dumps({"metric": "primal+grad walks/s (2D shape gradient)", "value": P * W / ms * 1e3,
                      "unit": "walks/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                      "data": "synthetic",
                      "config": {"workload": "2d-shape-gradient-clover-4096seg-126820pts-16walks", "points": P,
                                 "walks_per_point": W, "segments": int(c["indices"].shape[0])},
                      "forward_ms": fwd, "gradient_ms": bwd}))


def shape3d_workload(args, torch, wos_ext, C):
    """3D shape gradient (200k-triangle decimated displaced sphere, 262,144
    points, 16 walks): forward walk_reduce, loss L = mean (u - g)^2, Ours3D
    backward onto all vertex positions.  One "primal+grad walk" is a forward
    walk plus a gradient sample (primary walk, LIB, two secondary walks)."""
    c = C.config4()
    dv = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dv)
    V, F, G, pts, target = t(c["vertices"]), t(c["indices"]), t(c["values"]), t(c["points"]), t(
        c["target"].astype(np.float32))
    P, W = pts.shape[0], c["nwalks"]
    torch.cuda.synchronize()
    tb = time.perf_counter()
    scene = wos_ext.Scene3D(V, F, G, use_bvh=True)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - tb) * 1e3
    cw = wos_ext.WoS3D(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"])
    kw = dict(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"], epsilon2=c["epsilon2"], clamping=c["clamping"])

    def step(seed):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        ev[0].record()
        u = cw.walk_reduce(pts, scene, seed=seed)["mean"]
        g = (2.0 / P) * (u - target)
        ev[1].record()
        gv, _ = wos_ext.shape_gradient(scene, pts, seed, g, **kw)
        ev[2].record()
        return ev, gv

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs = [step(100 + i)[0] for i in range(args.steps)]
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    fwd = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    bwd = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    log(f"shape3d: {ms:.2f} ms/step (forward {fwd:.2f}, gradient {bwd:.2f}), {P * W / ms * 1e3:.3e} walks/s")
    print(json.dumps({"metric": "primal+grad walks/s (3D shape gradient)", "value": P * W / ms * 1e3,
                      "unit": "walks/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                      "data": "synthetic",
                      "config": {"workload": "3d-shape-gradient-200k-tris-262144pts-16walks", "points": P,
                                 "walks_per_point": W, "triangles": int(c["indices"].shape[0])},
                      "forward_ms": fwd, "gradient_ms": bwd, "scene_build_ms": build_ms}))


def batch_workload(args, torch, wos_ext, C):
    """64 uneven scenes (32 2D, 32 3D; sizes log-uniform; points scaled by
    --batch-scale): one batched launch per dimension, eager and replayed from
    a captured graph, against the same items walked one call at a time."""
    items_np = C.config5(scale=args.batch_scale)
    dv = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dv)
    items, total = [], 0
    for it in items_np:
        if it["dim"] == 2:
            sc = wos_ext.Scene(t(it["vertices"]), t(it["indices"]), t(it["values"]), use_bvh=True,
                               source_type=it.get("source_type", 0),
                               source_params=t(it["source_params"]) if "source_params" in it else None)
            cw = wos_ext.WoS(nwalks=it["nwalks"], nsteps=it["nsteps"], epsilon=it["epsilon"])
        else:
            sc = wos_ext.Scene3D(t(it["vertices"]), t(it["indices"]), t(it["values"]), use_bvh=True)
            cw = wos_ext.WoS3D(nwalks=it["nwalks"], nsteps=it["nsteps"], epsilon=it["epsilon"])
        items.append(dict(scene=sc, wos=cw, points=t(it["points"]), seed=it["seed"], records=True,
                          basis="source_params" in it))
        total += it["points"].shape[0] * it["nwalks"]
    batch = wos_ext.Batch(items)
    tuning = batch.autotune() if args.autotune else None
    if tuning:
        log(f"autotuned: {tuning}")

    def sequential():
        for it in items:
            it["wos"].walk_reduce(it["points"], it["scene"], seed=it["seed"], records=True, basis=it["basis"])

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps

    seq_ms = timed(sequential)
    eager_ms = timed(batch.run)
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    batch.run()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        batch.run()
    graph_ms = timed(g.replay)
    log(f"batch: {total} walks; sequential {seq_ms:.2f} ms, batched {eager_ms:.2f} ms, graph {graph_ms:.2f} ms")
    print(json.dumps({"metric": "walks/s (64 batched scenes)", "value": total / graph_ms * 1e3, "unit": "walks/s",
                      "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": graph_ms,
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                      "data": "synthetic",
                      "config": {"workload": f"64-scene-batch-scale-{args.batch_scale}", "walks": total},
                      "sequential_ms": seq_ms, "batched_eager_ms": eager_ms, "graph_replay_ms": graph_ms,
                      "tuning": tuning,
                      "sequential_walks_per_s": total / seq_ms * 1e3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
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
                         "batch: 64 uneven scenes")
    ap.add_argument("--autotune", action="store_true",
                    help="batch workload: choose the launch tuning first (cached per device)")
    ap.add_argument("--batch-scale", type=float, default=0.1,
                    help="query points per batched scene, as a fraction of the full workload's")
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
    if args.workload == "shape2d":
        return shape2d_workload(args, torch, wos_ext, C)
    if args.workload == "shape3d":
        return shape3d_workload(args, torch, wos_ext, C)
    if args.workload == "batch":
        return batch_workload(args, torch, wos_ext, C)
    if args.workload != "clover":
        return primal_workload(args, torch, wos_ext, C)
    c, d = build_inputs(torch, C)
    P, W = c["points"].shape[0], c["nwalks"]
    log(f"workload: {P} points x {W} walks, 4096 segments, eps 1e-4")
    scene = wos_ext.Scene(d["vertices"], d["indices"], d["values"], use_bvh=True, source_type=3,
                          source_params=d["source"])
    cwos = wos_ext.WoS(nwalks=W, nsteps=c["nsteps"], epsilon=c["epsilon"])
    # fixed target: the solution at 1.1 g (solved once, untimed)
    tscene = wos_ext.Scene(d["vertices"], d["indices"], d["values"] * 1.1, use_bvh=True, source_type=3,
                           source_params=d["source"])
    target = cwos.walk_reduce(d["points"], tscene, seed=12345)["mean"]
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    for i in range(args.warmup):
        fused_step(torch, wos_ext, scene, cwos, d["points"], target, seed=i)
    torch.cuda.synchronize()
    log("warmup done")

    # timed region: K full steps
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kernel_ms = []
    for i in range(args.steps):
        # HIP events on the stream the walk kernel runs on, around the walk launch
        ev[2 * i].record(stream)
        out = cwos.walk_reduce(d["points"], scene, seed=100 + i, records=True, basis=True)
        ev[2 * i + 1].record(stream)
        u = out["mean"]
        g = (2.0 / P) * (u - target)
        gv, gc = wos_ext.adjoint_values(scene, P, W, out["rec_prim"], out["rec_t"], g, basis=out["basis"],
                                        want_values=True, want_coeffs=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(args.steps):
        kernel_ms.append(ev[2 * i].elapsed_time(ev[2 * i + 1]))
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    walks = P * W
    value = walks / (ms_per_step * 1e-3)
    log(f"fused: {ms_per_step:.2f} ms/step, {value:.3e} walks/s; walk launch {np.mean(kernel_ms):.2f} ms")

    # ---- algorithmic work of one walk launch (untimed counting run) ----------
    counts = None if args.no_counts else cwos.walk_counts(d["points"], scene, seed=100)
    roofline = None
    if counts is not None:
        K = c["source_params"].shape[0]
        flops = (counts["box_tests"] * C_BOX_2D + counts["prim_tests"] * C_PRIM_2D
                 + counts["steps"] * (C_STEP + C_SRC + C_TERM * K) + walks * C_WALK)
        launch_s = np.mean(kernel_ms) * 1e-3
        achieved = flops / launch_s / 1e12
        roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32 / 1e12,
                    "unit": "TFLOP/s", "frac": round(achieved * 1e12 / PEAK_FP32, 4), "traffic": measured_traffic(),
                    "kernel": "k_walk_persistent<2,true,8,false>", "flops_per_launch": float(flops),
                    "launch_ms": round(np.mean(kernel_ms), 3),
                    "per_query": {"box_tests": counts["box_tests"] / counts["queries"],
                                  "prim_tests": counts["prim_tests"] / counts["queries"]},
                    "mean_steps": counts["steps"] / walks}
        if roofline["traffic"]:
            roofline["hbm_GBs"] = round(roofline["traffic"] / launch_s / 1e9, 1)
            roofline["hbm_frac"] = round(roofline["traffic"] / launch_s / PEAK_HBM, 4)

    # ---- accuracy: manufactured twin, RMSE vs exact ---------------------------
    ct = C.config2(twin=True)
    tw = wos_ext.Scene(d["vertices"], d["indices"], torch.from_numpy(ct["values"]).cuda(), use_bvh=True,
                       source_type=3, source_params=d["source"])
    o = cwos.walk_reduce(d["points"], tw, seed=7, stats=True)
    rmse = float(np.sqrt(np.mean((o["mean64"].cpu().numpy() - ct["exact"]) ** 2)))

    # ---- reference-style path ------------------------------------------------
    plain = None
    if not args.no_plain:
        scene_bf = wos_ext.Scene(d["vertices"], d["indices"], d["values"], use_bvh=False, source_type=3,
                                 source_params=d["source"])
        sub = d["points"][:: 8].contiguous()   # 1/8 of the grid keeps the brute-force run short
        tsub = target[:: 8].contiguous()
        plain_step(torch, wos_ext, scene_bf, cwos, sub, tsub, 0, W)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        plain_step(torch, wos_ext, scene_bf, cwos, sub, tsub, 1, W)
        torch.cuda.synchronize()
        tp = time.perf_counter() - ta
        plain = {"walks_per_s": sub.shape[0] * W / tp, "ms_per_step_sample": tp * 1e3,
                 "sample": f"{sub.shape[0]} points (every 8th) x {W} walks, brute-force scan, walks re-run in backward",
                 "speedup_fused_vs_plain": value / (sub.shape[0] * W / tp)}
        log(f"plain: {plain['walks_per_s']:.3e} walks/s")

    # ---- CPU baseline: the oracle on a bounded sample ------------------------
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        ref = O.Scene(c["vertices"], c["indices"], c["values"], 3, c["source_params"])
        threads = min(os.cpu_count() or 1, 16)
        done, tc, blk = 0, 0.0, 64
        pts_np = c["points"]
        while tc < args.cpu_seconds and done < P:
            q = pts_np[done: done + blk]
            ta = time.perf_counter()
            ref.walk_reduce(q, W, c["nsteps"], c["epsilon"], seed=0, threads=threads)
            tc += time.perf_counter() - ta
            done += q.shape[0]
        cpu = {"value": done * W / tc, "unit": "walks/s", "cores": threads, "kind": "port",
               "sample": f"first {done} grid points x {W} walks, primal only, brute-force scan, {tc:.1f} s"}
        log(f"cpu: {cpu['value']:.3e} walks/s on {threads} threads")

    line = {"metric": "primal+grad walks/s at 1 GPU (2D Poisson clover, values + source-coefficient adjoint)",
            "value": value, "unit": "walks/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": "2d-poisson-clover-4096seg-512grid-128walks-eps1e-4-fourier8",
                       "points": P, "walks_per_point": W, "segments": 4096, "parallelism": "single-gpu"},
            "rmse_twin": rmse, "roofline": roofline, "cpu_baseline": cpu, "plain": plain}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
