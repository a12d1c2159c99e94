"""GPU parity of the paths the other files leave open: the reference's own
source types (Gaussian, sinusoid) and importance-sampled interior terms per
walk against the oracle, the 3D single walk per walk, WoS.solve, Dirichlet
interpolation bit for bit, the 3D walk against the exact harmonic solution,
the re-walked source-parameter adjoint against central differences, the
workspace error path, and the scene's read-write fields."""
import numpy as np
import pytest
import torch

import oracle as O
from wos import configs as C

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


SOURCES = {
    1: np.array([0.2, -0.1, 0.3, 0.25, 2.0], np.float32),    # Gaussian: mu_x, mu_y, sigma_x, sigma_y, A
    2: np.array([2.0, -1.0, 0.3, 1.5, 1.0], np.float32),     # sinusoid: A, B, C, amplitude, offset
    3: C.fourier_params(),                                    # Fourier: (ax, ay, phase, coeff) per term
}


def _source_scene(stype, n=512):
    import wos_ext
    v, ix = C.circle_polyline(n)
    vals = C.g_saddle(v).astype(np.float32)
    sp = SOURCES[stype]
    sc = wos_ext.Scene(_t(v), _t(ix), _t(vals), use_bvh=True, source_type=stype, source_params=_t(sp))
    ref = O.Scene(v, ix, vals, stype, sp)
    return sc, ref


def _disk_points(n, r=0.85, seed=0):
    rng = np.random.default_rng(seed)
    rr, t = r * np.sqrt(rng.random(n)), 2 * np.pi * rng.random(n)
    return np.stack([rr * np.cos(t), rr * np.sin(t)], 1).astype(np.float32)


@pytest.mark.parametrize("stype,use_is", [(1, False), (1, True), (2, False), (2, True), (3, True)])
def test_source_walks_match_oracle_per_walk(stype, use_is):
    """walk_reduce with the reference's source types and both interior
    samplers: each walk's value against the oracle's (same key, same
    arithmetic), and every point within 3 sigma."""
    import wos_ext
    sc, ref = _source_scene(stype)
    p = _disk_points(400)
    W = 32
    w = wos_ext.WoS(nwalks=W, nsteps=256, epsilon=1e-3, use_IS_for_greens=use_is)
    out = w.walk_reduce(_t(p), sc, seed=7, stats=True)
    r = ref.walk_reduce(p, W, 256, 1e-3, seed=7, walk_values=True, use_is=use_is)
    gv = out["walk_value"].cpu().numpy()
    same = np.abs(gv - r["walk_value"]) <= 1e-5 * (1 + np.abs(r["walk_value"]))
    print(f"source {stype} is={use_is}: {same.mean():.5f} of {gv.size} walks agree with the oracle")
    assert same.all()
    m, v = out["mean64"].cpu().numpy(), out["var64"].cpu().numpy()
    se = np.sqrt((v + r["var"]) / W) + 1e-9
    assert np.mean(np.abs(m - r["mean"]) / se > 3) < 0.01


def test_importance_sampling_lowers_the_interior_variance():
    """Same scene and walks: sampling the interior term by the Green's
    function's density gives a smaller per-walk variance than uniform
    sampling, and both means agree within their errors."""
    import wos_ext
    sc, _ = _source_scene(2)
    p = _disk_points(2000, seed=3)
    W = 64
    o = [wos_ext.WoS(nwalks=W, nsteps=256, epsilon=1e-3, use_IS_for_greens=f).walk_reduce(_t(p), sc, seed=1,
                                                                                           stats=True)
         for f in (False, True)]
    mu, vu = o[0]["mean64"].cpu().numpy(), o[0]["var64"].cpu().numpy()
    mi, vi = o[1]["mean64"].cpu().numpy(), o[1]["var64"].cpu().numpy()
    assert vi.mean() < vu.mean()
    z = np.abs(mu - mi) / np.sqrt((vu + vi) / W + 1e-12)
    assert np.mean(z > 3) < 0.01


@pytest.mark.parametrize("use_is", [False, True])
def test_sinusoid_source_rmse_against_manufactured_solution(use_is):
    """-Lap u = a sin(Ax+By+C) + o with g = the manufactured solution on the
    boundary: the GPU means are within the Monte Carlo bound of it."""
    import wos_ext
    A, B, Cc, a, o = SOURCES[2].astype(np.float64)
    u = lambda q: (a * np.sin(A * q[:, 0] + B * q[:, 1] + Cc) / (A * A + B * B)
                   - o * (q[:, 0] ** 2 + q[:, 1] ** 2) / 4 + q[:, 0] ** 2 - q[:, 1] ** 2)
    v, ix = C.circle_polyline(1024)
    sc = wos_ext.Scene(_t(v), _t(ix), _t(u(v.astype(np.float64)).astype(np.float32)), use_bvh=True,
                       source_type=2, source_params=_t(SOURCES[2]))
    p = _disk_points(20000, r=0.9, seed=5)
    W = 128
    out = wos_ext.WoS(nwalks=W, nsteps=256, epsilon=1e-4, use_IS_for_greens=use_is).walk_reduce(
        _t(p), sc, seed=2, stats=True)
    m, var = out["mean64"].cpu().numpy(), out["var64"].cpu().numpy()
    rmse = np.sqrt(np.mean((m - u(p)) ** 2))
    bound = np.sqrt(np.mean(var / W))
    print(f"sinusoid source, is={use_is}: rmse {rmse:.5f}, Monte Carlo bound {bound:.5f}")
    assert rmse < 1.1 * bound + 5e-4


def test_single_walk_3d_matches_oracle_per_walk():
    import wos_ext
    from wos.fwd import PCG32
    v, f, noise = C.displaced_sphere(3)
    vals = C.g_harmonic3(v).astype(np.float32)
    sc = wos_ext.Scene3D(_t(v), _t(f), _t(vals), use_bvh=True)
    ref = O.Scene(v, f, vals)
    p = np.repeat(C.points_in_star(1500, noise, seed=4), 2, axis=0)
    n = p.shape[0]
    smp = PCG32(n, 777, 99)
    st0, inc0 = smp.state.cpu().numpy().view(np.uint64).copy(), smp.inc.cpu().numpy().view(np.uint64).copy()
    rec = wos_ext.WoS3D(nwalks=1, nsteps=256, epsilon=1e-3).single_walk(_t(p), sc, smp)
    r = ref.single_walk(p, st0, inc0, nsteps=256, epsilon=1e-3)
    same = rec.prim_id.cpu().numpy() == r["prim"]
    print(f"3D single walks: {same.mean():.5f} of {n} exit on the oracle's triangle")
    assert same.all()
    np.testing.assert_array_equal(smp.state.cpu().numpy().view(np.uint64)[same], r["state"][same])
    np.testing.assert_allclose(rec.p.cpu().numpy()[same], r["p"][same], atol=1e-5)
    np.testing.assert_allclose(rec.uv.cpu().numpy()[same], r["t"][same], atol=1e-5)
    np.testing.assert_array_equal(rec.valid.cpu().numpy(), r["valid"].astype(bool))


@pytest.mark.parametrize("dim", [2, 3])
def test_solve_is_the_walk_reduce_mean(dim):
    """WoS.solve (include/wos.cuh:320-364) runs all nwalks walks per point:
    the mean equals walk_reduce's and the oracle's within 3 sigma."""
    import wos_ext
    if dim == 2:
        c = C.config1()
        sc = wos_ext.Scene(_t(c["vertices"]), _t(c["indices"]), _t(c["values"]), use_bvh=True)
        w = wos_ext.WoS(nwalks=150, nsteps=256, epsilon=1e-3)
        p, ref = c["points"][::53], O.Scene(c["vertices"], c["indices"], c["values"])
    else:
        v, f, noise = C.displaced_sphere(2)
        vals = C.g_harmonic3(v).astype(np.float32)
        sc = wos_ext.Scene3D(_t(v), _t(f), _t(vals), use_bvh=True)
        w = wos_ext.WoS3D(nwalks=150, nsteps=256, epsilon=1e-3)
        p, ref = C.points_in_star(200, noise, seed=8), O.Scene(v, f, vals)
    u = w.solve(_t(p), sc, seed=3)
    o = w.walk_reduce(_t(p), sc, seed=3, stats=True)
    assert torch.equal(u, o["mean"])
    r = ref.walk_reduce(p, 150, 256, 1e-3, seed=3)
    m, var = o["mean64"].cpu().numpy(), o["var64"].cpu().numpy()
    z = np.abs(m - r["mean"]) / np.sqrt((var + r["var"]) / 150 + 1e-12)
    assert np.mean(z > 3) < 0.02


@pytest.mark.parametrize("dim", [2, 3])
def test_dirichlet_bit_exact(dim):
    """Scene.dirichlet (wos_dirichlet) equals the oracle's interpolation bit
    for bit on random records, invalid and out-of-range ones included."""
    import wos_ext
    rng = np.random.default_rng(dim)
    if dim == 2:
        v, f = C.star_polyline(777)
        vals = rng.normal(size=v.shape[0]).astype(np.float32)
        sc = wos_ext.Scene(_t(v), _t(f), _t(vals))
        n = 50000
        t = rng.random(n).astype(np.float32)
        rec = wos_ext.ClosestPointRecord
    else:
        v, f, _ = C.displaced_sphere(2)
        vals = rng.normal(size=v.shape[0]).astype(np.float32)
        sc = wos_ext.Scene3D(_t(v), _t(f), _t(vals))
        n = 50000
        uv = rng.random((n, 2)).astype(np.float32)
        t = np.where(uv.sum(1, keepdims=True) > 1, 1 - uv, uv).astype(np.float32)
        rec = wos_ext.ClosestPointRecord3D
    prim = rng.integers(-2, f.shape[0] + 3, size=n).astype(np.int32)
    valid = rng.random(n) < 0.9
    r = rec(valid=_t(valid), prim_id=_t(prim), **({"t": _t(t)} if dim == 2 else {"uv": _t(t)}))
    got = sc.dirichlet(r).cpu().numpy()
    want = O.Scene(v, f, vals).dirichlet(prim, t, valid.astype(np.int32))
    np.testing.assert_array_equal(got, want)


def test_walk_3d_rmse_against_exact_harmonic():
    """3D Laplace on the displaced sphere with g = x^2 - y^2 + z (harmonic):
    RMSE of the per-point means against the exact solution at the Monte
    Carlo bound; the oracle agrees per point on a subset."""
    import wos_ext
    v, f, noise = C.displaced_sphere(4)
    vals = C.g_harmonic3(v).astype(np.float32)
    sc = wos_ext.Scene3D(_t(v), _t(f), _t(vals), use_bvh=True)
    p = C.points_in_star(40000, noise, seed=6)
    W = 32
    out = wos_ext.WoS3D(nwalks=W, nsteps=256, epsilon=1e-3).walk_reduce(_t(p), sc, seed=0, stats=True)
    m, var = out["mean64"].cpu().numpy(), out["var64"].cpu().numpy()
    rmse = np.sqrt(np.mean((m - C.g_harmonic3(p)) ** 2))
    bound = np.sqrt(np.mean(var / W))
    print(f"3D sphere (5,120 triangles): rmse {rmse:.5f}, Monte Carlo bound {bound:.5f}")
    assert rmse < 1.1 * bound + 2e-3


@pytest.mark.parametrize("stype,use_is", [(1, False), (2, True), (3, False)])
def test_source_adjoint_matches_central_differences(stype, use_is):
    """dL/d theta for every source parameter, re-walked from the keys,
    against central differences of the FP64 means under the same seed (the
    paths do not depend on the source, so the estimate is smooth in it)."""
    import wos_ext
    sc, _ = _source_scene(stype, n=256)
    p = _disk_points(600, seed=2)
    W = 32
    w = wos_ext.WoS(nwalks=W, nsteps=256, epsilon=1e-3, use_IS_for_greens=use_is)
    g = np.random.default_rng(1).normal(size=p.shape[0])
    gs = wos_ext.adjoint_source(sc, w, _t(p), 5, _t(g.astype(np.float32))).cpu().numpy().reshape(-1)
    sp0 = SOURCES[stype].reshape(-1).copy()
    idx = range(sp0.size) if stype != 3 else [0, 1, 2, 3, 6, 7, 11]
    h = 1e-3
    for j in idx:
        fd = 0.0
        for sgn in (1, -1):
            sp = sp0.copy()
            sp[j] += sgn * h
            sc.set_source(stype, _t(sp.reshape(SOURCES[stype].shape)))
            m = w.walk_reduce(_t(p), sc, seed=5, stats=True)["mean64"].cpu().numpy()
            fd += sgn * float(g @ m) / (2 * h)
        assert abs(gs[j] - fd) <= 3e-3 * abs(fd) + 2e-4 * np.abs(gs).max() + 1e-6, (j, gs[j], fd)
    sc.set_source(stype, _t(SOURCES[stype]))



def test_source_adjoint_is_bit_reproducible():
    """The source-parameter gradient sums fixed slices of walks in a fixed
    order: repeated calls, and calls under another launch shape, give the
    same bits."""
    import wos_ext
    sc, _ = _source_scene(3, n=256)
    p = _t(_disk_points(3000, seed=4))
    w = wos_ext.WoS(nwalks=32, nsteps=256, epsilon=1e-3)
    g = _t(np.random.default_rng(2).normal(size=3000).astype(np.float32))
    a = wos_ext.adjoint_source(sc, w, p, 9, g)
    b = wos_ext.adjoint_source(sc, w, p, 9, g)
    old = wos_ext.get_tuning()
    try:
        wos_ext.set_tuning(wgs_per_cu=3)
        c = wos_ext.adjoint_source(sc, w, p, 9, g)
    finally:
        wos_ext.set_tuning(chunk=old[0], wgs_per_cu=old[1], top_levels=old[2], window=old[3])
    assert torch.equal(a, b)
    assert torch.equal(a, c)

def test_workspace_failure_is_an_error_not_a_launch():
    """A traversal-spill request that fails makes the entry point return an
    error (WoSError) before any launch; the next call works."""
    import wos_ext
    v, f, noise = C.displaced_sphere(4)   # deep enough to need the spill
    sc = wos_ext.Scene3D(_t(v), _t(f), _t(C.g_harmonic3(v).astype(np.float32)), use_bvh=True)
    assert sc.max_depth + 2 > 8
    p = _t(C.points_in_star(512, noise))
    w = wos_ext.WoS3D(nwalks=4, nsteps=64, epsilon=1e-3)
    for call in (lambda: w.walk_reduce(p, sc, seed=1), lambda: sc.closest_point_preliminary(p),
                 lambda: sc.sdf_native(p)):
        wos_ext.debug_fail_workspace(1)
        with pytest.raises(wos_ext.WoSError, match="injected"):
            call()
    wos_ext.debug_fail_workspace(0)
    a = w.walk_reduce(p, sc, seed=1)["mean"]
    torch.cuda.synchronize()
    assert torch.isfinite(a).all()


def test_scene_fields_follow_assignment():
    """Polyline.values / .source_params are the native scene's read-write
    fields (src/python/scene.cu:35-47): assigning them changes the walks
    exactly as building a new scene with them; vertices cannot be reassigned."""
    import wos_ext
    from wos.scene import Polyline
    c = C.config2(res=40)
    sp2 = C.fourier_params(np.array([0.2, -0.4, 0.1, 0, 0, 0.3, 0, 0]))
    pts = _t(c["points"])
    cw = wos_ext.WoS(nwalks=16, nsteps=256, epsilon=1e-4)
    sc = Polyline(_t(c["vertices"]), _t(c["indices"]), _t(c["values"]), source_type=3,
                  source_params=_t(c["source_params"]))
    g2 = _t((c["values"] * 0.5 - 0.3).astype(np.float32))
    sc.values = g2
    sc.source_params = _t(sp2)
    fresh = Polyline(_t(c["vertices"]), _t(c["indices"]), g2, source_type=3, source_params=_t(sp2))
    a = cw.walk_reduce(pts, sc, seed=2, records=True)
    b = cw.walk_reduce(pts, fresh, seed=2, records=True)
    for k in a:
        assert torch.equal(a[k], b[k]), k
    with pytest.raises(wos_ext.WoSError, match="vertices"):
        sc.vertices = _t(c["vertices"] * 1.1)
