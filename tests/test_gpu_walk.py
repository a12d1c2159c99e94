"""GPU parity of the walks: single walks from explicit PCG32 states against
the oracle (per walk), walk_reduce per point within 3 sigma of the oracle,
RMSE against the exact harmonic solutions, and the value / coefficient
adjoints against the oracle and central finite differences."""
import numpy as np
import pytest
import torch

import oracle as O
from wos import configs as C

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _z(m1, v1, m2, v2, W):
    se = np.sqrt((v1 + v2) / W) + 1e-9
    return np.abs(m1 - m2) / se


@pytest.mark.parametrize("source,use_is", [(False, False), (True, False), (True, True)])
def test_single_walk_matches_oracle_per_walk(source, use_is):
    import wos_ext
    from wos.fwd import PCG32
    c = C.config2(res=64, twin=True)
    kw = dict(source_type=3, source_params=_t(c["source_params"])) if source else {}
    sc = wos_ext.Scene(_t(c["vertices"]), _t(c["indices"]), _t(c["values"]), use_bvh=True, **kw)
    ref = O.Scene(c["vertices"], c["indices"], c["values"], source_type=3 if source else 0,
                  source_params=c["source_params"] if source else None)
    p = np.repeat(c["points"][::3], 4, axis=0)
    n = p.shape[0]
    smp = PCG32(n, 12345, 678)
    st0, inc0 = smp.state.cpu().numpy().view(np.uint64).copy(), smp.inc.cpu().numpy().view(np.uint64).copy()
    w = wos_ext.WoS(nwalks=1, nsteps=128, epsilon=1e-3, use_IS_for_greens=use_is)
    rec = w.single_walk(_t(p), sc, smp)
    r = ref.single_walk(p, st0, inc0, nsteps=128, epsilon=1e-3, use_is=use_is)
    same = (rec.prim_id.cpu().numpy() == r["prim"])
    # identical float expressions in the same order on both sides: every walk agrees
    print(f"per-walk exit agreement: {same.mean():.5f} of {n} walks")
    assert same.all()
    np.testing.assert_array_equal(smp.state.cpu().numpy().view(np.uint64)[same], r["state"][same])
    np.testing.assert_allclose(rec.p.cpu().numpy()[same], r["p"][same], atol=1e-5)
    if source:
        np.testing.assert_allclose(rec.contrib.cpu().numpy()[same], r["contrib"][same], rtol=1e-3, atol=1e-4)


def test_pcg32_torch_matches_oracle():
    from wos.fwd import PCG32
    s = PCG32(5, 42, 54)
    got = [s.next_uint32().cpu().numpy() for _ in range(4)]
    for k in range(5):
        ref = O.pcg32_stream(k, 42, 54, 4)
        np.testing.assert_array_equal(np.array([g[k] for g in got], dtype=np.uint32), ref)


def test_walk_reduce_config1_parity_and_rmse():
    import wos_ext
    c = C.config1()
    sc = wos_ext.Scene(_t(c["vertices"]), _t(c["indices"]), _t(c["values"]), use_bvh=True)
    w = wos_ext.WoS(nwalks=64, nsteps=256, epsilon=1e-3)
    out = w.walk_reduce(_t(c["points"]), sc, seed=0, stats=True)
    m, v = out["mean64"].cpu().numpy(), out["var64"].cpu().numpy()
    rmse = np.sqrt(np.mean((m - c["exact"]) ** 2))
    # RMSE of a 64-walk estimate of a function of size ~1: Monte Carlo bound
    assert rmse < 1.2 * np.sqrt(np.mean(v / 64)) + 1e-3, rmse
    # oracle on a subset (walk keys use the index within the array passed in)
    ps = np.ascontiguousarray(c["points"][::37])
    ref = O.Scene(c["vertices"], c["indices"], c["values"]).walk_reduce(ps, 64, 256, 1e-3, seed=0)
    o2 = w.walk_reduce(_t(ps), sc, seed=0, stats=True)
    m2, v2 = o2["mean64"].cpu().numpy(), o2["var64"].cpu().numpy()
    z = _z(m2, v2, ref["mean"], ref["var"], 64)
    assert np.mean(z > 3) < 0.01, np.mean(z > 3)
    # identical streams: most points agree to rounding
    assert np.mean(np.abs(m2 - ref["mean"]) < 1e-4) > 0.9


def test_walk_reduce_bvh_equals_bruteforce():
    import wos_ext
    c = C.config2(res=48)
    sc = wos_ext.Scene(_t(c["vertices"]), _t(c["indices"]), _t(c["values"]), use_bvh=True)
    w = wos_ext.WoS(nwalks=16, nsteps=256, epsilon=1e-4)
    a = w.walk_reduce(_t(c["points"]), sc, seed=3, records=True)
    sc.use_bvh = False
    b = w.walk_reduce(_t(c["points"]), sc, seed=3, records=True)
    np.testing.assert_array_equal(a["walk_value"].cpu().numpy(), b["walk_value"].cpu().numpy())
    np.testing.assert_array_equal(a["rec_prim"].cpu().numpy(), b["rec_prim"].cpu().numpy())


def test_config2_twin_poisson_parity():
    import wos_ext
    c = C.config2(res=96, twin=True)
    sc = wos_ext.Scene(_t(c["vertices"]), _t(c["indices"]), _t(c["values"]), use_bvh=True, source_type=3,
                       source_params=_t(c["source_params"]))
    w = wos_ext.WoS(nwalks=128, nsteps=256, epsilon=1e-4)
    out = w.walk_reduce(_t(c["points"]), sc, seed=0, stats=True)
    m, v = out["mean64"].cpu().numpy(), out["var64"].cpu().numpy()
    err = m - c["exact"]
    assert np.sqrt(np.mean(err ** 2)) < 1.2 * np.sqrt(np.mean(v / 128)) + 2e-3
    ref = O.Scene(c["vertices"], c["indices"], c["values"], 3, c["source_params"])
    sub = slice(None, None, 23)
    r = ref.walk_reduce(c["points"][sub], 128, 256, 1e-4, seed=0)
    z = _z(m[sub], v[sub], r["mean"], r["var"], 128)
    assert np.mean(z > 3) < 0.01


def test_walk_reduce_3d_parity():
    import wos_ext
    v, f, noise = C.displaced_sphere(3)
    vals = C.g_harmonic3(v).astype(np.float32)
    p = C.points_in_star(2000, noise)
    sc = wos_ext.Scene3D(_t(v), _t(f), _t(vals), use_bvh=True)
    w = wos_ext.WoS3D(nwalks=32, nsteps=256, epsilon=1e-3)
    out = w.walk_reduce(_t(p), sc, seed=0, stats=True)
    m, var = out["mean64"].cpu().numpy(), out["var64"].cpu().numpy()
    r = O.Scene(v, f, vals).walk_reduce(p[::5], 32, 256, 1e-3, seed=0)
    z = _z(m[::5], var[::5], r["mean"], r["var"], 32)
    assert np.mean(z > 3) < 0.01


def test_adjoint_values_and_coeffs():
    import wos_ext
    c = C.config2(res=40)
    sc = wos_ext.Scene(_t(c["vertices"]), _t(c["indices"]), _t(c["values"]), use_bvh=True, source_type=3,
                       source_params=_t(c["source_params"]))
    W = 32
    w = wos_ext.WoS(nwalks=W, nsteps=256, epsilon=1e-4)
    P = c["points"].shape[0]
    out = w.walk_reduce(_t(c["points"]), sc, seed=5, records=True, basis=True)
    g = np.random.default_rng(0).normal(size=P).astype(np.float32)
    gv, gc = wos_ext.adjoint_values(sc, P, W, out["rec_prim"], out["rec_t"], _t(g), basis=out["basis"],
                                    want_values=True, want_coeffs=True)
    ref = O.Scene(c["vertices"], c["indices"], c["values"], 3, c["source_params"])
    gv_ref = ref.adjoint_values(P, W, out["rec_prim"].cpu().numpy(), out["rec_t"].cpu().numpy(), g)
    np.testing.assert_allclose(gv.cpu().numpy(), gv_ref, rtol=1e-4, atol=1e-6)
    # the value sums are 64-bit fixed point: repeated launches agree bit for bit
    for _ in range(3):
        gv2, gc2 = wos_ext.adjoint_values(sc, P, W, out["rec_prim"], out["rec_t"], _t(g), basis=out["basis"],
                                          want_values=True, want_coeffs=True)
        assert torch.equal(gv, gv2) and torch.equal(gc, gc2)
    # coefficients: d mean / d c_k by central differences under common random numbers (exactly linear)
    K = c["source_params"].shape[0]
    for k in (0, 2, 5):
        h = 1e-2
        sp = c["source_params"].copy()
        sp[k, 3] += h
        sc.set_source(3, _t(sp))
        up = w.walk_reduce(_t(c["points"]), sc, seed=5)["mean"].cpu().numpy().astype(np.float64)
        sp[k, 3] -= 2 * h
        sc.set_source(3, _t(sp))
        dn = w.walk_reduce(_t(c["points"]), sc, seed=5)["mean"].cpu().numpy().astype(np.float64)
        fd = float(g.astype(np.float64) @ ((up - dn) / (2 * h)))
        assert abs(fd - float(gc[k])) <= 2e-3 * (abs(fd) + 1.0), (k, fd, float(gc[k]))
    assert gc.shape[0] == K


def test_basis_sums_match_oracle_and_coefficient_gradients_are_reproducible():
    """Per-point Fourier basis sums are 64-bit fixed-point sums: equal to the
    oracle's (same walks, same arithmetic, same fixed-point units) at every
    point whose walks all end on the oracle's records, and the coefficient
    gradient built from them is bit-identical across runs and launch tunings
    (the walks finish in a different order)."""
    import wos_ext
    c = C.config2(res=40)
    sp = c["source_params"]
    K = sp.shape[0]
    sc = wos_ext.Scene(_t(c["vertices"]), _t(c["indices"]), _t(c["values"]), use_bvh=True, source_type=3,
                       source_params=_t(sp))
    W = 32
    w = wos_ext.WoS(nwalks=W, nsteps=256, epsilon=1e-4, use_IS_for_greens=False)
    P = c["points"].shape[0]
    out = w.walk_reduce(_t(c["points"]), sc, seed=5, records=True, basis=True)
    ref = O.Scene(c["vertices"], c["indices"], c["values"], 3, sp).walk_reduce(
        c["points"], W, 256, 1e-4, seed=5, records=True, basis=True, use_is=False)
    assert out["basis"].dtype == torch.float64 and out["basis"].numel() == K * P
    gb = out["basis"].cpu().numpy().reshape(K, P).T
    same = (out["rec_prim"].cpu().numpy() == ref["rec_prim"]).reshape(P, W).all(axis=1)
    err = np.abs(gb[same] - ref["basis"][same]).max()
    print(f"points whose walks all exit where the oracle's do: {same.mean():.4f} of {P}; "
          f"max basis difference there {err:.3g} (scale {np.abs(ref['basis']).max():.3g})")
    assert same.mean() > 0.97
    np.testing.assert_array_equal(gb[same], ref["basis"][same])
    g = _t(np.random.default_rng(1).normal(size=P).astype(np.float32))
    _, gc1 = wos_ext.adjoint_values(sc, P, W, out["rec_prim"], out["rec_t"], g, basis=out["basis"],
                                    want_values=False, want_coeffs=True)
    before = wos_ext.get_tuning()
    try:
        wos_ext.set_tuning(chunk=96, wgs_per_cu=5, window=5)
        out2 = w.walk_reduce(_t(c["points"]), sc, seed=5, records=True, basis=True)
    finally:
        wos_ext.set_tuning(*before)
    _, gc2 = wos_ext.adjoint_values(sc, P, W, out2["rec_prim"], out2["rec_t"], g, basis=out2["basis"],
                                    want_values=False, want_coeffs=True)
    assert torch.equal(out["basis"], out2["basis"])
    assert torch.equal(gc1, gc2)
    # and against the oracle's basis in FP64: sum_i g_i / W * basis[i]
    gd = g.cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(gc1.cpu().numpy(), (gd @ ref["basis"]) / W, rtol=2e-3, atol=1e-6)


@pytest.mark.parametrize("dim", [2, 3])
def test_spatial_order_leaves_results_unchanged(dim):
    """Visiting the points in Morton order only reschedules walks: every
    walk's value, exit record and the per-point means are bit-identical."""
    import wos_ext
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    if dim == 2:
        c = C.config2(res=96)
        sc = wos_ext.Scene(t(c["vertices"]), t(c["indices"]), t(c["values"]), use_bvh=True, source_type=3,
                           source_params=t(c["source_params"]))
        w = wos_ext.WoS(nwalks=16, nsteps=256, epsilon=1e-4)
        pts = t(c["points"])
    else:
        v, f, noise = C.displaced_sphere(3)
        sc = wos_ext.Scene3D(t(v), t(f), t(C.g_harmonic3(v).astype(np.float32)), use_bvh=True)
        w = wos_ext.WoS3D(nwalks=16, nsteps=256, epsilon=1e-3)
        pts = t(C.points_in_star(8000, noise))
    try:
        wos_ext.set_spatial_order(True)
        a = w.walk_reduce(pts, sc, seed=4, records=True, basis=dim == 2)
        wos_ext.set_spatial_order(False)
        b = w.walk_reduce(pts, sc, seed=4, records=True, basis=dim == 2)
    finally:
        wos_ext.set_spatial_order(True)
    for k in a:
        assert torch.equal(a[k], b[k]), k


@pytest.mark.parametrize("dim", [2, 3])
def test_query_window_leaves_results_unchanged(dim):
    """The node visits per round (window; 0 = the per-dimension default)
    only reschedule queries: walk values, exit records and means are
    bit-identical for every setting."""
    import wos_ext
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    if dim == 2:
        c = C.config2(res=64)
        sc = wos_ext.Scene(t(c["vertices"]), t(c["indices"]), t(c["values"]), use_bvh=True, source_type=3,
                           source_params=t(c["source_params"]))
        w = wos_ext.WoS(nwalks=16, nsteps=256, epsilon=1e-4)
        pts = t(c["points"])
    else:
        v, f, noise = C.displaced_sphere(3)
        sc = wos_ext.Scene3D(t(v), t(f), t(C.g_harmonic3(v).astype(np.float32)), use_bvh=True)
        w = wos_ext.WoS3D(nwalks=16, nsteps=256, epsilon=1e-3)
        pts = t(C.points_in_star(3000, noise))
    before = wos_ext.get_tuning()
    try:
        outs = []
        for win in (0, 1, 8, 16, 37):
            wos_ext.set_tuning(window=win)
            assert wos_ext.get_tuning()[3] == win
            outs.append(w.walk_reduce(pts, sc, seed=9, records=True, basis=dim == 2))
    finally:
        wos_ext.set_tuning(*before)
    for o in outs[1:]:
        for k in outs[0]:
            assert torch.equal(outs[0][k], o[k]), k


def test_changed_values_match_a_fresh_scene():
    """Scene.set_values(g') (wos_scene_set_values) gives the same walks as a
    scene built with g': the leaf copies that carry the segment end-point
    values follow the change."""
    import wos_ext
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    c = C.config2(res=48)
    v, ix = t(c["vertices"]), t(c["indices"])
    g1 = t(c["values"])
    g2 = t((c["values"] * 1.7 + 0.25).astype(np.float32))
    w = wos_ext.WoS(nwalks=32, nsteps=256, epsilon=1e-4)
    pts = t(c["points"])
    sc = wos_ext.Scene(v, ix, g1, use_bvh=True)
    a1 = w.walk_reduce(pts, sc, seed=5, records=True)
    sc.set_values(g2)
    a2 = w.walk_reduce(pts, sc, seed=5, records=True)
    b2 = w.walk_reduce(pts, wos_ext.Scene(v, ix, g2, use_bvh=True), seed=5, records=True)
    for k in a2:
        assert torch.equal(a2[k], b2[k]), k
    assert not torch.equal(a1["mean"], a2["mean"])
    assert torch.equal(a1["rec_prim"], a2["rec_prim"])
