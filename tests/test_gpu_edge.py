"""Edge cases of the entry points: empty point sets (every query, walk and
gradient entry point returns empty or zero outputs without launching work),
batches holding empty items, and scenes of a single primitive (2D segment,
3D triangle) checked against the oracle's brute-force scan."""
import numpy as np
import pytest
import torch

import oracle as O
from wos import configs as C

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _scenes():
    import wos_ext
    c2 = C.config2(res=32)
    s2 = wos_ext.Scene(_t(c2["vertices"]), _t(c2["indices"]), _t(c2["values"]), use_bvh=True, source_type=3,
                       source_params=_t(c2["source_params"]))
    v3, f3 = C.icosphere(2)
    s3 = wos_ext.Scene3D(_t(v3.astype(np.float32)), _t(f3.astype(np.int32)),
                         _t(C.g_harmonic3(v3).astype(np.float32)), use_bvh=True)
    return (s2, wos_ext.WoS, 2), (s3, wos_ext.WoS3D, 3)


def test_empty_point_sets():
    import wos_ext
    from wos.fwd import PCG32
    for sc, WoS, dim in _scenes():
        e = torch.zeros((0, dim), dtype=torch.float32, device="cuda")
        assert sc.closest_point_preliminary(e).shape == (0,)
        idx, d2 = sc.closest_point_index_d2(e)
        assert idx.shape == (0,) and d2.shape == (0,)
        assert sc.sdf_native(e).shape == (0,)
        rec = sc.closest_point(e)
        assert rec.p.shape[0] == 0
        assert sc.largest_inscribed_ball(rec).shape == (0,)
        assert sc.dirichlet(rec).shape == (0,)
        w = WoS(nwalks=16, nsteps=64, epsilon=1e-3)
        out = w.walk_reduce(e, sc, seed=1, records=True, stats=True, basis=(dim == 2))
        for k, v in out.items():
            assert v.numel() == 0, k
        cnt = w.walk_counts(e, sc, seed=1, basis=(dim == 2))
        assert cnt["walks"] == 0 and cnt["steps"] == 0 and cnt["queries"] == 0
        r = w.single_walk(e, sc, PCG32(0, 1, 2))
        assert r.p.shape[0] == 0
        gv, _ = wos_ext.shape_gradient(sc, e, 3, torch.zeros(0, device="cuda"), nwalks=8, nsteps=64, epsilon=1e-3,
                                       epsilon2=1e-3, clamping=0.1)
        assert gv.shape == sc.c_vertices.shape and float(gv.abs().sum()) == 0.0
        if dim == 2:
            gvals, gc = wos_ext.adjoint_values(sc, 0, 16, out["rec_prim"], out["rec_t"], torch.zeros(0, device="cuda"),
                                               basis=out["basis"], want_coeffs=True)
            assert float(gvals.abs().sum()) == 0.0 and float(gc.abs().sum()) == 0.0


def test_batch_with_empty_items():
    import wos_ext
    (s2, WoS, _), (s3, WoS3D, _) = _scenes()
    c2 = C.config2(res=32)
    w2, w3 = WoS(nwalks=32, nsteps=128, epsilon=1e-3), WoS3D(nwalks=16, nsteps=64, epsilon=1e-3)
    p2 = _t(c2["points"][:300])
    rng = np.random.default_rng(3)
    p3 = _t((rng.uniform(-0.5, 0.5, (200, 3))).astype(np.float32))
    e2 = torch.zeros((0, 2), dtype=torch.float32, device="cuda")
    e3 = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
    items = [dict(scene=s2, wos=w2, points=e2, seed=1), dict(scene=s2, wos=w2, points=p2, seed=5, records=True),
             dict(scene=s3, wos=w3, points=e3, seed=2), dict(scene=s2, wos=w2, points=e2, seed=4),
             dict(scene=s3, wos=w3, points=p3, seed=7), dict(scene=s3, wos=w3, points=e3, seed=9)]
    b = wos_ext.Batch(items)
    outs = b.run()
    torch.cuda.synchronize()
    for j in (0, 2, 3, 5):
        for k, v in outs[j].items():
            assert v.numel() == 0, (j, k)
    ref2 = w2.walk_reduce(p2, s2, seed=5, records=True, stats=True)
    ref3 = w3.walk_reduce(p3, s3, seed=7, stats=True)
    for k in ("walk_value", "rec_prim", "rec_t", "mean64", "var64"):
        np.testing.assert_array_equal(outs[1][k].cpu().numpy(), ref2[k].cpu().numpy(), err_msg=k)
    for k in ("walk_value", "mean64", "var64"):
        np.testing.assert_array_equal(outs[4][k].cpu().numpy(), ref3[k].cpu().numpy(), err_msg=k)


@pytest.mark.parametrize("dim", [2, 3])
def test_single_primitive_scene(dim):
    import wos_ext
    rng = np.random.default_rng(dim)
    if dim == 2:
        v = np.array([[-0.4, -0.1], [0.5, 0.2]], np.float32)
        f = np.array([[0, 1]], np.int32)
        sc = wos_ext.Scene(_t(v), _t(f), _t(np.array([1.0, 3.0], np.float32)), use_bvh=True)
    else:
        v = np.array([[-0.4, -0.1, 0.0], [0.5, 0.2, 0.1], [0.0, 0.6, -0.2]], np.float32)
        f = np.array([[0, 1, 2]], np.int32)
        sc = wos_ext.Scene3D(_t(v), _t(f), _t(np.array([1.0, 3.0, 2.0], np.float32)), use_bvh=True)
    p = rng.uniform(-1.0, 1.0, (5000, dim)).astype(np.float32)
    idx, d2 = sc.closest_point_index_d2(_t(p))
    i_ref, d_ref = O.Scene(v, f, np.zeros(len(v), np.float32)).closest_index(p)
    np.testing.assert_array_equal(idx.cpu().numpy(), i_ref)
    np.testing.assert_array_equal(d2.cpu().numpy(), d_ref)
    WoS = wos_ext.WoS if dim == 2 else wos_ext.WoS3D
    out = WoS(nwalks=8, nsteps=32, epsilon=1e-2, double_sided=True).walk_reduce(_t(p[:256]), sc, seed=1, stats=True)
    assert np.isfinite(out["mean64"].cpu().numpy()).all()
