"""GPU parity of the scene queries: closest primitive (bit-exact against the
in-order brute-force scan of the oracle and of the device), records, sdf,
largest inscribed ball, LBVH topology (bit-exact against the oracle's Karras
build)."""
import numpy as np
import pytest
import torch

import oracle as O
from wos import configs as C

pytestmark = pytest.mark.gpu


def _scene2(n=1024, kind="circle", use_bvh=True):
    import wos_ext
    v, ix = C.circle_polyline(n) if kind == "circle" else C.star_polyline(n)
    vals = C.g_circle(v).astype(np.float32)
    sc = wos_ext.Scene(torch.from_numpy(v).cuda(), torch.from_numpy(ix).cuda(), torch.from_numpy(vals).cuda(),
                       use_bvh=use_bvh)
    return sc, O.Scene(v, ix, vals), v


def _scene3(subdiv=3, use_bvh=True):
    import wos_ext
    v, f, noise = C.displaced_sphere(subdiv)
    vals = C.g_harmonic3(v).astype(np.float32)
    sc = wos_ext.Scene3D(torch.from_numpy(v).cuda(), torch.from_numpy(f).cuda(), torch.from_numpy(vals).cuda(),
                         use_bvh=use_bvh)
    return sc, O.Scene(v, f, vals), v, noise


def _query_points_2d(rng, n):
    p = rng.uniform(-1.6, 1.6, size=(n, 2)).astype(np.float32)
    # exact ties: points on the bisector of two adjacent segments
    t = 2 * np.pi * (np.arange(64) + 0.5) / 1024
    q = np.stack([0.5 * np.cos(t), 0.5 * np.sin(t)], 1).astype(np.float32)
    return np.concatenate([p, q, np.zeros((1, 2), np.float32)])


@pytest.mark.parametrize("kind,n", [("circle", 1024), ("clover", 4096), ("clover", 7)])
def test_closest_index_bitexact_2d(kind, n):
    rng = np.random.default_rng(1)
    sc, ref, _ = _scene2(n, kind)
    p = _query_points_2d(rng, 20000)
    pt = torch.from_numpy(p).cuda()
    i_bvh, d_bvh = sc.closest_point_index_d2(pt)
    sc.use_bvh = False
    i_bf, d_bf = sc.closest_point_index_d2(pt)
    i_ref, d_ref = ref.closest_index(p)
    assert (i_bvh.cpu().numpy() == i_bf.cpu().numpy()).all()
    assert (i_bf.cpu().numpy() == i_ref).all()
    np.testing.assert_array_equal(d_bvh.cpu().numpy(), d_ref)


@pytest.mark.parametrize("res", [0, 12, 96])
def test_closest_index_bitexact_3d_start_grid(res):
    rng = np.random.default_rng(5)
    sc, ref, v, noise = _scene3(4)
    assert sc.start_grid_info()["nz"] == 0    # 3D scenes are built without one
    sc.set_start_grid(res)
    assert (sc.start_grid_info()["nz"] > 0) == (res > 0)
    p = np.concatenate([rng.uniform(-1.3, 1.3, size=(6000, 3)), C.points_in_star(4000, noise),
                        v[:500] * (1 + rng.normal(scale=1e-3, size=(500, 1))), rng.uniform(-9, 9, size=(300, 3))]
                       ).astype(np.float32)
    idx, d2 = sc.closest_point_index_d2(torch.from_numpy(p).cuda())
    i_ref, d_ref = ref.closest_index(p)
    np.testing.assert_array_equal(idx.cpu().numpy(), i_ref)
    np.testing.assert_array_equal(d2.cpu().numpy(), d_ref)


def test_closest_index_bitexact_3d():
    rng = np.random.default_rng(2)
    sc, ref, v, noise = _scene3(4)
    p = np.concatenate([rng.uniform(-1.3, 1.3, size=(6000, 3)), C.points_in_star(4000, noise),
                        v[:200] * 0.999]).astype(np.float32)
    pt = torch.from_numpy(p).cuda()
    i_bvh, d_bvh = sc.closest_point_index_d2(pt)
    sc.use_bvh = False
    i_bf, _ = sc.closest_point_index_d2(pt)
    i_ref, d_ref = ref.closest_index(p)
    assert (i_bvh.cpu().numpy() == i_bf.cpu().numpy()).all()
    assert (i_bf.cpu().numpy() == i_ref).all()
    np.testing.assert_array_equal(d_bvh.cpu().numpy(), d_ref)


@pytest.mark.parametrize("dim", [2, 3])
def test_lbvh_topology_matches_oracle(dim):
    if dim == 2:
        sc, ref, _ = _scene2(4096, "clover")
    else:
        sc, ref, _, _ = _scene3(5)
    prim, codes, ch = sc.bvh_topology()
    b = ref.lbvh()
    np.testing.assert_array_equal(prim.cpu().numpy(), b["prim"])
    np.testing.assert_array_equal(codes.cpu().numpy().view(np.uint32), b["codes"])
    np.testing.assert_array_equal(ch.cpu().numpy(), b["children"])
    assert sc.max_depth == b["max_depth"]


@pytest.mark.parametrize("dim", [2, 3])
def test_records_sdf_lib_match_oracle(dim):
    rng = np.random.default_rng(3)
    if dim == 2:
        sc, ref, _ = _scene2(4096, "clover")
        p = rng.uniform(-1.2, 1.2, size=(3000, 2)).astype(np.float32)
    else:
        sc, ref, _, noise = _scene3(4)
        p = C.points_in_star(3000, noise)
    pt = torch.from_numpy(p).cuda()
    rec = sc.closest_point(pt)
    r = ref.closest_point(p)
    np.testing.assert_array_equal(rec.prim_id.cpu().numpy(), r["prim"])
    np.testing.assert_array_equal(rec.p.cpu().numpy(), r["p"])
    np.testing.assert_array_equal(rec.n.cpu().numpy(), r["n"])
    tt = rec.t if dim == 2 else rec.uv
    np.testing.assert_array_equal(tt.cpu().numpy(), r["t"])
    np.testing.assert_array_equal(sc.sdf_native(pt).cpu().numpy(), ref.sdf(p))
    R = sc.largest_inscribed_ball(rec).cpu().numpy()
    Rr = ref.largest_inscribed_ball(r["valid"], r["p"], r["n"], 1e-3)
    np.testing.assert_array_equal(R, Rr)
    val = sc.dirichlet(rec).cpu().numpy()
    assert np.isfinite(val).all()



@pytest.mark.parametrize("n", [1, 3, 65])
def test_persistent_3d_queries_small_sizes(n):
    """The 3D point queries run on persistent lanes that claim points in
    chunks: point counts below, at and just above a wave, against the oracle
    (closest index and squared distance, sdf, records, and LIB with one record
    marked invalid)."""
    rng = np.random.default_rng(n)
    sc, ref, _, noise = _scene3(4)
    p = C.points_in_star(n, noise, seed=int(rng.integers(1 << 30)))
    pt = torch.from_numpy(p).cuda()
    idx, d2 = sc.closest_point_index_d2(pt)
    i_ref, d_ref = ref.closest_index(p)
    np.testing.assert_array_equal(idx.cpu().numpy(), i_ref)
    np.testing.assert_array_equal(d2.cpu().numpy(), d_ref)
    np.testing.assert_array_equal(sc.sdf_native(pt).cpu().numpy(), ref.sdf(p))
    rec = sc.closest_point(pt)
    r = ref.closest_point(p)
    np.testing.assert_array_equal(rec.prim_id.cpu().numpy(), r["prim"])
    rec.valid[0] = False
    r["valid"][0] = 0
    R = sc.largest_inscribed_ball(rec).cpu().numpy()
    Rr = ref.largest_inscribed_ball(r["valid"], r["p"], r["n"], 1e-3)
    np.testing.assert_array_equal(R, Rr)
    assert R[0] == 0.0

def test_scene_errors():
    import wos_ext
    v = torch.tensor([[0.0, 0.0], [1.0, 0.0], [1.0, 0.0]]).cuda()
    ix = torch.tensor([[0, 1], [1, 2], [2, 0]], dtype=torch.int32).cuda()
    with pytest.raises(wos_ext.WoSError, match="degenerate"):
        wos_ext.Scene(v, ix, torch.zeros(3).cuda())
    ix2 = torch.tensor([[0, 1], [1, 5]], dtype=torch.int32).cuda()
    with pytest.raises(wos_ext.WoSError, match="out of range"):
        wos_ext.Scene(v, ix2, torch.zeros(3).cuda())


def test_vertex_normals_3d_angle_weighted_and_reproducible():
    """3D vertex normals: the angle-weighted sum of unit face normals
    (include/scene.cuh:381-408) against an FP64 numpy restatement, and the
    same bits from two builds (the sums are order-independent)."""
    import wos_ext
    v, f, _ = C.displaced_sphere(4)
    vals = C.g_harmonic3(v).astype(np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    n1 = wos_ext.Scene3D(t(v), t(f), t(vals)).normals
    n2 = wos_ext.Scene3D(t(v), t(f), t(vals)).normals
    assert torch.equal(n1, n2)
    p = v.astype(np.float64)[f]                                   # (F, 3 corners, 3)
    fn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    ref = np.zeros_like(v, dtype=np.float64)
    for c in range(3):
        a = p[:, (c + 1) % 3] - p[:, c]
        b = p[:, (c + 2) % 3] - p[:, c]
        ang = np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), np.einsum("ij,ij->i", a, b))
        np.add.at(ref, f[:, c], ang[:, None] * fn)
    ref /= np.linalg.norm(ref, axis=1, keepdims=True)
    np.testing.assert_allclose(n1.cpu().numpy(), ref, atol=2e-6)
