"""GPU parity of the 2D query accelerators (near-boundary candidate grid,
start-node grid): closest primitive and squared distance bit-exact against
the oracle's in-order brute-force scan for points at distances 1e-7 .. 1e-1
on both sides of the boundary, under several grid shapes, and walks
identical with and without them."""
import numpy as np
import pytest
import torch

import oracle as O
from wos import configs as C

pytestmark = pytest.mark.gpu


def _polyline(kind, n):
    if kind == "circle":
        return C.circle_polyline(n)
    if kind == "star":
        return C.star_polyline(n)
    c = C.config2()
    if n == c["indices"].shape[0]:
        return c["vertices"], c["indices"]
    return C.star_polyline(n)


def _near_boundary_points(v, ix, rng, n):
    f = rng.integers(0, ix.shape[0], n)
    t = rng.random(n)
    a, b = v[ix[f, 0]].astype(np.float64), v[ix[f, 1]].astype(np.float64)
    x = a + t[:, None] * (b - a)
    nrm = np.stack([(b - a)[:, 1], -(b - a)[:, 0]], 1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = 10.0 ** rng.uniform(-7, -1, n) * rng.choice([-1.0, 1.0], n)
    p = x + d[:, None] * nrm
    # vertices themselves (exact ties between the two adjacent segments) and
    # points slightly off them
    k = rng.integers(0, v.shape[0], 256)
    q = v[k] + rng.normal(scale=1e-6, size=(256, 2))
    return np.concatenate([p, v[k], q]).astype(np.float32)


@pytest.mark.parametrize("kind,n", [("circle", 1024), ("clover", 4096), ("star", 2048), ("star", 7)])
@pytest.mark.parametrize("grid", [(1.0, 16, 48), (4.0, 4, 16), (0.25, 40, 96)])
def test_grid_closest_bitexact(kind, n, grid):
    import wos_ext
    v, ix = _polyline(kind, n)
    vals = C.g_circle(v).astype(np.float32)
    sc = wos_ext.Scene(torch.from_numpy(v).cuda(), torch.from_numpy(ix).cuda(), torch.from_numpy(vals).cuda())
    sc.set_grid(*grid)
    info = sc.grid_info()
    assert info["nx"] > 0 and info["entries"] > 0, info
    rng = np.random.default_rng(n)
    p = np.concatenate([_near_boundary_points(v, ix, rng, 20000),
                        rng.uniform(-1.6, 1.6, size=(4000, 2)).astype(np.float32)])
    idx, d2 = sc.closest_point_index_d2(torch.from_numpy(p).cuda())
    i_ref, d_ref = O.Scene(v, ix, vals).closest_index(p)
    np.testing.assert_array_equal(idx.cpu().numpy(), i_ref)
    np.testing.assert_array_equal(d2.cpu().numpy(), d_ref)


@pytest.mark.parametrize("kind,n", [("circle", 1024), ("clover", 4096), ("star", 2048), ("star", 7)])
@pytest.mark.parametrize("res", [0, 16, 512, 3000, -1])
def test_start_grid_closest_bitexact(kind, n, res):
    import wos_ext
    v, ix = _polyline(kind, n)
    vals = C.g_circle(v).astype(np.float32)
    sc = wos_ext.Scene(torch.from_numpy(v).cuda(), torch.from_numpy(ix).cuda(), torch.from_numpy(vals).cuda())
    sc.set_grid(0.0)           # isolate the start-node grid
    sc.set_start_grid(res)
    info = sc.start_grid_info()
    assert (info["nx"] > 0) == (res != 0), info
    if res < 0:
        # automatic: two cells per mean segment length along the longer side,
        # within [64, 2048] (plus a one-cell border on each side)
        seg = np.linalg.norm(v[ix[:, 1]] - v[ix[:, 0]], axis=1).mean()
        side = (v.max(0) - v.min(0)).max()
        assert max(info["nx"], info["ny"]) - 2 == pytest.approx(min(max(2 * side / seg, 64), 2048), abs=2), info
    rng = np.random.default_rng(n + abs(res))
    p = np.concatenate([_near_boundary_points(v, ix, rng, 20000),
                        rng.uniform(-1.6, 1.6, size=(4000, 2)).astype(np.float32),
                        rng.uniform(-40, 40, size=(500, 2)).astype(np.float32)])
    idx, d2 = sc.closest_point_index_d2(torch.from_numpy(p).cuda())
    i_ref, d_ref = O.Scene(v, ix, vals).closest_index(p)
    np.testing.assert_array_equal(idx.cpu().numpy(), i_ref)
    np.testing.assert_array_equal(d2.cpu().numpy(), d_ref)


def test_grid_removal():
    import wos_ext
    v, ix = C.circle_polyline(256)
    vals = C.g_circle(v).astype(np.float32)
    sc = wos_ext.Scene(torch.from_numpy(v).cuda(), torch.from_numpy(ix).cuda(), torch.from_numpy(vals).cuda())
    assert sc.grid_info()["nx"] > 0
    sc.set_grid(0.0)
    assert sc.grid_info() == {"nx": 0, "ny": 0, "cell": 0.0, "entries": 0}


def test_walks_identical_with_and_without_grid():
    import wos_ext
    c = C.config2(res=64)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    sc = wos_ext.Scene(t(c["vertices"]), t(c["indices"]), t(c["values"]), source_type=3,
                       source_params=t(c["source_params"]))
    w = wos_ext.WoS(nwalks=64, nsteps=256, epsilon=1e-4)
    pts = t(c["points"])
    a = w.walk_reduce(pts, sc, seed=3, records=True, basis=True)
    sc.set_grid(0.0)
    b = w.walk_reduce(pts, sc, seed=3, records=True, basis=True)
    sc.set_start_grid(0)
    c0 = w.walk_reduce(pts, sc, seed=3, records=True, basis=True)
    for k in ("mean", "walk_value", "rec_prim", "rec_t", "basis"):
        np.testing.assert_array_equal(a[k].cpu().numpy(), b[k].cpu().numpy(), err_msg=k)
        np.testing.assert_array_equal(a[k].cpu().numpy(), c0[k].cpu().numpy(), err_msg=k)


@pytest.mark.parametrize("n,res", [(70000, -1), (70000, 700), (65537, -1), (65538, -1)])
def test_start_grid_wide_node_ids_bitexact(n, res):
    """Trees at and beyond 16-bit internal node ids: 65,537 segments (ids up
    to 0xffff) still use the packed 8-byte cells, 65,538 and more the 16-byte
    ones."""
    import wos_ext
    v, ix = C.star_polyline(n)
    vals = C.g_circle(v).astype(np.float32)
    sc = wos_ext.Scene(torch.from_numpy(v).cuda(), torch.from_numpy(ix).cuda(), torch.from_numpy(vals).cuda())
    sc.set_grid(0.0)
    sc.set_start_grid(res)
    assert sc.start_grid_info()["nx"] > 0
    rng = np.random.default_rng(n + res)
    p = np.concatenate([_near_boundary_points(v, ix, rng, 3000),
                        rng.uniform(-1.6, 1.6, size=(500, 2)).astype(np.float32)])
    idx, d2 = sc.closest_point_index_d2(torch.from_numpy(p).cuda())
    i_ref, d_ref = O.Scene(v, ix, vals).closest_index(p)
    np.testing.assert_array_equal(idx.cpu().numpy(), i_ref)
    np.testing.assert_array_equal(d2.cpu().numpy(), d_ref)
