"""wos_walk_reduce_noalloc, the allocation-free walk_reduce, against
wos_walk_reduce through the C ABI: every output equal, on two calls in a
row with the second smaller, in 2D (Fourier source, basis and records) and
in 3D (a BVH deep enough to spill; 4 million points make the launch shrink
its grid to fit the spill into the static array); and its size limit."""
import numpy as np
import pytest
import torch

from wos import configs as C

pytestmark = pytest.mark.gpu

KEYS = ("mean", "mean64", "var64", "walk_value", "rec_prim", "rec_t")


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _same(a, b, keys):
    for k in keys:
        assert torch.equal(a[k], b[k]), k


def test_noalloc_walk_reduce_2d_source_equals_walk_reduce():
    import wos_ext
    c = C.config2(res=256)
    sc = wos_ext.Scene(_t(c["vertices"]), _t(c["indices"]), _t(c["values"]), use_bvh=True, source_type=3,
                       source_params=_t(c["source_params"]))
    cw = wos_ext.WoS(nwalks=16, nsteps=256, epsilon=1e-4, use_IS_for_greens=False)
    pts = _t(c["points"])
    for n, seed in ((pts.shape[0], 7), (3001, 8)):
        print(f"2D noalloc: {n} points, seed {seed}", flush=True)
        q = pts[:n].contiguous()
        ref = cw.walk_reduce(q, sc, seed=seed, records=True, basis=True, stats=True)
        got = cw.walk_reduce(q, sc, seed=seed, records=True, basis=True, stats=True, noalloc=True)
        torch.cuda.synchronize()
        _same(got, ref, KEYS + ("basis",))


def test_noalloc_walk_reduce_3d_spill_equals_walk_reduce():
    import wos_ext
    v, f, noise = C.displaced_sphere(6)
    sc = wos_ext.Scene3D(_t(v), _t(f), _t(C.g_harmonic3(v).astype(np.float32)), use_bvh=True)
    cw = wos_ext.WoS3D(nwalks=2, nsteps=128, epsilon=1e-3)
    big = _t(C.points_in_star(4_000_000, noise, seed=4).astype(np.float32))
    for q, W, seed in ((big, 2, 3), (big[:5000].contiguous(), 32, 4)):
        print(f"3D noalloc: {q.shape[0]} points, {W} walks, seed {seed}", flush=True)
        ref = cw.walk_reduce(q, sc, seed=seed, nwalks=W, records=True, stats=True)
        got = cw.walk_reduce(q, sc, seed=seed, nwalks=W, records=True, stats=True, noalloc=True)
        torch.cuda.synchronize()
        _same(got, ref, KEYS)
        del ref, got


def test_noalloc_walk_reduce_rejects_more_points_than_its_scratch_holds():
    import wos_ext
    v, ix = C.star_polyline(256)
    sc = wos_ext.Scene(_t(v), _t(ix), _t(C.g_saddle(v).astype(np.float32)), use_bvh=True)
    cw = wos_ext.WoS(nwalks=1, nsteps=8, epsilon=1e-3, use_IS_for_greens=False)
    q = torch.zeros((9_000_000, 2), dtype=torch.float32, device="cuda")
    with pytest.raises(wos_ext.WoSError, match="scratch"):
        cw.walk_reduce(q, sc, seed=0, noalloc=True)
    # the refused call launched nothing; the entry point still works
    out = cw.walk_reduce(q[:100], sc, seed=0, noalloc=True)
    ref = cw.walk_reduce(q[:100], sc, seed=0)
    torch.cuda.synchronize()
    assert torch.equal(out["mean"], ref["mean"])
