"""Batched multi-scene launch: every item equals its single-scene
walk_reduce bit for bit, and graph replays equal the eager run."""
import numpy as np
import pytest
import torch

from wos import configs as C

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


# 2D items with and without a Fourier source (odd items carry one) and 3D items
PICK = [0, 1, 5, 8, 13, 16, 21, 24, 29, 32, 36, 40, 44, 48, 52, 56, 60]


def _build(scale):
    import wos_ext
    items = []
    cfg = C.config5(scale=scale, seed=3)
    for c in (cfg[j] for j in PICK):
        if c["dim"] == 2:
            kw = dict(source_type=3, source_params=_t(c["source_params"])) if "source_type" in c else {}
            sc = wos_ext.Scene(_t(c["vertices"]), _t(c["indices"]), _t(c["values"]), use_bvh=True, **kw)
            cw = wos_ext.WoS(nwalks=c["nwalks"], nsteps=c["nsteps"], epsilon=c["epsilon"])
        else:
            sc = wos_ext.Scene3D(_t(c["vertices"]), _t(c["indices"]), _t(c["values"]), use_bvh=True)
            cw = wos_ext.WoS3D(nwalks=c["nwalks"], nsteps=c["nsteps"], epsilon=c["epsilon"])
        items.append(dict(scene=sc, wos=cw, points=_t(c["points"]), seed=c["seed"], records=True,
                          basis="source_type" in c))
    return items


def test_batch_matches_single_scene_runs():
    import wos_ext
    items = _build(0.01)
    b = wos_ext.Batch(items)
    outs = b.run()
    torch.cuda.synchronize()
    for it, o in zip(items, outs):
        ref = it["wos"].walk_reduce(it["points"], it["scene"], seed=it["seed"], records=True, basis=it["basis"])
        np.testing.assert_array_equal(o["walk_value"].cpu().numpy(), ref["walk_value"].cpu().numpy())
        np.testing.assert_array_equal(o["rec_prim"].cpu().numpy(), ref["rec_prim"].cpu().numpy())
        np.testing.assert_array_equal(o["mean"].cpu().numpy(), ref["mean"].cpu().numpy())
        if it["basis"]:
            np.testing.assert_array_equal(o["basis"].cpu().numpy(), ref["basis"].cpu().numpy())


def test_batch_graph_replay_matches_eager():
    import wos_ext
    items = _build(0.005)
    b = wos_ext.Batch(items)
    eager = [o["walk_value"].clone() for o in b.run()]
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        b.run()
    for rep in range(3):
        for o in b.outputs:
            o["walk_value"].fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for e, o in zip(eager, b.outputs):
            assert torch.equal(e, o["walk_value"]), rep


def test_autotune_keeps_results_and_caches(tmp_path, monkeypatch):
    """Tuning changes only speed: outputs are identical under the tuned
    setting, and a second call is answered from the cache."""
    import wos_ext
    from wos_ext.autotune import autotune
    monkeypatch.setenv("WOS_AMD_TUNING_CACHE", str(tmp_path / "tuning.json"))
    before = wos_ext.get_tuning()
    items = _build(0.005)
    b = wos_ext.Batch(items)
    ref = [o["walk_value"].clone() for o in b.run()]
    small = {"chunk": (64, 256), "wgs_per_cu": (4, 8), "top_levels": (0, 6), "window": (4, 16)}
    try:
        t = autotune(b.run, "test-batch", reps=1, candidates=small)
        assert not t["cached"]
        for r, o in zip(ref, b.run()):
            assert torch.equal(r, o["walk_value"])
        t2 = autotune(b.run, "test-batch", reps=1, candidates=small)
        assert t2["cached"] and [t2[k] for k in ("chunk", "wgs_per_cu", "top_levels", "window")] == \
            [t[k] for k in ("chunk", "wgs_per_cu", "top_levels", "window")]
    finally:
        wos_ext.set_tuning(*before)


def _with_grads(items):
    for it in items:
        it["grad_values"] = True
        it["grad_coeffs"] = it["basis"]
    return items


def test_batch_adjoint_matches_single_scene_adjoint():
    """The batched reverse pass gives each item the vertex-value and
    coefficient gradients wos_adjoint_values computes from the same records,
    bit for bit (fixed-point value sums, fixed-order coefficient sums)."""
    import wos_ext
    items = _with_grads(_build(0.01))
    b = wos_ext.Batch(items)
    b.run()
    rng = np.random.default_rng(2)
    gs = []
    for o in b.outputs:
        g = _t(rng.normal(size=o["mean"].shape[0]).astype(np.float32))
        o["grad_out"].copy_(g)
        gs.append(g)
    b.zero_grad()
    b.adjoint()
    torch.cuda.synchronize()
    for it, o, g in zip(items, b.outputs, gs):
        P, W = o["mean"].shape[0], it["wos"].nwalks
        gv, gc = wos_ext.adjoint_values(it["scene"], P, W, o["rec_prim"], o["rec_t"], g,
                                        basis=o.get("basis"), want_values=True, want_coeffs=it["basis"])
        # both sum the same fixed-point terms (and the coefficient sums in the same order): bit-identical
        assert torch.equal(o["grad_values"], gv)
        if it["basis"]:
            assert torch.equal(o["grad_coeffs"], gc)


def test_batch_primal_loss_adjoint_in_one_graph():
    """run + L2-loss gradient + zero_grad + adjoint captured together: every
    replay reproduces the eager step."""
    import wos_ext
    items = _with_grads(_build(0.005))
    b = wos_ext.Batch(items)
    targets = [torch.zeros_like(o["mean"]) + 0.1 for o in b.outputs]

    def step():
        b.run()
        for o, tg in zip(b.outputs, targets):
            o["grad_out"].copy_((2.0 / tg.numel()) * (o["mean"] - tg))
        b.zero_grad()
        b.adjoint()

    step()
    torch.cuda.synchronize()
    eager = [(o["mean"].clone(), o["grad_values"].clone()) for o in b.outputs]
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        step()
    for rep in range(3):
        for o in b.outputs:
            o["mean"].fill_(-1.0)
            o["grad_values"].fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for (m, gv), o in zip(eager, b.outputs):
            assert torch.equal(m, o["mean"]), rep
            scale = float(gv.abs().max()) + 1e-12
            assert float((gv - o["grad_values"]).abs().max()) <= 1e-5 * scale, rep


def test_batch_refuses_a_changed_scene():
    """After an item's scene changes its source (the item table holds the
    old buffer), run and adjoint raise until the batch is created again;
    new vertex values are simply read by the next run."""
    import wos_ext
    items = _with_grads(_build(0.005))
    b = wos_ext.Batch(items)
    b.run()
    j = next(i for i, it in enumerate(items) if it["basis"])
    sc = items[j]["scene"]
    sc.set_values(sc.c_values * 2.0)
    b.run()   # values are read in place
    sp = sc.c_source_params.clone()
    sc.set_source(3, torch.cat([sp, sp]))   # more terms: a new buffer
    with pytest.raises(wos_ext.WoSError, match="changed"):
        b.run()
    with pytest.raises(wos_ext.WoSError, match="changed"):
        b.adjoint()
    items[j]["scene"] = sc
    b2 = wos_ext.Batch(items)
    b2.run()
    torch.cuda.synchronize()


GRAD2 = dict(nwalks=4, nsteps=128, epsilon=1e-3, epsilon2=5e-4, clamping=1e-1)
GRAD3 = dict(nwalks=4, nsteps=128, epsilon=1e-3, epsilon2=1e-3, clamping=1e-1)


def _with_shape_grads(items):
    for it in items:
        it["grad_vertices"] = True
        it["grad_params"] = GRAD2 if it["scene"].dim == 2 else GRAD3
    return items


def test_batch_shape_gradient_matches_single_scene():
    """One persistent launch per dimension computes every item's OursCUDA /
    Ours3D vertex gradient: the samples are the single-scene launch's (same
    seed, same streams) and both sum them in 64-bit fixed point, so each item
    equals wos_shape_gradient on it alone bit for bit."""
    import wos_ext
    items = _with_shape_grads(_build(0.01))
    b = wos_ext.Batch(items)
    rng = np.random.default_rng(4)
    gs = []
    for o in b.outputs:
        g = _t(rng.normal(size=o["mean"].shape[0]).astype(np.float32))
        o["grad_out"].copy_(g)
        gs.append(g)
    b.zero_grad()
    b.shape_gradient()
    torch.cuda.synchronize()
    for it, o, g in zip(items, b.outputs, gs):
        ref, _ = wos_ext.shape_gradient(it["scene"], it["points"], it["seed"], g, **it["grad_params"])
        assert float(ref.abs().max()) > 1e-9
        # the same samples summed in the same fixed-point units: bit-identical
        assert torch.equal(o["grad_vertices"], ref), it["scene"].dim


def test_batch_primal_adjoint_and_shape_gradient_in_one_graph():
    """run + loss gradient + zero_grad + adjoint + shape_gradient captured as
    one graph: every replay reproduces the eager step (means bit for bit,
    gradients up to the order of float additions)."""
    import wos_ext
    items = _with_shape_grads(_with_grads(_build(0.005)))
    b = wos_ext.Batch(items)
    targets = [torch.zeros_like(o["mean"]) + 0.1 for o in b.outputs]

    def step():
        b.run()
        for o, tg in zip(b.outputs, targets):
            o["grad_out"].copy_((2.0 / tg.numel()) * (o["mean"] - tg))
        b.zero_grad()
        b.adjoint()
        b.shape_gradient()

    step()
    torch.cuda.synchronize()
    eager = [(o["mean"].clone(), o["grad_values"].clone(), o["grad_vertices"].clone()) for o in b.outputs]
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        step()
    for rep in range(3):
        for o in b.outputs:
            o["mean"].fill_(-1.0)
            o["grad_values"].fill_(7.0)
            o["grad_vertices"].fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for (m, gv, gx), o in zip(eager, b.outputs):
            assert torch.equal(m, o["mean"]), rep
            for ref, got in ((gv, o["grad_values"]), (gx, o["grad_vertices"])):
                scale = float(ref.abs().max()) + 1e-12
                assert float((ref - got).abs().max()) <= 1e-4 * scale, rep
