"""GPU parity of the gradients: the shape-gradient samples against the oracle,
torch autograd through the drop-in solvers, and a directional finite
difference of the shape derivative."""
import numpy as np
import pytest
import torch

import oracle as O
from wos import configs as C

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("dim", [2, 3])
def test_shape_gradient_samples_match_oracle(dim):
    import wos_ext
    rng = np.random.default_rng(0)
    if dim == 2:
        v, ix = C.star_polyline(1024)
        vals = C.g_saddle(v).astype(np.float32)
        p = rng.uniform(-0.5, 0.5, size=(200, 2)).astype(np.float32)
        sc = wos_ext.Scene(_t(v), _t(ix), _t(vals), use_bvh=True)
    else:
        v, ix, noise = C.displaced_sphere(2)
        vals = C.g_harmonic3(v).astype(np.float32)
        p = C.points_in_star(100, noise)
        sc = wos_ext.Scene3D(_t(v), _t(ix), _t(vals), use_bvh=True)
    W = 8
    kw = dict(nwalks=W, nsteps=128, epsilon=1e-3, epsilon2=1e-3, clamping=0.1)
    g = rng.normal(size=p.shape[0]).astype(np.float32)
    gv, sg = wos_ext.shape_gradient(sc, _t(p), 3, _t(g), sample_grad_u=True, **kw)
    ref = O.Scene(v, ix, vals)
    gv_ref, sg_ref = ref.shape_gradient(p, W, 128, 1e-3, 1e-3, 0.1, seed=3, grad_out=g, sample_grad_u=True)
    sg = sg.cpu().numpy()
    close = np.all(np.abs(sg - sg_ref) <= 1e-3 * (1 + np.abs(sg_ref)), axis=1)
    print(f"{dim}D gradient samples equal to the oracle's: {close.mean():.5f} of {close.size}")
    assert close.mean() > 0.99, close.mean()
    gv = gv.cpu().numpy()
    scale = np.abs(gv_ref).max()
    assert np.abs(gv - gv_ref).max() < 0.1 * scale


def test_autograd_values_and_source_coefficients():
    from wos.scene import Polyline
    from wos.wos_with_source import WoSWithSourceCUDA
    import wos_ext
    c = C.config2(res=40)
    vals = _t(c["values"]).requires_grad_(True)
    sp = _t(c["source_params"]).requires_grad_(True)
    scene = Polyline(_t(c["vertices"]), _t(c["indices"]), vals, source_type=3, source_params=sp)
    solver = WoSWithSourceCUDA(nwalks=16, nsteps=256, epsilon=1e-4)
    pts = _t(c["points"])
    u = solver.walk(pts, scene, seed=4)
    loss = (u ** 2).mean()
    loss.backward()
    out = solver.cwos.walk_reduce(pts, scene, seed=4, records=True, basis=True)
    np.testing.assert_array_equal(out["mean"].cpu().numpy(), u.detach().cpu().numpy())
    g = 2 * u.detach() / u.shape[0]
    gv, gc = wos_ext.adjoint_values(scene, pts.shape[0], 16, out["rec_prim"], out["rec_t"], g, basis=out["basis"],
                                    want_values=True, want_coeffs=True)
    np.testing.assert_allclose(vals.grad.cpu().numpy(), gv.cpu().numpy(), rtol=1e-5, atol=1e-8)
    # coefficient column: the re-walked source adjoint equals the stored-basis adjoint
    np.testing.assert_allclose(sp.grad[:, 3].cpu().numpy(), gc.cpu().numpy(), rtol=2e-4, atol=1e-7)
    # frequency and phase columns: central differences of the loss under common random numbers (the paths
    # do not depend on the source, so the estimate is a smooth function of its parameters)
    g64 = g.double()
    for (k, col) in ((0, 0), (1, 2), (1, 1), (0, 2)):   # terms 0 and 1 carry the nonzero coefficients
        h = 1e-2
        fd = 0.0
        for sgn in (1, -1):
            spp = sp.detach().clone()
            spp[k, col] += sgn * h
            scene.set_source(3, spp)
            up = solver.cwos.walk_reduce(pts, scene, seed=4, stats=True)["mean64"]
            fd += sgn * float((g64 * up).sum()) / (2 * h)
        scene.set_source(3, sp.detach())
        ad = float(sp.grad[k, col])
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-6, (k, col, ad, fd)


def test_custom_solver_shape_gradient_directional_fd():
    """d L / d s for a uniform scaling s of the boundary: adjoint (OursCUDA
    through CustomSolver) against the closed form -4 mean g^2 (g = x^2 - y^2
    is harmonic on the plane, so u_s(x) = g(x / s)) and a central difference
    of the primal under common seeds.  CustomSolver replays the forward seed
    in the backward pass (wos/solver.py:105-110), so its gradient also
    carries Cov(u, du)/W, a term that falls as 1/W: 2,048 walks keep it
    near 1%."""
    from wos.scene import Polyline
    from wos.solver import CustomSolver
    from wos.wos import WoSCUDA
    from wos.wos_grad import OursCUDA
    v, ix = C.star_polyline(1024)
    vals = _t(C.g_saddle(v).astype(np.float32))
    rng = np.random.default_rng(1)
    r, t = 0.55 * np.sqrt(rng.random(2048)), 2 * np.pi * rng.random(2048)
    pts = _t(np.stack([r * np.cos(t), r * np.sin(t)], 1).astype(np.float32))   # inside for all scalings used
    target = torch.zeros(pts.shape[0], device="cuda")
    base = _t(v)

    def loss_at(s, solver, grad=False):
        sv = torch.tensor(s, device="cuda", requires_grad=grad)
        scene = Polyline(base * sv, _t(ix), vals)
        u = solver.walk(pts, scene, seed=11)
        L = ((u - target) ** 2).mean()
        if grad:
            L.backward()
            return L.item(), sv.grad.item()
        return L.item()

    fwd = WoSCUDA(nwalks=2048, nsteps=256, epsilon=1e-4)
    bwd = OursCUDA(nwalks=2048, nsteps=256, epsilon=1e-4, epsilon2=1e-4, clamping=0.1)
    _, dL = loss_at(1.0, CustomSolver(fwd, bwd), grad=True)
    h = 2e-2
    fd = (loss_at(1 + h, fwd) - loss_at(1 - h, fwd)) / (2 * h)
    exact = -4.0 * float(np.mean(C.g_saddle(np.stack([r * np.cos(t), r * np.sin(t)], 1)) ** 2))
    print(f"CustomSolver scaling: adjoint {dL:+.5e}, closed form {exact:+.5e}, central difference {fd:+.5e}")
    assert np.isfinite(dL)
    assert abs(dL - exact) < 0.05 * abs(exact), (dL, exact)
    assert abs(fd - exact) < 0.03 * abs(exact), (fd, exact)


def test_ours3d_directional_fd_on_64_vertices():
    """Scaled-down shape-gradient workload: gradient w.r.t. vertex positions
    (Ours3D through CustomSolver) against central differences of the loss,
    common seeds, along two deformations: 64 random vertices pushed along
    their normals, and a smooth anisotropic stretch of the whole mesh.

    The backward pass replays the forward seed (wos/solver.py:105-110), so
    the adjoint estimates the derivative of E[L] at the backward walk count;
    that bias shrinks as 1/W, hence 512 backward walks.  Directions with
    large random tangential moves fold the 0.05-edge mesh at these step sizes
    and make the difference quotient meaningless, so they are not used."""
    from wos.scene3d import Scene3D
    from wos.solver import CustomSolver
    from wos.wos3d import WoS3DCUDA
    from wos.wos_grad_3d import Ours3D
    c = C.config4(n_tris=5000, n_points=4096)
    V0 = _t(c["vertices"])
    vals = _t(c["values"])
    pts = _t(c["points"])
    ix = _t(c["indices"])
    target = _t(c["target"].astype(np.float32))
    W = 512
    fwd = WoS3DCUDA(nwalks=W, nsteps=256, epsilon=1e-3)
    bwd = Ours3D(nwalks=W, nsteps=256, epsilon=1e-3, epsilon2=1e-3, clamping=0.1)

    def loss(V, solver, seed):
        u = solver.walk(pts, Scene3D(V, ix, vals), seed=seed)
        return ((u - target) ** 2).mean()

    grads = []
    for seed in range(5, 9):
        V = V0.clone().requires_grad_(True)
        loss(V, CustomSolver(fwd, bwd), seed).backward()
        grads.append(V.grad.detach().clone())
    rng = np.random.default_rng(9)
    sel = torch.from_numpy(rng.choice(V0.shape[0], 64, replace=False)).cuda()
    bump = torch.zeros_like(V0)
    bump[sel] = V0[sel] / V0[sel].norm(dim=1, keepdim=True)
    stretch = torch.zeros_like(V0)
    stretch[:, 0] = 0.5 * V0[:, 0]
    stretch[:, 2] = -0.3 * V0[:, 2]
    fwd_fd = WoS3DCUDA(nwalks=4096, nsteps=256, epsilon=1e-3)
    h = 2e-2
    for name, dirn in (("bump64", bump), ("stretch", stretch)):
        ads = np.array([float((g * dirn).sum()) for g in grads])
        ad, se = ads.mean(), ads.std(ddof=1) / np.sqrt(len(ads))
        fd = np.mean([(loss(V0 + h * dirn, fwd_fd, s).item() - loss(V0 - h * dirn, fwd_fd, s).item()) / (2 * h)
                      for s in (5, 6)])
        print(f"{name}: adjoint {ad:+.4e} +- {se:.1e}, central difference {fd:+.4e}")
        assert np.isfinite(ad)
        assert abs(ad - fd) < 0.15 * abs(fd) + 3 * se, (name, ad, se, fd)


@pytest.mark.parametrize("dim", [2, 3])
def test_shape_gradient_is_bit_reproducible(dim):
    """The vertex gradient is summed in 64-bit fixed point from staged
    per-sample terms: repeated launches, with a different launch tuning (the
    samples end in a different order), give identical bits."""
    import wos_ext
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    if dim == 2:
        c = C.config2(res=48)
        sc = wos_ext.Scene(t(c["vertices"]), t(c["indices"]), t(c["values"]), use_bvh=True)
        pts = t(c["points"])
        kw = dict(nwalks=8, nsteps=256, epsilon=1e-3, epsilon2=1e-3, clamping=0.1)
    else:
        v, f, noise = C.displaced_sphere(3)
        sc = wos_ext.Scene3D(t(v), t(f), t(C.g_harmonic3(v).astype(np.float32)), use_bvh=True)
        pts = t(C.points_in_star(4000, noise))
        kw = dict(nwalks=8, nsteps=256, epsilon=1e-3, epsilon2=1e-3, clamping=0.1)
    g = t(np.random.default_rng(7).normal(size=pts.shape[0]).astype(np.float32))
    a, _ = wos_ext.shape_gradient(sc, pts, 11, g, **kw)
    b, _ = wos_ext.shape_gradient(sc, pts, 11, g, **kw)
    before = wos_ext.get_tuning()
    try:
        wos_ext.set_tuning(chunk=96, wgs_per_cu=3, window=5)
        c2, _ = wos_ext.shape_gradient(sc, pts, 11, g, **kw)
    finally:
        wos_ext.set_tuning(*before)
    assert float(a.abs().max()) > 0
    assert torch.equal(a, b) and torch.equal(a, c2)
