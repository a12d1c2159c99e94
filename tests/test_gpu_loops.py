"""The reference's loop estimators on torch + native queries: the 2D / 3D
Laplace loop walks against exact harmonic solutions, the Poisson loop walk
(uniform and importance-sampled interior term) against a manufactured
solution, the spatial-gradient and normal-derivative estimators against
closed forms, and the autograd integrators (Ours, Baseline) against the
native estimator and finite differences.  All comparisons are statistical:
per-point means within a few standard errors (stated in each test)."""
import math

import numpy as np
import pytest
import torch

from wos import configs as C

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _circle_scene(n=1024, g=C.g_circle):
    from wos.scene import Polyline
    v, ix = C.circle_polyline(n)
    return Polyline(_t(v), _t(ix), _t(g(v).astype(np.float32))), v


def _disk_points(n, r, seed=0):
    rng = np.random.default_rng(seed)
    rr, t = r * np.sqrt(rng.random(n)), 2 * np.pi * rng.random(n)
    return np.stack([rr * np.cos(t), rr * np.sin(t)], 1).astype(np.float32)


def test_loop_wos_2d_matches_harmonic_and_native():
    from wos.wos import WoS, WoSCUDA
    sc, _ = _circle_scene()
    p = _disk_points(400, 0.9)
    exact = C.g_circle(p)
    W = 256
    u_loop = WoS(nwalks=W, nsteps=256, epsilon=1e-3).walk(_t(p), sc, seed=3).cpu().numpy()
    u_cuda = WoSCUDA(nwalks=W, nsteps=256, epsilon=1e-3).walk(_t(p), sc, seed=3).cpu().numpy()
    # each walk value lies in [min g, max g] (|g| <= 1.5 on the circle): sd of a mean <= 1.5/sqrt(W)
    tol = 5 * 1.5 / math.sqrt(W)
    assert np.abs(u_loop - exact).max() < tol
    assert np.sqrt(np.mean((u_loop - exact) ** 2)) < 0.25 * tol
    assert np.sqrt(np.mean((u_loop - u_cuda) ** 2)) < 0.35 * tol


def test_loop_wos_with_source_uniform_and_importance_sampled():
    from wos.scene import Polyline
    from wos.wos_with_source import WoSWithSource
    c = C.config2(res=24, twin=True)
    sc = Polyline(_t(c["vertices"]), _t(c["indices"]), _t(c["values"]), source_type=3,
                  source_params=_t(c["source_params"]))
    p = c["points"][::2]
    exact = C.u_clover_twin(p)
    for is_g in (False, True):
        s = WoSWithSource(nwalks=512, nsteps=256, epsilon=1e-4, use_IS_for_greens=is_g)
        u = s.walk(_t(p), sc, seed=5).cpu().numpy()
        rmse = np.sqrt(np.mean((u - exact) ** 2))
        assert rmse < 0.05, (is_g, rmse)


def test_loop_wos3d_matches_harmonic():
    from wos.scene3d import Scene3D
    from wos.wos3d import WoS3D
    v, f, noise = C.displaced_sphere(3)
    sc = Scene3D(_t(v), _t(f), _t(C.g_harmonic3(v).astype(np.float32)))
    p = C.points_in_star(300, noise, shrink=0.8)
    exact = C.g_harmonic3(p)
    u = WoS3D(nwalks=256, nsteps=256, epsilon=1e-3).walk(_t(p), sc, seed=2).cpu().numpy()
    assert np.sqrt(np.mean((u - exact) ** 2)) < 0.06


def test_grad_estimators_match_closed_forms():
    """grad u of a harmonic u: 2D (x^2 - y^2 + 0.5x on the disk) and 3D
    (x^2 - y^2 + z on the displaced sphere), mean of many estimates."""
    from wos.fwd import PCG32
    from wos.wos import WoS
    from wos.wos3d import WoS3D
    sc, _ = _circle_scene()
    x = np.repeat(np.array([[0.3, -0.2], [-0.4, 0.1]], np.float32), 8192, axis=0)
    s = PCG32(x.shape[0], 11, 7)
    gx, gy = WoS(nwalks=1, nsteps=256, epsilon=1e-3).grad(_t(x), sc, s)
    est = np.stack([gx.cpu().numpy(), gy.cpu().numpy()], 1).reshape(2, 8192, 2).mean(1)
    ref = np.array([[2 * 0.3 + 0.5, 0.4], [2 * -0.4 + 0.5, -0.2]])
    assert np.abs(est - ref).max() < 0.06, (est, ref)

    from wos.scene3d import Scene3D
    v, f, _ = C.displaced_sphere(3)
    s3 = Scene3D(_t(v), _t(f), _t(C.g_harmonic3(v).astype(np.float32)))
    x3 = np.repeat(np.array([[0.2, 0.1, -0.1]], np.float32), 16384, axis=0)
    g3 = WoS3D(nwalks=1, nsteps=256, epsilon=1e-3).grad(_t(x3), s3, PCG32(x3.shape[0], 5, 9))
    est3 = np.array([g.mean().item() for g in g3])
    assert np.abs(est3 - np.array([0.4, -0.2, 1.0])).max() < 0.08, est3


@pytest.mark.parametrize("k", [2, 4])
def test_normal_derivative_on_the_disk(k):
    """u = r^k cos k theta on the unit disk (ablations/test_disk.py:23-37):
    du/dn = k cos k theta on the boundary."""
    from wos.fwd import PCG32
    from wos.wos import WoS
    sc, _ = _circle_scene(2048, g=lambda p: np.cos(k * np.arctan2(p[:, 1], p[:, 0])))
    th = np.array([0.3, 1.1, 2.5, 4.0])
    q = np.repeat(np.stack([0.999 * np.cos(th), 0.999 * np.sin(th)], 1).astype(np.float32), 4096, axis=0)
    its = sc.closest_point(_t(q))
    gx, gy = WoS(nwalks=1, nsteps=256, epsilon=1e-3).normal_derivative(its, sc, PCG32(q.shape[0], 3, 4))
    n = np.stack([np.cos(th), np.sin(th)], 1)
    dn = ((gx.cpu().numpy() * np.repeat(n[:, 0], 4096)) + (gy.cpu().numpy() * np.repeat(n[:, 1], 4096)))
    est = dn.reshape(4, 4096).mean(1)
    assert np.abs(est - k * np.cos(k * th)).max() < 0.15 * k, (est, k * np.cos(k * th))


def test_loop_ours_agrees_with_native_and_baseline_runs():
    """d L / d s for a uniform scaling of a star boundary: the loop
    integrator Ours (projected exits, loop normal derivative) against the
    native OursCUDA; the soft-edge Baseline returns finite gradients."""
    from wos.scene import Polyline
    from wos.solver import CustomSolver
    from wos.wos import Baseline, WoSCUDA
    from wos.wos_grad import Ours, OursCUDA
    v, ix = C.star_polyline(512)
    vals = _t(C.g_saddle(v).astype(np.float32))
    pts = _t(_disk_points(1024, 0.5, seed=4))
    base = _t(v)

    def dL(bwd, W):
        sv = torch.tensor(1.0, device="cuda", requires_grad=True)
        scene = Polyline(base * sv, _t(ix), vals)
        u = CustomSolver(WoSCUDA(nwalks=W, nsteps=256, epsilon=1e-3), bwd).walk(pts, scene, seed=1)
        (u ** 2).mean().backward()
        return sv.grad.item()

    g_loop = dL(Ours(nwalks=128, nsteps=256, epsilon=1e-3), 128)
    g_cuda = dL(OursCUDA(nwalks=128, nsteps=256, epsilon=1e-3, epsilon2=1e-3, clamping=0.1), 128)
    assert np.isfinite(g_loop) and np.isfinite(g_cuda)
    assert abs(g_loop - g_cuda) < 0.2 * abs(g_cuda) + 1e-3, (g_loop, g_cuda)

    sv = torch.tensor(1.0, device="cuda", requires_grad=True)
    scene = Polyline(base * sv, _t(ix), vals)
    u = Baseline(nwalks=4, nsteps=64, epsilon=5e-3).walk(pts[:64], scene, seed=2)
    u.sum().backward()
    assert torch.isfinite(sv.grad) and torch.isfinite(u).all()


def test_boundary_sampling_and_boundary_term():
    """Arc-length sampling: 1/pdf equals the polyline length for every
    sample; the boundary term's derivative under a uniform scaling s of the
    unit circle with g = x^2 integrates g (x . n) = -g over the boundary:
    d/ds sum(image) = npixels * (closed integral of x^2) / detector size."""
    from wos.fwd import PCG32
    from wos.scene import Detector, Polyline
    from wos.wos_boundary import WoSBoundary
    v, ix = C.circle_polyline(1024)
    base = _t(v)
    L = float(np.linalg.norm(v[ix[:, 1]] - v[ix[:, 0]], axis=1).sum())
    sc = Polyline(base, _t(ix), _t(v[:, 0] ** 2))
    b = sc.sample_boundary(PCG32(4096, 3))
    np.testing.assert_allclose((1.0 / b.pdf).cpu().numpy(), L, rtol=1e-5)

    s = torch.tensor(1.0, device="cuda", requires_grad=True)
    scene = Polyline(base * s, _t(ix), _t(v[:, 0] ** 2))
    det = Detector(vmin=(-1.2, -1.2), vmax=(1.2, 1.2), res=(32, 32))
    img = WoSBoundary(nwalks=64).walk_detector(scene, det, seed=7)
    assert float(img.abs().max()) == 0.0
    img.sum().backward()
    expect = 32 * 32 * math.pi / det.size()
    assert abs(s.grad.item() - expect) < 0.05 * expect, (s.grad.item(), expect)


def test_boundary_sampling_3d_uniform_on_area():
    from wos.fwd import PCG32
    from wos.scene3d import Scene3D
    v, f = C.icosphere(2)
    v = v.astype(np.float32)
    sc = Scene3D(_t(v), _t(f.astype(np.int32)), _t(v[:, 2]))
    area = float(np.linalg.norm(np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]]), axis=1).sum())
    b = sc.sample_boundary(PCG32(20000, 5))
    # 1/pdf = total |cross| (twice the area) for every sample; z is uniform on the sphere in the mean
    np.testing.assert_allclose((1.0 / b.pdf).cpu().numpy(), area, rtol=1e-4)
    assert abs(float(b.p[:, 2].mean())) < 0.03
    assert torch.allclose((b.p * b.n).sum(1), -b.p.norm(dim=1), atol=0.06)


def test_inverse_solve_recovers_boundary_scale(tmp_path):
    """inv_solve-style loop: recover the uniform scale of a star boundary
    from the solution at interior points (native forward walk, OursCUDA
    shape gradient, Adam), from 1.0 to the target 1.1.  The backward pass
    replays the forward seed (wos/solver.py:105-110), so the gradient also
    carries d/ds of the walk variance / W; 1024 walks keep that shift of the
    optimum well inside the tolerance."""
    from wos.optimize import Model, TrainRunner
    from wos.scene import Polyline
    from wos.solver import CustomSolver
    from wos.wos import WoSCUDA
    from wos.wos_grad import OursCUDA
    v, ix = C.star_polyline(512)
    base, idx = _t(v), _t(ix)
    vals = _t(2.0 * C.g_saddle(v).astype(np.float32) + 1.0)
    pts = _t(_disk_points(2048, 0.75, seed=8))
    W = 1024

    class ScaleModel(Model):
        def __init__(self, s, lr=0.01):
            self.s = torch.tensor(float(s), device="cuda", requires_grad=True)
            self.opt = torch.optim.Adam([self.s], lr=lr)
            self.solver = CustomSolver(WoSCUDA(nwalks=W, nsteps=256, epsilon=1e-3),
                                       OursCUDA(nwalks=W, nsteps=256, epsilon=1e-3, epsilon2=1e-3, clamping=0.1))

        def forward(self, i=0):
            scene = Polyline(base * self.s, idx, vals)
            return self.solver.walk(pts, scene, seed=100 + i)

        def parameters(self):
            return self.s.detach().reshape(1)

        def step(self):
            self.opt.step()
            self.opt.zero_grad()

    runner = TrainRunner(out_dir=str(tmp_path), model_target=ScaleModel(1.1), model=ScaleModel(1.0, lr=0.005),
                         niters=60, is_l1=False, is_mask=False)
    hist = runner.run()
    assert abs(runner.model.s.item() - 1.1) < 0.03, runner.model.s.item()
    assert hist["loss"][-1] < 0.3 * hist["loss"][0]
    assert (tmp_path / "loss.txt").exists() and (tmp_path / "target.exr").exists()
