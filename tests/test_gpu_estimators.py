"""The reference's remaining drop-in estimators against closed forms:
WoSWithSource.grad and normal_derivative on a manufactured Poisson
solution (the clover ablation's u = sin 5x + cos 5y, ablations/
test_clover.py:16-46), the 3D loop normal and tangential derivatives of a
harmonic quadratic, the torch triangle closest point against the native
query, a primal detector image against the exact solution, and the 2D / 3D
baseline integrators against the exact derivative of a uniform scaling.
Monte Carlo comparisons state their tolerance in standard errors (se) of
the sample mean plus a stated relative allowance for known biases."""
import math

import numpy as np
import pytest
import torch

from wos import configs as C

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


# u = sin 5x + cos 5y solves -Lap u = 25 sin 5x + 25 cos 5y, the source as
# two Fourier terms c sin(ax x + ay y + ph): (5, 0, 0, 25) and (0, 5, pi/2, 25)
def _u5(p):
    return np.sin(5 * p[:, 0]) + np.cos(5 * p[:, 1])


def _grad_u5(p):
    return np.stack([5 * np.cos(5 * p[:, 0]), -5 * np.sin(5 * p[:, 1])], 1)


def _clover_u5():
    from wos.scene import Polyline
    v, ix = C.star_polyline(4096)
    sp = np.array([[5, 0, 0, 25], [0, 5, np.pi / 2, 25]], np.float32)
    return Polyline(_t(v), _t(ix), _t(_u5(v).astype(np.float32)), source_type=3, source_params=_t(sp))


def _mean_se(x, groups):
    x = np.asarray(x, np.float64).reshape(groups, -1)
    return x.mean(1), x.std(1, ddof=1) / math.sqrt(x.shape[1])


def test_source_grad_against_manufactured_solution():
    """WoSWithSource.grad (wos/wos_with_source.py:109-152): boundary term on
    the first ball plus the source's interior term; the mean of 131,072
    estimates per point equals grad u within 4 se."""
    from wos.fwd import PCG32
    from wos.wos_with_source import WoSWithSource
    sc = _clover_u5()
    x = np.array([[0.3, -0.2], [-0.4, 0.25], [0.1, 0.6]], np.float32)
    N = 131072
    xx = np.repeat(x, N, axis=0)
    ws = WoSWithSource(nwalks=1, nsteps=256, epsilon=1e-4, use_IS_for_greens=False)
    gx, gy = ws.grad(_t(xx), sc, PCG32(xx.shape[0], 3, 5))
    ex = _grad_u5(x.astype(np.float64))
    for k, g in enumerate((gx, gy)):
        m, se = _mean_se(g.cpu().numpy(), 3)
        print(f"source grad component {k}: {m} +- {se}, exact {ex[:, k]}")
        assert np.all(np.abs(m - ex[:, k]) <= 4 * se + 0.01 * np.abs(ex).max()), (k, m, se, ex[:, k])


def test_source_normal_derivative_on_the_clover():
    """WoSWithSource.normal_derivative (wos/wos_with_source.py:154-232) at
    five boundary points of the clover: the mean of 65,536 estimates equals
    grad u . n within 4 se (the 2D angle clamp at 0.1 costs O(c^3))."""
    from wos.fwd import PCG32
    from wos.wos_with_source import WoSWithSource
    sc = _clover_u5()
    th = np.array([0.2, 1.0, 2.2, 3.9, 5.3])
    rb = 1 + 0.3 * np.cos(4 * th)
    q = np.stack([0.995 * rb * np.cos(th), 0.995 * rb * np.sin(th)], 1).astype(np.float32)
    N = 65536
    its = sc.closest_point(_t(np.repeat(q, N, axis=0)))
    ws = WoSWithSource(nwalks=1, nsteps=256, epsilon=1e-4, use_IS_for_greens=False)
    gx, gy = ws.normal_derivative(its, sc, PCG32(5 * N, 7, 9), clamping=0.1)
    n_in = its.n.cpu().numpy().astype(np.float64)     # points to the walk's side: inward
    est = -(gx.cpu().numpy() * n_in[:, 0] + gy.cpu().numpy() * n_in[:, 1])
    m, se = _mean_se(est, 5)
    pb = its.p.cpu().numpy().astype(np.float64)[::N]
    exact = -(_grad_u5(pb) * n_in[::N]).sum(1)
    print(f"source normal derivative: {m} +- {se}, exact {exact}")
    assert np.all(np.abs(m - exact) <= 4 * se + 0.02 * np.abs(exact).max()), (m, se, exact)


def _sphere5_records(n_per):
    from wos.scene3d import Scene3D
    v, f, noise = C.displaced_sphere(5)
    sc = Scene3D(_t(v), _t(f), _t(C.g_harmonic3(v).astype(np.float32)))
    rng = np.random.default_rng(0)
    dirs = rng.normal(size=(5, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    q = (dirs * (0.99 * (1 + 0.15 * noise(dirs)))[:, None]).astype(np.float32)
    its = sc.closest_point(_t(np.repeat(q, n_per, axis=0)))
    return sc, its


def _grad_h3(p):
    return np.stack([2 * p[:, 0], -2 * p[:, 1], np.ones(p.shape[0])], 1)


def test_3d_loop_normal_derivative_closed_form():
    """WoS3D.normal_derivative (wos/wos3d.py:97-135) for u = x^2 - y^2 + z on
    the displaced sphere: at clamp 1e-2 the mean of 16,384 estimates equals
    du/dn within 4 se + 6% (the clamp's deficit is sqrt(c) / (2 sqrt 2) =
    3.5%); at the reference drivers' clamp 0.1 the estimate is low by about
    that formula's 11.2% (the summed estimates over the summed exact values
    lie in [0.80, 0.95])."""
    from wos.fwd import PCG32
    from wos.wos3d import WoS3D
    N = 16384
    sc, its = _sphere5_records(N)
    w = WoS3D(nwalks=1, nsteps=256, epsilon=1e-4)
    n_in = its.n.cpu().numpy().astype(np.float64)
    pb = its.p.cpu().numpy().astype(np.float64)[::N]
    exact = -(_grad_h3(pb) * n_in[::N]).sum(1)
    res = {}
    for clamp in (1e-2, 1e-1):
        g = w.normal_derivative(its, sc, PCG32(5 * N, 1, 2), clamping=clamp)
        est = -sum(g[k].cpu().numpy() * n_in[:, k] for k in range(3))
        res[clamp] = _mean_se(est, 5)
        print(f"3D normal derivative, clamp {clamp}: {res[clamp][0]} +- {res[clamp][1]}, exact {exact}")
    m, se = res[1e-2]
    assert np.all(np.abs(m - exact) <= 4 * se + 0.06 * np.abs(exact).max()), (m, se, exact)
    m, _ = res[1e-1]
    ratio = m.sum() / exact.sum()
    assert 0.80 <= ratio <= 0.95, ratio


def test_3d_loop_tangent_derivative():
    """WoS3D.tangent_derivative (wos/wos3d.py:137-156): the in-plane
    gradient of the linearly interpolated values equals the tangential part
    of grad u to the interpolation error of a 0.04-edge mesh."""
    from wos.wos3d import WoS3D
    sc, its = _sphere5_records(1)
    tg = np.stack([x.cpu().numpy() for x in WoS3D(nwalks=1).tangent_derivative(its, sc)], 1)
    n = its.n.cpu().numpy().astype(np.float64)
    g = _grad_h3(its.p.cpu().numpy().astype(np.float64))
    exact = g - (g * n).sum(1, keepdims=True) * n
    print(f"tangent derivative error {np.abs(tg - exact).max():.4f}")
    assert np.abs(tg - exact).max() < 0.05
    assert np.abs((tg * n).sum(1)).max() < 1e-5     # in the triangle's plane


def test_torch_closest_point_triangle_matches_native():
    """wos/utils.py:108-173 (the torch closest point behind Scene3D's
    attached queries) against the native query on the same triangles, and
    its autograd: the distance's gradient in p is the unit vector from the
    closest point."""
    from wos.scene3d import Scene3D
    from wos.utils import closest_point_triangle
    v, f, noise = C.displaced_sphere(4)
    sc = Scene3D(_t(v), _t(f), _t(C.g_harmonic3(v).astype(np.float32)))
    rng = np.random.default_rng(2)
    surf = v[f[rng.integers(0, f.shape[0], 4000)]]
    bary = rng.dirichlet([1, 1, 1], 4000)
    near = (surf * bary[:, :, None]).sum(1) * (1 + rng.uniform(-0.02, 0.02, (4000, 1)))
    pts = np.concatenate([rng.uniform(-1.4, 1.4, (8000, 3)), near, v[:1000] * 1.01]).astype(np.float32)
    rec = sc.closest_point(_t(pts))
    fi = sc.indices[rec.prim_id.long()].long()
    a, b, c = sc.vertices[fi[:, 0]], sc.vertices[fi[:, 1]], sc.vertices[fi[:, 2]]
    p = _t(pts).requires_grad_(True)
    pt, uv, d = closest_point_triangle(p, a, b, c)
    assert float((pt - rec.p).abs().max()) < 1e-5
    assert float((uv - rec.uv).abs().max()) < 1e-4
    assert float((d - sc.sdf(_t(pts)).abs()).abs().max()) < 1e-5
    d.sum().backward()
    unit = (p.detach() - pt.detach()) / d.detach()[:, None]
    far = d.detach() > 1e-3
    assert float((p.grad[far] - unit[far]).abs().max()) < 1e-3


def test_walk_detector_image_against_exact():
    """Solver.walk_detector (wos/solver.py:67-84): a 64 x 64 primal image of
    the circle with g = x^2 - y^2 + 0.5x, 256 jittered walks per pixel.  The
    mean of g over a square pixel equals g at its centre (the quadratic
    terms' pixel means cancel), so the image's RMSE against g inside r < 0.9
    is the Monte Carlo error: within 1.2x the bound from the per-walk
    variance, with no bias beyond 3 se of the mean error; pixels outside the
    disk are 0."""
    from wos.scene import Detector, Polyline
    from wos.wos import WoSCUDA
    v, ix = C.circle_polyline(1024)
    sc = Polyline(_t(v), _t(ix), _t(C.g_circle(v).astype(np.float32)))
    det = Detector(vmin=(-1.0, -1.0), vmax=(1.0, 1.0), res=(64, 64))
    W = 256
    solver = WoSCUDA(nwalks=W, nsteps=256, epsilon=1e-3)
    img = solver.walk_detector(sc, det, seed=2).cpu().numpy()
    P = det.make_points()
    Pn = P.cpu().numpy().astype(np.float64)
    r = np.hypot(Pn[:, 0], Pn[:, 1])
    inner = r < 0.9
    err = img[inner] - C.g_circle(Pn[inner])
    var = solver.cwos.walk_reduce(P[_t(inner)], sc, seed=5, stats=True)["var64"].cpu().numpy()
    bound = math.sqrt(np.mean(var / W))
    rmse = math.sqrt(np.mean(err ** 2))
    print(f"detector image: rmse {rmse:.4f}, Monte Carlo bound {bound:.4f}, mean error {err.mean():+.5f}")
    assert rmse < 1.2 * bound
    assert abs(err.mean()) < 3 * bound / math.sqrt(err.size)
    np.testing.assert_array_equal(img[r > 1.05], 0.0)


def _scaling_2d():
    v, ix = C.star_polyline(1024)
    rng = np.random.default_rng(1)
    rad, ang = 0.55 * np.sqrt(rng.random(8192)), 2 * np.pi * rng.random(8192)
    p = np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)
    return v, ix, p, -4.0 * float(np.mean(C.g_saddle(p) ** 2))


def _adjoint_of_scaling(make_scene, base, pts, u, integrator, seeds):
    """d/ds of mean(u^2) for vertices s * base through a zero-valued
    integrator: sum_i 2 u_i / P du_i/ds, with u walked independently of the
    integrator's samples (so no Cov(u, du)/W term)."""
    out = []
    for s in seeds:
        sv = torch.tensor(1.0, device="cuda", requires_grad=True)
        r = integrator.walk(pts, make_scene(base * sv), seed=s)
        ((2.0 / pts.shape[0]) * u * r).sum().backward()
        out.append(sv.grad.item())
    return np.array(out)


@pytest.mark.parametrize("name", ["Baseline2", "Baseline2_normal_only", "Baseline2CUDA"])
def test_baselines_2d_scaling_closed_form(name):
    """wos/wos_grad.py:15-54, 89-110: the spatial gradient l = 5e-3 inside
    the exit point (all of it, or its normal part plus the tangential
    derivative) gives d/ds of mean u^2 for the star scaled by s with
    harmonic values g = x^2 - y^2, exact -4 mean g^2.  These are the
    reference's baselines, biased by evaluating the gradient off the
    boundary (measured 4-5% low here, against 0.04% for OursCUDA in the
    closed-form test of the configurations); within 3 se + 10% over four
    seeds, which still fails a wrong sign, scale or missing term."""
    from wos.scene import Polyline
    from wos.wos import WoSCUDA
    from wos.wos_grad import Baseline2, Baseline2CUDA
    v, ix, p, exact = _scaling_2d()
    vals, I, pts = _t(C.g_saddle(v).astype(np.float32)), _t(ix), _t(p.astype(np.float32))
    u = WoSCUDA(nwalks=1024, nsteps=256, epsilon=1e-4).walk(pts, Polyline(_t(v), I, vals), seed=99).detach()
    kw = dict(nwalks=256, nsteps=256, epsilon=1e-4)
    integ = {"Baseline2": lambda: Baseline2(**kw),
             "Baseline2_normal_only": lambda: Baseline2(normal_derivative_only=True, **kw),
             "Baseline2CUDA": lambda: Baseline2CUDA(**kw)}[name]()
    g = _adjoint_of_scaling(lambda V: Polyline(V, I, vals), _t(v), pts, u, integ, (11, 12, 13, 14))
    m, se = g.mean(), g.std(ddof=1) / 2.0
    print(f"{name}: {m:+.5e} +- {se:.1e}, exact {exact:+.5e}")
    assert abs(m - exact) <= 3 * se + 0.10 * abs(exact), (m, se, exact)


def test_baselines_3d_scaling_closed_form():
    """wos/wos_grad_3d.py:60-101 (Baseline23D, both modes) against d/ds of
    mean u^2 for the displaced sphere scaled by s with g = x^2 - y^2 + z,
    exact -2 mean(u (2x^2 - 2y^2 + z)), within 3 se + 12% (biased low as in
    2D: 7-8% measured); Baseline3D
    (wos/wos_grad_3d.py:23-57, soft-edge weights min(|d|, eps)/eps,
    differentiated through the torch closest point) runs with the same sign
    and within 50% of it at eps = 1e-2, its soft edge's bias; WoSGradient3D's
    native walk (wos/wos_grad_3d.py:10-20) solves the primal."""
    from wos.scene3d import Scene3D
    from wos.wos3d import WoS3DCUDA
    from wos.wos_grad_3d import Baseline3D, Baseline23D, WoSGradient3D
    v, f, noise = C.displaced_sphere(3)
    p = C.points_in_star(2048, noise, seed=3, shrink=0.8).astype(np.float64)
    uex = C.g_harmonic3(p)
    exact = -2.0 * np.mean(uex * (2 * p[:, 0] ** 2 - 2 * p[:, 1] ** 2 + p[:, 2]))
    G, F, pts = _t(C.g_harmonic3(v).astype(np.float32)), _t(f), _t(p.astype(np.float32))
    mk = lambda V: Scene3D(V, F, G)
    u = WoS3DCUDA(nwalks=1024, nsteps=256, epsilon=1e-3).walk(pts, mk(_t(v)), seed=99).detach()
    for nd in (False, True):
        integ = Baseline23D(nwalks=256, nsteps=256, epsilon=1e-3, normal_derivative_only=nd)
        g = _adjoint_of_scaling(mk, _t(v), pts, u, integ, (11, 12, 13, 14))
        m, se = g.mean(), g.std(ddof=1) / 2.0
        print(f"Baseline23D (normal only {nd}): {m:+.5e} +- {se:.1e}, exact {exact:+.5e}")
        assert abs(m - exact) <= 3 * se + 0.12 * abs(exact), (nd, m, se, exact)
    sub = slice(0, 256)
    g = _adjoint_of_scaling(mk, _t(v), pts[sub], u[sub], Baseline3D(nwalks=16, nsteps=64, epsilon=1e-2), (3,))[0]
    ex_sub = -2.0 * np.mean(uex[sub] * (2 * p[sub, 0] ** 2 - 2 * p[sub, 1] ** 2 + p[sub, 2]))
    print(f"Baseline3D: {g:+.5e}, exact {ex_sub:+.5e}")
    assert np.isfinite(g) and 0.5 * abs(ex_sub) <= abs(g) <= 1.5 * abs(ex_sub) and g * ex_sub > 0
    uw = WoSGradient3D(nwalks=64, nsteps=256, epsilon=1e-3).wos.walk(pts, mk(_t(v)), seed=1).cpu().numpy()
    assert math.sqrt(np.mean((uw - uex) ** 2)) < 0.1
