"""Forward-mode shape derivatives (the reference's renderD, which seeds a
scene parameter's tangent and reads dr.grad of the image: diff_solve/
runner.py:75-109): wos_shape_derivative against the reverse-mode shape
gradient on the same samples (an exact identity up to float rounding), the
torch forward-AD plumbing through the shape solvers, and a central
difference of the primal."""
import numpy as np
import pytest
import torch
import torch.autograd.forward_ad as fwAD

from wos import configs as C

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _scene(dim):
    import wos_ext
    if dim == 2:
        v, ix = C.star_polyline(1024)
        sc = wos_ext.Scene(_t(v), _t(ix), _t(C.g_saddle(v).astype(np.float32)), use_bvh=True)
        rng = np.random.default_rng(5)
        r, t = 0.55 * np.sqrt(rng.random(1500)), 2 * np.pi * rng.random(1500)
        pts = np.stack([r * np.cos(t), r * np.sin(t)], 1).astype(np.float32)
    else:
        v, f = C.icosphere(3)
        v, f = v.astype(np.float32), f.astype(np.int32)
        sc = wos_ext.Scene3D(_t(v), _t(f), _t(C.g_harmonic3(v).astype(np.float32)), use_bvh=True)
        rng = np.random.default_rng(6)
        d = rng.normal(size=(800, 3))
        pts = (0.6 * rng.random((800, 1)) ** (1 / 3) * d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return sc, v, pts


@pytest.mark.parametrize("dim", [2, 3])
def test_shape_derivative_matches_reverse_mode(dim):
    """sum_i g_i du_i/dV = grad_vertices(g) . V for the same seed: both sum
    -grad u . V(exit) over the same samples."""
    import wos_ext
    sc, v, pts = _scene(dim)
    rng = np.random.default_rng(dim)
    g = rng.normal(size=pts.shape[0]).astype(np.float32)
    V = rng.normal(size=v.shape).astype(np.float32)
    kw = dict(nwalks=32, nsteps=128, epsilon=1e-3, epsilon2=1e-3, clamping=0.1)
    mean, m64, var64 = wos_ext.shape_derivative(sc, _t(pts), 17, _t(V), stats=True, **kw)
    gv, _ = wos_ext.shape_gradient(sc, _t(pts), 17, _t(g), **kw)
    fwd = float((g.astype(np.float64) * m64.cpu().numpy()).sum())
    rev = float((gv.cpu().numpy().astype(np.float64) * V).sum())
    scale = float((np.abs(g) * np.abs(m64.cpu().numpy())).sum())
    assert np.isfinite(fwd) and scale > 0
    assert abs(fwd - rev) <= 1e-4 * scale, (fwd, rev, scale)
    np.testing.assert_allclose(mean.cpu().numpy(), m64.cpu().numpy(), rtol=1e-5, atol=1e-6)
    assert (var64.cpu().numpy() >= 0).all()


def test_forward_ad_through_the_shape_solver_and_central_difference():
    """renderD in torch: a dual deformation parameter s (uniform scaling of
    the star, vertices = s * base) through OursCUDA gives per-point tangents
    equal to wos_shape_derivative with V = base; d/ds of the mean square of
    u then agrees with the closed form -4 mean g^2 (g = x^2 - y^2 is
    harmonic on the plane, so u_s(x) = g(x / s)) and with a central
    difference of the primal (common seeds).  The value u that multiplies
    the tangent is walked with another seed: with the same seed their
    product also carries Cov(u, du)/W."""
    import wos_ext
    from wos.scene import Polyline
    from wos.wos import WoSCUDA
    from wos.wos_grad import OursCUDA
    v, ix = C.star_polyline(1024)
    vals = _t(C.g_saddle(v).astype(np.float32))
    rng = np.random.default_rng(1)
    r, t = 0.55 * np.sqrt(rng.random(2048)), 2 * np.pi * rng.random(2048)
    pts = _t(np.stack([r * np.cos(t), r * np.sin(t)], 1).astype(np.float32))
    base = _t(v)
    ours = OursCUDA(nwalks=256, nsteps=256, epsilon=1e-4, epsilon2=1e-4, clamping=0.1)
    with fwAD.dual_level():
        s = fwAD.make_dual(torch.tensor(1.0, device="cuda"), torch.tensor(1.0, device="cuda"))
        scene = Polyline(base * s, _t(ix), vals)
        out = ours.walk(pts, scene, seed=11)
        tangent = fwAD.unpack_dual(out).tangent
    assert tangent is not None and tangent.shape == (pts.shape[0],)
    direct = wos_ext.shape_derivative(Polyline(base, _t(ix), vals), pts, 11, base, nwalks=256, nsteps=256,
                                      epsilon=1e-4, epsilon2=1e-4, clamping=0.1)
    np.testing.assert_array_equal(tangent.cpu().numpy(), direct.cpu().numpy())

    fwd = WoSCUDA(nwalks=1024, nsteps=256, epsilon=1e-4)
    u = fwd.walk(pts, Polyline(base, _t(ix), vals), seed=12)
    dL = float((2.0 * u.double() * tangent.double()).mean())
    p64 = pts.double().cpu().numpy()
    exact = -4.0 * float(np.mean(C.g_saddle(p64) ** 2))
    h = 2e-2
    L = lambda sc: float((fwd.walk(pts, Polyline(base * sc, _t(ix), vals), seed=11).double() ** 2).mean())
    fd = (L(1 + h) - L(1 - h)) / (2 * h)
    print(f"forward mode: dL/ds {dL:+.5e}, closed form {exact:+.5e}, central difference {fd:+.5e}")
    assert abs(dL - exact) < 0.04 * abs(exact), (dL, exact)
    assert abs(fd - exact) < 0.03 * abs(exact), (fd, exact)


def test_derivative_image_on_a_detector():
    """walk_detector with a shape solver under forward AD: one derivative
    per pixel (nwalks jittered walks each), finite, and zero outside."""
    from wos.scene import Detector, Polyline
    from wos.wos_grad import OursCUDA
    v, ix = C.star_polyline(512)
    base = _t(v)
    det = Detector(vmin=(-1.5, -1.5), vmax=(1.5, 1.5), res=(32, 32))
    ours = OursCUDA(nwalks=16, nsteps=128, epsilon=1e-3, epsilon2=1e-3, clamping=0.1)
    with fwAD.dual_level():
        s = fwAD.make_dual(torch.tensor(0.0, device="cuda"), torch.tensor(1.0, device="cuda"))
        shift = torch.stack([torch.zeros_like(s), s])
        scene = Polyline(base + shift, _t(ix), _t(C.g_saddle(v).astype(np.float32)))
        img = ours.walk_detector(scene, det, seed=3)
        d_img = fwAD.unpack_dual(img).tangent
    assert d_img.shape == (32 * 32,)
    d = d_img.cpu().numpy()
    assert np.isfinite(d).all() and np.abs(d).max() > 0
    corner = d.reshape(32, 32)[[0, 0, -1, -1], [0, -1, 0, -1]]
    np.testing.assert_array_equal(corner, 0.0)


def test_runner_render_tasks(tmp_path):
    """TestRunner (diff_solve/runner.py): the star moved up by delta.
    renderD (OursCUDA + WoSBoundary, forward mode) correlates with renderFD
    (difference of two primal images under common seeds) over the interior,
    and every task writes its EXR image."""
    from dataclasses import dataclass
    from wos.io import read_exr
    from wos.runner import RenderC, RenderD, RenderFD, TestRunner
    from wos.scene import Detector, Polyline
    from wos.wos import WoSCUDA
    from wos.wos_boundary import WoSBoundary
    from wos.wos_grad import OursCUDA
    v, ix = C.star_polyline(1024)
    base, I, vals = _t(v), _t(ix), _t(C.g_saddle(v).astype(np.float32))

    @dataclass
    class Star(TestRunner):
        def make_scene(self, delta=0.0):
            if not isinstance(delta, torch.Tensor):
                delta = torch.tensor(float(delta), device="cuda")
            return Polyline(base + torch.stack([torch.zeros_like(delta), delta]), I, vals)

    r = Star(detector=Detector(vmin=(-1.2, -1.2), vmax=(1.2, 1.2), res=(64, 64)), npasses=4, out_dir=str(tmp_path),
             delta=2e-3, wos_boundary=WoSBoundary(nwalks=16),
             tasks=[RenderC(WoSCUDA(nwalks=256, nsteps=256, epsilon=1e-4)),
                    RenderD(OursCUDA(nwalks=256, nsteps=256, epsilon=1e-4, epsilon2=1e-4, clamping=0.1)),
                    RenderFD(WoSCUDA(nwalks=256, nsteps=256, epsilon=1e-4, prevent_fd_artifacts=True))])
    c, d, fd = r.run()
    for name in ("renderC.exr", "renderD.exr", "renderFD.exr"):
        assert read_exr(str(tmp_path / name)).shape == (64, 64, 1)
    inner = (r.make_scene().sdf(r.p).cpu().numpy() < -0.05)
    a, b = d.reshape(-1)[inner], fd.reshape(-1)[inner]
    assert np.isfinite(a).all() and np.isfinite(c).all()
    assert np.corrcoef(a, b)[0, 1] > 0.4
