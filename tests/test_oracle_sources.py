"""Pins of the oracle's source terms and interior sampling against what the
mathematics fixes (CPU only).

* Sinusoid source (include/scene.cuh:278-301, type 2): f = a sin(Ax + By +
  C) + o has the particular solution a sin(.)/(A^2 + B^2) - o (x^2 + y^2)/4 of
  -Lap u = f, so with matching boundary data the walk's mean is that
  solution plus a harmonic part.
* Gaussian source (type 1): at the centre of the unit disk with zero boundary
  data u(0) = integral of G(0, y) f(y) over the disk, G(0, y) = ln(1/|y|)/2pi,
  evaluated by FP64 quadrature.
* Importance-sampled interior term (include/wos.cuh:49-88): the table is the
  normalised integral of r ln(1/r); the bin search is the first entry above
  the draw; the estimate of the integral of G over the disk is R^2/4, the
  closed form, and the sampled radii follow the density r ln(R/r).
"""
import numpy as np
import pytest
from scipy import integrate

import oracle as O
from wos import configs as C


def _disk_scene(n, values_fn, source_type, params):
    v, ix = C.circle_polyline(n)
    return O.Scene(v, ix, values_fn(v.astype(np.float64)).astype(np.float32), source_type, params), v


SIN = np.array([2.0, -1.0, 0.3, 1.5, 1.0], np.float32)   # A, B, C, amplitude, offset


def _u_sin(p):
    A, B, Cc, a, o = SIN.astype(np.float64)
    p = np.asarray(p, np.float64)
    return (a * np.sin(A * p[:, 0] + B * p[:, 1] + Cc) / (A * A + B * B) - o * (p[:, 0] ** 2 + p[:, 1] ** 2) / 4
            + p[:, 0] ** 2 - p[:, 1] ** 2)


@pytest.mark.parametrize("use_is", [False, True])
def test_sinusoid_source_manufactured_solution(use_is):
    s, _ = _disk_scene(512, _u_sin, 2, SIN)
    rng = np.random.default_rng(0)
    r, t = 0.8 * np.sqrt(rng.random(48)), 2 * np.pi * rng.random(48)
    p = np.stack([r * np.cos(t), r * np.sin(t)], 1).astype(np.float32)
    W = 4000
    out = s.walk_reduce(p, W, 256, 1e-4, seed=3, use_is=use_is)
    se = np.sqrt(out["var"] / W)
    err = out["mean"] - _u_sin(p)
    assert np.all(np.abs(err) < 4.5 * se + 2e-4), (np.abs(err) / se).max()
    # the offset term matters: without it the same estimates are far off
    wrong = _u_sin(p) + SIN[4] * (p[:, 0] ** 2 + p[:, 1] ** 2) / 4
    assert np.mean(np.abs(out["mean"] - wrong) > 4.5 * se) > 0.5


GAUSS = np.array([0.2, -0.1, 0.3, 0.25, 2.0], np.float32)   # mu_x, mu_y, sigma_x, sigma_y, A


def _gauss_center_value():
    mx, my, sx, sy, A = GAUSS.astype(np.float64)

    def f(r, th):
        x, y = r * np.cos(th), r * np.sin(th)
        return A * np.exp(-(x - mx) ** 2 / (2 * sx * sx) - (y - my) ** 2 / (2 * sy * sy)) * np.log(1 / r) / (
            2 * np.pi) * r
    val, err = integrate.dblquad(f, 0, 2 * np.pi, 1e-12, 1.0, epsabs=1e-10, epsrel=1e-10)
    return val


@pytest.mark.parametrize("use_is", [False, True])
def test_gaussian_source_center_value_quadrature(use_is):
    s, _ = _disk_scene(1024, lambda v: np.zeros(v.shape[0]), 1, GAUSS)
    p = np.zeros((64, 2), np.float32)   # 64 copies of the centre, independent streams per copy
    W = 2000
    out = s.walk_reduce(p, W, 256, 1e-4, seed=1, use_is=use_is)
    m = out["mean"].mean()
    se = np.sqrt(out["var"].mean() / (W * 64))
    exact = _gauss_center_value()
    assert abs(m - exact) < 4 * se + 2e-4 * abs(exact), (m, exact, se)
    # a sign error in the Gaussian's exponent or a dropped amplitude fails
    assert abs(m - exact / GAUSS[4]) > 4 * se


def test_green_cdf_table_is_the_normalised_integral():
    cdf = O.green_cdf()
    n = cdf.size
    assert n == 10000
    total = integrate.quad(lambda r: r * np.log(1 / r), 0, 1)[0]
    assert abs(total - 0.25) < 1e-12
    for i in (0, 1, 9, 99, 1234, 5000, 9000, 9998, 9999):
        x = (i + 1) / n
        num = integrate.quad(lambda r: r * np.log(1 / r), 0, x, epsabs=1e-14)[0] / total
        assert abs(cdf[i] - num) < 1e-7 * max(1.0, num) + 2e-9, (i, cdf[i], num)
    assert np.all(np.diff(cdf) >= 0) and cdf[-1] == 1.0


def test_green_bin_search_is_first_entry_above_the_draw():
    cdf = O.green_cdf()
    rng = np.random.default_rng(5)
    # draws are next_1d values in [0, 1): table entries below 1 (exact bin edges), 0 and the largest draw
    edges = cdf[[0, 1, 17, 500, 9000, 9997]]
    assert (edges < 1.0).all()
    u = np.concatenate([rng.random(20000).astype(np.float32), edges,
                        np.array([0.0, 1.0 - 2.0 ** -23], np.float32)])
    rad, pdf, b = O.sample_green_radius(u, 1.0)
    ref = np.minimum(np.searchsorted(cdf, u, side="right"), cdf.size - 1)
    np.testing.assert_array_equal(b, ref)
    # the draw is reused inside its bin: the radius lies in [l, l+1] / N
    assert np.all(rad >= b / cdf.size - 1e-6) and np.all(rad <= (b + 1) / cdf.size + 1e-6)
    # radial pdf = pmf * N / R, pmf the bin's probability
    lo = np.where(b > 0, cdf[np.maximum(b - 1, 0)], 0.0)
    np.testing.assert_allclose(pdf, (cdf[b] - lo) * cdf.size, rtol=1e-5)


@pytest.mark.parametrize("use_is", [False, True])
def test_interior_estimate_of_integral_of_G_is_R2_over_4(use_is):
    """For f = 1 the interior term estimates the integral of G over the disk,
    R^2 / 4.  Importance sampling has almost no variance here (G r / pdf is
    constant within a bin), so its mean agrees to the table's resolution."""
    R = 0.7
    n = 1 << 16
    g = np.random.default_rng(7)
    # stratified draws over [0,1)^2
    k = np.arange(n)
    u0 = ((k + g.random(n)) / n).astype(np.float32)
    u1 = g.random(n).astype(np.float32)
    dx, dy, G, w, take = O.interior_sample(u0, u1, R, use_is)
    est = np.where(take, G.astype(np.float64) * w, 0.0)
    if use_is:
        assert abs(est.mean() - R * R / 4) < 1e-5 * R * R
        # offsets lie in the disk, with radius from the table and angle 2 pi u1
        r = np.hypot(dx, dy)
        assert r.max() <= R * (1 + 1e-6)
        ang = np.mod(np.arctan2(dy, dx) / (2 * np.pi), 1.0)
        ok = r > 1e-6
        assert np.abs(np.mod(ang[ok] - u1[ok] + 0.5, 1.0) - 0.5).max() < 1e-5
    else:
        assert abs(est.mean() - R * R / 4) < 4 * est.std() / np.sqrt(n)


def test_importance_sampled_radii_follow_r_log_R_over_r():
    R = 1.0
    u = np.random.default_rng(11).random(100000).astype(np.float32)
    rad, _, _ = O.sample_green_radius(u, R)
    # CDF of the density r log(R/r) / (R^2 / 4) is x^2 (1 - 2 ln x), x = r/R: KS distance
    x = np.sort(rad.astype(np.float64) / R)
    F = x * x * (1 - 2 * np.log(np.maximum(x, 1e-300)))
    emp = np.arange(1, x.size + 1) / x.size
    assert np.abs(emp - F).max() < 1.63 / np.sqrt(x.size)   # KS at the 1% level
    # the uniform-area density (x^2) is rejected by the same test
    assert np.abs(emp - x * x).max() > 0.05


def test_fourier_basis_sums_are_the_coefficient_derivative():
    """The per-point basis sums (64-bit fixed point, units of 2^-e) are the
    derivative of W times the point's mean with respect to each Fourier
    coefficient: the walk is linear in the coefficients under common random
    numbers, so central differences of the mean reproduce basis / W."""
    c = C.config2(res=24)
    sp = c["source_params"].astype(np.float32)
    pts = c["points"][::3]
    W, h = 16, 1e-2
    sc = O.Scene(c["vertices"], c["indices"], c["values"], 3, sp)
    r = sc.walk_reduce(pts, W, 256, 1e-4, seed=3, basis=True)
    assert r["basis"].shape == (pts.shape[0], sp.shape[0])
    for k in (0, 3, 7):
        up, dn = sp.copy(), sp.copy()
        up[k, 3] += h
        dn[k, 3] -= h
        mu = O.Scene(c["vertices"], c["indices"], c["values"], 3, up).walk_reduce(pts, W, 256, 1e-4, seed=3)["mean"]
        md = O.Scene(c["vertices"], c["indices"], c["values"], 3, dn).walk_reduce(pts, W, 256, 1e-4, seed=3)["mean"]
        fd = (mu - md) / (2 * h)
        np.testing.assert_allclose(r["basis"][:, k] / W, fd, rtol=1e-3, atol=2e-5)


def test_basis_exponent_bounds_the_fixed_point_sums():
    """2^e times the bound on a point's basis sum (nwalks walks of at most
    nsteps samples, each below 16 R^2 for R = 178.5) stays below 2^61, and
    e is the largest such exponent."""
    for W, n in ((1, 1), (16, 256), (128, 256), (4096, 1024), (1 << 20, 4096)):
        e = O.lib.or_basis_exponent(W, n)
        bound = W * n * 178.5 ** 2 * 16.0
        assert bound * 2.0 ** e < 2.0 ** 61 <= bound * 2.0 ** (e + 1)
