"""CPU pins of the oracle: golden RNG vectors, elementary-function error
bounds, closed forms of the geometry, an independent top-down restatement of
the Karras hierarchy, and estimator checks against exact harmonic solutions.
Each check is chosen so that a plausible slip (wrong shift, dropped term,
swapped barycentric, sign error) fails it."""
import json
import math
import os

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

import oracle as O
from wos import configs as C

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rng_vectors.json")))


# ---- random numbers --------------------------------------------------------
def test_pcg32_published_demo_sequence():
    got = O.pcg32_stream(0, 42, 54, 6)
    assert [hex(x) for x in got] == GOLD["pcg32_demo_42_54"]


def test_pcg32_golden_streams():
    for c in GOLD["pcg32"]:
        got = O.pcg32_stream(c["index"], c["initstate"], c["initseq"], len(c["out"]))
        assert got.tolist() == c["out"]


def test_next_1d_and_next_f64_from_raw_words():
    r = O.Pcg32.seeded(3, 11, 17)
    raw = O.pcg32_stream(3, 11, 17, 6)
    f = [r.next_1d() for _ in range(2)]
    for k in range(2):
        assert f[k] == np.float32((int(raw[k]) >> 9) * 2.0 ** -23)
    d = r.next_f64()
    word = int(raw[2]) | (int(raw[3]) << 32)
    assert d == (word >> 12) * 2.0 ** -52


def test_tea_golden():
    for c in GOLD["tea"]:
        assert list(O.tea(c["v0"], c["v1"])) == c["out"]


@settings(max_examples=200, deadline=None)
@given(st.integers(-2 ** 31, 2 ** 31 - 1), st.integers(-2 ** 31, 2 ** 31 - 1))
def test_tea_matches_bigint_model(a, b):
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(os.path.dirname(__file__), "golden",
                                                                     "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    assert O.tea(a, b) == mg.tea(a, b)


# ---- elementary functions ------------------------------------------------------
def test_elementary_functions_error_bounds():
    rng = np.random.default_rng(0)
    u = rng.random(100000).astype(np.float32)
    ref = 2 * np.pi * u.astype(np.float64)
    assert np.abs(O.elementary("sin_turns", u) - np.sin(ref)).max() < 2e-7
    assert np.abs(O.elementary("cos_turns", u) - np.cos(ref)).max() < 2e-7
    # exact quarter turns
    q = np.array([0, 0.25, 0.5, 0.75], np.float32)
    np.testing.assert_array_equal(O.elementary("sin_turns", q), [0, 1, 0, -1])
    np.testing.assert_array_equal(O.elementary("cos_turns", q), [1, 0, -1, 0])
    x = rng.uniform(-40, 40, 100000).astype(np.float32)
    assert np.abs(O.elementary("sin", x) - np.sin(x.astype(np.float64))).max() < 3e-7
    y = np.exp(rng.uniform(-15, 15, 100000)).astype(np.float32)
    ly = np.log(y.astype(np.float64))
    assert (np.abs(O.elementary("log", y) - ly) <= 2e-7 * np.maximum(1, np.abs(ly))).all()
    z = rng.uniform(-80, 80, 100000).astype(np.float32)
    assert np.abs(O.elementary("exp", z) / np.exp(z.astype(np.float64)) - 1).max() < 3e-7


# ---- geometry ----------------------------------------------------------------
def _square():
    v = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    ix = np.array([[0, 1], [1, 2], [2, 3], [3, 0]], np.int32)
    return O.Scene(v, ix, np.array([0, 1, 2, 3], np.float32))


def test_segment_distance_closed_form():
    s = _square()
    rng = np.random.default_rng(1)
    p = rng.uniform(-0.5, 1.5, size=(5000, 2)).astype(np.float32)
    idx, d2 = s.closest_index(p)
    # distance to the unit square's boundary, float64
    x, y = p[:, 0].astype(np.float64), p[:, 1].astype(np.float64)
    inside = (x > 0) & (x < 1) & (y > 0) & (y < 1)
    din = np.minimum.reduce([x, 1 - x, y, 1 - y])
    dx = np.maximum.reduce([-x, np.zeros_like(x), x - 1])
    dy = np.maximum.reduce([-y, np.zeros_like(y), y - 1])
    dout = np.hypot(dx, dy)
    d = np.where(inside, din, dout)
    np.testing.assert_allclose(np.sqrt(d2), d, atol=2e-7)
    sd = s.sdf(p)
    assert (sd[inside] < 0).all() and (sd[~inside & (d > 1e-6)] > 0).all()


def test_closest_index_tie_rule_smallest_index():
    s = _square()
    # the centre is equidistant from all four sides: index 0 wins
    idx, _ = s.closest_index(np.array([[0.5, 0.5]], np.float32))
    assert idx[0] == 0
    # on the diagonal near (1,1): sides 1 and 2 tie -> 1
    idx, _ = s.closest_index(np.array([[0.75, 0.75]], np.float32))
    assert idx[0] == 1


def test_closest_index_matches_float64_bruteforce():
    v, ix = C.star_polyline(257)
    s = O.Scene(v, ix, np.zeros(257, np.float32))
    p = np.random.default_rng(2).uniform(-1.5, 1.5, size=(3000, 2)).astype(np.float32)
    idx, d2 = s.closest_index(p)
    a = v[ix[:, 0]].astype(np.float64)
    b = v[ix[:, 1]].astype(np.float64)
    pa = p[:, None, :] - a[None]
    ba = (b - a)[None]
    h = np.clip((pa * ba).sum(-1) / (ba * ba).sum(-1), 0, 1)
    dd = ((pa - ba * h[..., None]) ** 2).sum(-1)
    best = dd.min(axis=1)
    # the chosen segment is optimal up to float32 rounding
    assert (dd[np.arange(len(p)), idx] <= best * (1 + 1e-5) + 1e-12).all()
    np.testing.assert_allclose(d2, best, rtol=1e-5, atol=1e-12)


def test_triangle_closest_point_against_dense_sampling():
    rng = np.random.default_rng(3)
    tri = np.array([[0, 0, 0], [1, 0, 0], [0.2, 0.9, 0.1]], np.float32)
    s = O.Scene(tri, np.array([[0, 1, 2]], np.int32), np.array([1.0, 2.0, 3.0], np.float32))
    p = rng.uniform(-1, 1.5, size=(400, 3)).astype(np.float32)
    r = s.closest_point(p)
    n = 300
    uu, vv = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n))
    m = uu + vv <= 1
    bary = np.stack([1 - uu[m] - vv[m], uu[m], vv[m]], 1)
    pts = bary @ tri.astype(np.float64)
    dmin = np.sqrt(((p[:, None, :] - pts[None]) ** 2).sum(-1)).min(1)
    d = np.linalg.norm(p - r["p"], axis=1)
    assert (d <= dmin + 1e-6).all()
    assert (dmin - d < 5e-3).all()
    # record point = interpolate(a, b, c, uv) with uv = weights of (b, c)
    u, w = r["t"][:, 0], r["t"][:, 1]
    q = tri[1] * u[:, None] + tri[2] * w[:, None] + tri[0] * (1 - u - w)[:, None]
    np.testing.assert_allclose(q, r["p"], atol=1e-6)
    # Dirichlet value of the record is the same interpolation of the values:
    # for vertex values of a linear function the interpolant is that function
    # itself, so the value at the record equals the function at the record point
    lin = lambda x: 1.0 + 2.0 * x[:, 0] - 3.0 * x[:, 1] + 0.5 * x[:, 2]
    sl = O.Scene(tri, np.array([[0, 1, 2]], np.int32), lin(tri.astype(np.float64)).astype(np.float32))
    val = sl.dirichlet(r["prim"], r["t"], r["valid"])
    np.testing.assert_allclose(val, lin(r["p"].astype(np.float64)), atol=2e-6)
    # a wrong weight order (uv as the weights of (a, b)) would fail the same check
    swapped = sl.dirichlet(r["prim"], r["t"][:, ::-1].copy(), r["valid"])
    assert np.abs(swapped - lin(r["p"].astype(np.float64))).max() > 1e-2
    # invalid records and out-of-range primitives give 0
    assert (sl.dirichlet(np.array([0, -1, 5]), np.zeros((3, 2)), np.array([0, 1, 1])) == 0).all()


def test_dirichlet_2d_linear_values_exact():
    """2D records: t is the segment parameter, value = lerp of the end values,
    which for linear vertex data equals the function at the record point."""
    v, ix = C.circle_polyline(64)
    lin = lambda x: 0.3 - 1.5 * x[:, 0] + 0.25 * x[:, 1]
    s = O.Scene(v, ix, lin(v.astype(np.float64)).astype(np.float32))
    p = np.random.default_rng(4).uniform(-0.9, 0.9, size=(500, 2)).astype(np.float32)
    r = s.closest_point(p)
    val = s.dirichlet(r["prim"], r["t"], r["valid"])
    np.testing.assert_allclose(val, lin(r["p"].astype(np.float64)), atol=2e-6)


def test_largest_inscribed_ball_circle_closed_form():
    v, ix = C.circle_polyline(4096)
    s = O.Scene(v, ix, np.zeros(4096, np.float32))
    p = np.array([[1.0, 0.0], [0.0, -1.0]], np.float32)
    nrm = -p
    R = s.largest_inscribed_ball(np.ones(2, np.int32), p, nrm, 1e-3)
    # every m up to the radius passes the test, so 10 bisections from
    # [1e-3, 1] end at 1 - (1 - 1e-3) / 2^10
    np.testing.assert_allclose(R, 1 - (1 - 1e-3) / 1024, rtol=1e-5)


def test_morton_codes_hand_values():
    assert O.morton(2, np.array([0, 0], np.float32)) == 0
    assert O.morton(2, np.array([1, 1], np.float32)) == 0xFFFFFFFF
    assert O.morton(2, np.array([0.5, 0], np.float32)) == 0x80000000 >> 1   # x MSB -> bit 30
    assert O.morton(3, np.array([1, 1, 1], np.float32)) == 0x3FFFFFFF
    assert O.morton(3, np.array([0.5, 0, 0], np.float32)) == 1 << 29


def _topdown(keys):
    """Binary radix tree of sorted unique keys, split where the highest
    differing bit changes (the tree Karras's construction emits)."""
    ranges = []

    def rec(lo, hi):
        if lo == hi:
            return
        bit = (int(keys[lo]) ^ int(keys[hi])).bit_length()
        s = lo
        while s + 1 <= hi and (int(keys[lo]) ^ int(keys[s + 1])).bit_length() < bit:
            s += 1
        ranges.append(((lo, hi), (lo, s), (s + 1, hi)))
        rec(lo, s)
        rec(s + 1, hi)

    rec(0, len(keys) - 1)
    return ranges


@pytest.mark.parametrize("n,seed", [(2, 0), (3, 1), (57, 2), (1000, 3)])
def test_lbvh_matches_topdown_radix_tree(n, seed):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.random(n)) * 2 * np.pi
    v = np.stack([np.cos(t) * (1 + 0.2 * rng.random(n)), np.sin(t)], 1).astype(np.float32)
    ix = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1).astype(np.int32)
    if n == 2:
        ix = ix[:1]
        n = 1
        return
    s = O.Scene(v, ix, np.zeros(len(v), np.float32))
    b = s.lbvh()
    keys = [(int(c) << 32) | j for j, c in enumerate(b["codes"])]   # code then sorted position
    assert all(keys[i] < keys[i + 1] for i in range(len(keys) - 1))
    # range of every Karras node
    N = len(keys)
    rng_of = {}

    def span(c):
        if c < 0:
            return (~c, ~c)
        if c not in rng_of:
            l, r = b["children"][c]
            rng_of[c] = (span(l)[0], span(r)[1])
        return rng_of[c]

    got = sorted((span(i), span(b["children"][i][0]), span(b["children"][i][1])) for i in range(N - 1))
    assert got == sorted(_topdown(keys))
    # each primitive once, parents consistent
    assert sorted(b["prim"].tolist()) == list(range(N))
    for i, (l, r) in enumerate(b["children"]):
        for c in (l, r):
            if c < 0:
                assert b["parent_leaf"][~c] == i
            else:
                assert b["parent_internal"][c] == i


# ---- estimators against exact solutions -----------------------------------------
def _z(mean, var, exact, W):
    return np.abs(mean - exact) / np.sqrt(var / W + 1e-12)


def test_walk_circle_harmonic_exact():
    c = C.config1()
    s = O.Scene(c["vertices"], c["indices"], c["values"])
    p = c["points"][::120]
    r = s.walk_reduce(p, 64, 256, 1e-3, seed=0)
    z = _z(r["mean"], r["var"], c["exact"][::120], 64)
    assert np.mean(z > 3.5) < 0.03


def test_walk_disk_cos4theta():
    """ablations/test_disk.py:23-33: g = cos(4 theta) on the unit circle has
    u = r^4 cos(4 theta) inside."""
    v, ix = C.circle_polyline(2048)
    th = np.arctan2(v[:, 1], v[:, 0])
    s = O.Scene(v, ix, np.cos(4 * th).astype(np.float32))
    rng = np.random.default_rng(4)
    r_, t_ = np.sqrt(rng.random(60)) * 0.9, rng.random(60) * 2 * np.pi
    p = np.stack([r_ * np.cos(t_), r_ * np.sin(t_)], 1).astype(np.float32)
    out = s.walk_reduce(p, 256, 256, 1e-4, seed=1)
    exact = r_ ** 4 * np.cos(4 * t_)
    z = _z(out["mean"], out["var"], exact, 256)
    assert np.mean(z > 3.5) < 0.05


def test_walk_poisson_manufactured_twin():
    """-Lap u = A sin3x cos2y with g = x^2 - y^2 + A sin3x cos2y / 13 has
    u = g; A = 20 makes the interior term large against the MC noise."""
    A = 20.0
    c = C.config2(res=48)
    v = c["vertices"]
    vals = (C.g_saddle(v) + A * C.f_clover_source(v) / 13).astype(np.float32)
    sp = C.fourier_params(A * C.FOURIER_COEFFS)
    s = O.Scene(v, c["indices"], vals, 3, sp)
    p = c["points"][::25]
    exact = C.g_saddle(p) + A * C.f_clover_source(p) / 13
    out = s.walk_reduce(p, 128, 256, 1e-4, seed=2)
    z = _z(out["mean"], out["var"], exact, 128)
    assert np.mean(z > 3.5) < 0.05
    # dropping the interior term biases the estimate
    s0 = O.Scene(v, c["indices"], vals)
    o0 = s0.walk_reduce(p, 128, 256, 1e-4, seed=2)
    assert np.mean(_z(o0["mean"], o0["var"], exact, 128) > 3.5) > 0.3


def test_walk_outside_and_no_steps():
    c = C.config1()
    s = O.Scene(c["vertices"], c["indices"], c["values"])
    out = s.walk_reduce(np.array([[2.0, 0.0]], np.float32), 8, 64, 1e-3)
    assert out["mean"][0] == 0.0
    out = s.walk_reduce(np.array([[0.1, 0.2]], np.float32), 8, 0, 1e-3)
    assert out["mean"][0] == 0.0


def test_adjoint_values_is_exact_derivative():
    c = C.config2(res=24)
    s = O.Scene(c["vertices"], c["indices"], c["values"])
    p = c["points"]
    W = 8
    out = s.walk_reduce(p, W, 256, 1e-4, seed=3, records=True)
    g = np.random.default_rng(5).normal(size=len(p))
    adj = s.adjoint_values(len(p), W, out["rec_prim"], out["rec_t"], g)
    for v in (0, 1000, 2047, 3000):
        h = 0.25
        vals = c["values"].astype(np.float64).copy()
        vals[v] += h
        up = O.Scene(c["vertices"], c["indices"], vals).walk_reduce(p, W, 256, 1e-4, seed=3)["mean"]
        vals[v] -= 2 * h
        dn = O.Scene(c["vertices"], c["indices"], vals).walk_reduce(p, W, 256, 1e-4, seed=3)["mean"]
        assert abs(g @ (up - dn) / (2 * h) - adj[v]) < 1e-4 * (1 + abs(adj[v]))


@pytest.mark.parametrize("kind", ["circle", "clover"])
def test_shape_sample_gradient_of_linear_field(kind):
    """For g = a x + b y (harmonic, linear), grad u = (a, b) everywhere, so the
    per-sample estimate (normal derivative by the off-centred Poisson kernel
    with LIB + tangential derivative) averages to (a, b)."""
    a, b = 0.7, -1.3
    v, ix = C.circle_polyline(1024) if kind == "circle" else C.star_polyline(2048)
    s = O.Scene(v, ix, (a * v[:, 0] + b * v[:, 1]).astype(np.float32))
    p = np.random.default_rng(6).uniform(-0.3, 0.3, size=(64, 2)).astype(np.float32)
    _, sg = s.shape_gradient(p, 32, 256, 1e-3, 1e-3, 0.1, seed=4, sample_grad_u=True)
    ok = np.abs(sg).sum(1) > 0
    m = sg[ok].mean(0)
    se = sg[ok].std(0) / np.sqrt(ok.sum())
    assert ok.mean() > 0.95
    assert abs(m[0] - a) < 4 * se[0] + 0.01 and abs(m[1] - b) < 4 * se[1] + 0.01, (m, se)


def test_shape_sample_gradient_3d_linear_field():
    v, f, _ = C.displaced_sphere(2)
    gvec = np.array([0.4, -0.8, 1.1])
    s = O.Scene(v, f, (v @ gvec).astype(np.float32))
    p = np.random.default_rng(7).uniform(-0.2, 0.2, size=(24, 3)).astype(np.float32)
    _, sg = s.shape_gradient(p, 16, 256, 1e-3, 1e-3, 0.1, seed=5, sample_grad_u=True)
    ok = np.abs(sg).sum(1) > 0
    m = sg[ok].mean(0)
    se = sg[ok].std(0) / np.sqrt(ok.sum())
    assert (np.abs(m - gvec) < 4 * se + 0.02).all(), (m, se)


@pytest.mark.parametrize("dim", [2, 3])
def test_cpu_lbvh_equals_bruteforce(dim):
    """The oracle's CPU LBVH (the CPU baseline's query) returns the scan's
    primitive and squared distance everywhere, ties and far points included,
    so walks through it are identical."""
    rng = np.random.default_rng(dim + 10)
    if dim == 2:
        v, ix = C.star_polyline(2048, a=0.25, k=5)
        p = rng.uniform(-1.6, 1.6, size=(20000, 2)).astype(np.float32)
        p[:8] = 0.0   # equidistant from many segments
    else:
        v, ix, _ = C.displaced_sphere(3)
        p = rng.uniform(-1.5, 1.5, size=(8000, 3)).astype(np.float32)
        p[:8] = 0.0
    vals = np.zeros(v.shape[0], np.float32)
    a, b = O.Scene(v, ix, vals), O.Scene(v, ix, vals, use_bvh=True)
    i1, d1 = a.closest_index(p)
    i2, d2 = b.closest_index(p)
    np.testing.assert_array_equal(i1, i2)
    np.testing.assert_array_equal(d1, d2)
    q = p[np.abs(a.sdf(p)) < 0.9][:400] * 0.9
    r1 = a.walk_reduce(q, 8, 256, 1e-3, seed=3, walk_values=True)
    r2 = b.walk_reduce(q, 8, 256, 1e-3, seed=3, walk_values=True)
    np.testing.assert_array_equal(r1["walk_value"], r2["walk_value"])
