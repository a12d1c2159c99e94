"""Seeded scene and query generators for the five benchmark workloads.

Everything here is NumPy (float32 geometry, int32 indices) so the same inputs
feed the HIP path and the CPU oracle.  Boundary data are harmonic where a
closed form exists, so the exact solution is known:

* circle polyline, g = x^2 - y^2 + 0.5 x                       (u = g)
* clover polyline r = 1 + 0.3 cos 4t with source sin 3x cos 2y; the
  manufactured twin uses g = x^2 - y^2 + sin 3x cos 2y / 13    (u = g)
* displaced icosphere, g = x^2 - y^2 + z                       (u = g)
"""
from __future__ import annotations

import math

import numpy as np

# ---------------------------------------------------------------------------
# 2D polylines
# ---------------------------------------------------------------------------


def circle_polyline(n=1024, radius=1.0):
    """Closed CCW polyline on a circle (outward normals (dy, -dx))."""
    t = 2.0 * np.pi * np.arange(n) / n
    v = np.stack([radius * np.cos(t), radius * np.sin(t)], axis=1).astype(np.float32)
    i = np.arange(n, dtype=np.int32)
    return v, np.stack([i, (i + 1) % n], axis=1).astype(np.int32)


def star_polyline(n=4096, a=0.3, k=4, phase=0.0):
    """Closed CCW polyline r(t) = 1 + a cos(k t + phase) (the clover for k = 4)."""
    t = 2.0 * np.pi * np.arange(n) / n
    r = 1.0 + a * np.cos(k * t + phase)
    v = np.stack([r * np.cos(t), r * np.sin(t)], axis=1).astype(np.float32)
    i = np.arange(n, dtype=np.int32)
    return v, np.stack([i, (i + 1) % n], axis=1).astype(np.int32)


def grid_points(res, lo, hi):
    x = np.linspace(lo, hi, res, dtype=np.float64)
    X, Y = np.meshgrid(x, x, indexing="xy")
    return np.stack([X.ravel(), Y.ravel()], axis=1)


def g_circle(p):
    p = np.asarray(p, dtype=np.float64)
    return p[:, 0] ** 2 - p[:, 1] ** 2 + 0.5 * p[:, 0]


def g_saddle(p):
    p = np.asarray(p, dtype=np.float64)
    return p[:, 0] ** 2 - p[:, 1] ** 2


def f_clover_source(p):
    p = np.asarray(p, dtype=np.float64)
    return np.sin(3 * p[:, 0]) * np.cos(2 * p[:, 1])


def u_clover_twin(p):
    """Exact solution of -Lap u = sin3x cos2y with g = x^2-y^2 + sin3x cos2y/13."""
    return g_saddle(p) + f_clover_source(p) / 13.0


# 8-term Fourier basis for the clover source: sin(ax x + ay y + ph).
# sin3x cos2y = 0.5 sin(3x+2y) + 0.5 sin(3x-2y).
FOURIER_TERMS = np.array([
    [3.0, 2.0, 0.0], [3.0, -2.0, 0.0], [3.0, 2.0, np.pi / 2], [3.0, -2.0, np.pi / 2],
    [2.0, 3.0, 0.0], [2.0, -3.0, 0.0], [2.0, 3.0, np.pi / 2], [2.0, -3.0, np.pi / 2],
], dtype=np.float64)
FOURIER_COEFFS = np.array([0.5, 0.5, 0, 0, 0, 0, 0, 0], dtype=np.float64)


def fourier_params(coeffs=FOURIER_COEFFS):
    c = np.asarray(coeffs, dtype=np.float64).reshape(-1, 1)
    return np.concatenate([FOURIER_TERMS, c], axis=1).astype(np.float32)   # [K, 4]


def config1(res=128):
    """2D Laplace, unit circle of 1024 segments, g = x^2 - y^2 + 0.5x,
    128x128 grid on [-1,1]^2 clipped to r < 0.95, 64 walks, eps 1e-3, 256 steps."""
    v, ix = circle_polyline(1024)
    vals = g_circle(v).astype(np.float32)
    p = grid_points(res, -1.0, 1.0)
    p = p[np.hypot(p[:, 0], p[:, 1]) < 0.95].astype(np.float32)
    return dict(vertices=v, indices=ix, values=vals, points=p, nwalks=64, nsteps=256, epsilon=1e-3, seed=0,
                exact=g_circle(p))


def clover_inside(p, a=0.3, k=4, shrink=0.0):
    r = np.hypot(p[:, 0], p[:, 1])
    t = np.arctan2(p[:, 1], p[:, 0])
    return r < (1.0 + a * np.cos(k * t)) - shrink


def config2(res=512, twin=False):
    """2D Poisson on the clover r = 1 + 0.3 cos 4t (4096 segments), source
    sin3x cos2y as 8 Fourier terms, g = x^2 - y^2 (twin: exact g), 512x512
    grid on [-1.3,1.3]^2 inside, 128 walks, eps 1e-4."""
    v, ix = star_polyline(4096)
    vals = (u_clover_twin(v) if twin else g_saddle(v)).astype(np.float32)
    p = grid_points(res, -1.3, 1.3)
    p = p[clover_inside(p)].astype(np.float32)
    out = dict(vertices=v, indices=ix, values=vals, points=p, nwalks=128, nsteps=256, epsilon=1e-4, seed=0,
               source_type=3, source_params=fourier_params())
    if twin:
        out["exact"] = u_clover_twin(p)
    return out


# ---------------------------------------------------------------------------
# 3D meshes
# ---------------------------------------------------------------------------


def icosphere(subdiv):
    """Unit icosphere: 20 * 4^s triangles, 10 * 4^s + 2 vertices, outward CCW."""
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = [np.array(x, dtype=np.float64) / np.linalg.norm(x) for x in verts]
    F = np.array(faces, dtype=np.int64)
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        nf = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = np.array(nf, dtype=np.int64)
    return np.array(V), F


def _legendre(l, x):
    if l == 0:
        return np.ones_like(x)
    if l == 1:
        return x
    p0, p1 = np.ones_like(x), x
    for k in range(2, l + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    return p1


class SmoothNoise:
    """n(v) = sum_j c_j P_{l_j}(v . d_j): 8 zonal spherical harmonics about
    random directions d_j, degrees l_j in 1..4, c_j ~ N(0,1), scaled so that
    max |n| over a fine sphere sampling is 1."""

    def __init__(self, seed=0, terms=8, max_degree=4):
        rng = np.random.default_rng(seed)
        d = rng.normal(size=(terms, 3))
        self.dirs = d / np.linalg.norm(d, axis=1, keepdims=True)
        self.degrees = rng.integers(1, max_degree + 1, size=terms)
        self.coeffs = rng.normal(size=terms)
        probe, _ = icosphere(5)
        self.scale = 1.0
        self.scale = 1.0 / np.abs(self(probe)).max()

    def __call__(self, u):
        u = np.asarray(u, dtype=np.float64)
        u = u / np.linalg.norm(u, axis=-1, keepdims=True)
        out = np.zeros(u.shape[:-1])
        for d, l, c in zip(self.dirs, self.degrees, self.coeffs):
            out += c * _legendre(int(l), u @ d)
        return out * self.scale


def g_harmonic3(p):
    p = np.asarray(p, dtype=np.float64)
    return p[:, 0] ** 2 - p[:, 1] ** 2 + p[:, 2]


def displaced_sphere(subdiv=6, amp=0.15, seed=0):
    V, F = icosphere(subdiv)
    noise = SmoothNoise(seed)
    r = 1.0 + amp * noise(V)
    return (V * r[:, None]).astype(np.float32), F.astype(np.int32), noise


def points_in_star(n, noise, amp=0.15, seed=1, shrink=0.98):
    """Uniform points inside r < shrink * (1 + amp n(dir))."""
    rng = np.random.default_rng(seed)
    out = []
    have = 0
    rmax = 1.0 + amp
    while have < n:
        q = rng.uniform(-rmax, rmax, size=(2 * (n - have) + 1024, 3))
        r = np.linalg.norm(q, axis=1)
        ok = r < shrink * (1.0 + amp * noise(q))
        q = q[ok]
        out.append(q)
        have += q.shape[0]
    return np.concatenate(out)[:n].astype(np.float32)


def config3(n_points=1_000_000, subdiv=6):
    """3D Laplace on the displaced icosphere (subdiv 6: 81,920 triangles),
    g = x^2 - y^2 + z, uniform interior points, 32 walks, eps 1e-3."""
    v, f, noise = displaced_sphere(subdiv)
    vals = g_harmonic3(v).astype(np.float32)
    p = points_in_star(n_points, noise)
    return dict(vertices=v, indices=f, values=vals, points=p, nwalks=32, nsteps=256, epsilon=1e-3, seed=0,
                exact=g_harmonic3(p))


def decimated_star_mesh(n_tris=200_000, subdiv=7, amp=0.15, seed=0):
    """Seeded decimation of the subdivision-7 icosphere: keep n_tris/2 + 2 of
    its vertices (random, seeded), re-triangulate their directions by convex
    hull (2N - 4 triangles), displace radially by 1 + amp n."""
    from scipy.spatial import ConvexHull
    V, _ = icosphere(subdiv)
    rng = np.random.default_rng(seed)
    keep = np.sort(rng.choice(V.shape[0], size=n_tris // 2 + 2, replace=False))
    U = V[keep]
    hull = ConvexHull(U)
    F = hull.simplices.astype(np.int64)
    # orient outward
    c = U[F].mean(axis=1)
    nrm = np.cross(U[F[:, 1]] - U[F[:, 0]], U[F[:, 2]] - U[F[:, 0]])
    flip = (nrm * c).sum(axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    noise = SmoothNoise(seed)
    r = 1.0 + amp * noise(U)
    return (U * r[:, None]).astype(np.float32), U.astype(np.float32), F.astype(np.int32), noise


def random_star_mesh(n_tris, amp=0.15, seed=0):
    """Closed star-shaped triangle mesh with n_tris = 2N - 4 triangles: convex
    hull of N seeded random directions, displaced radially by 1 + amp n."""
    from scipy.spatial import ConvexHull
    rng = np.random.default_rng(seed)
    N = max(n_tris // 2 + 2, 4)
    U = rng.normal(size=(N, 3))
    U /= np.linalg.norm(U, axis=1, keepdims=True)
    F = ConvexHull(U).simplices.astype(np.int64)
    c = U[F].mean(axis=1)
    nrm = np.cross(U[F[:, 1]] - U[F[:, 0]], U[F[:, 2]] - U[F[:, 0]])
    flip = (nrm * c).sum(axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    noise = SmoothNoise(seed)
    return (U * (1.0 + amp * noise(U))[:, None]).astype(np.float32), F.astype(np.int32), noise


def icosphere_level(n_tris):
    """Subdivision level (3..6) of the icosphere whose 20 * 4^s triangles are
    nearest to n_tris on a log scale: 1,280 / 5,120 / 20,480 / 81,920."""
    return int(np.clip(np.rint(np.log(n_tris / 20.0) / np.log(4.0)), 3, 6))


def config5(scale=1.0, seed=0):
    """64 uneven scenes: 32 clover/star polylines with 256..16,384 segments
    and 32 displaced icospheres with 1k..100k triangles (log-uniform target
    sizes; each 3D scene is the icosphere of the nearest subdivision level,
    1,280..81,920 triangles, displaced radially by 1 + 0.15 n with its own
    noise seed), 4k..1M query points each (log-uniform, times `scale`),
    16..128 walks, half of the 2D scenes Poisson (Fourier source), the rest
    Laplace."""
    rng = np.random.default_rng(seed)
    out = []
    for j in range(64):
        n_pts = int(np.exp(rng.uniform(np.log(4e3), np.log(1e6))) * scale)
        n_pts = max(n_pts, 16)
        W = int(rng.integers(16, 129))
        if j < 32:
            nseg = int(np.exp(rng.uniform(np.log(256), np.log(16384))))
            k = int(rng.integers(3, 7))
            a = float(rng.uniform(0.1, 0.3))
            v, ix = star_polyline(nseg, a=a, k=k, phase=float(rng.uniform(0, 2 * np.pi)))
            vals = g_saddle(v).astype(np.float32)
            r = np.sqrt(rng.random(n_pts)) * (1 - a) * 0.98
            t = rng.random(n_pts) * 2 * np.pi
            p = np.stack([r * np.cos(t), r * np.sin(t)], 1).astype(np.float32)
            item = dict(dim=2, vertices=v, indices=ix, values=vals, points=p, nwalks=W, nsteps=256, epsilon=1e-3,
                        seed=j, exact=g_saddle(p))
            if j % 2:
                # Poisson with the clover source; boundary data of its manufactured twin, so u is known
                item.update(source_type=3, source_params=fourier_params(), values=u_clover_twin(v).astype(np.float32),
                            exact=u_clover_twin(p))
        else:
            ntri = int(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
            v, f, noise = displaced_sphere(icosphere_level(ntri), seed=j)
            p = points_in_star(n_pts, noise, seed=j)
            item = dict(dim=3, vertices=v, indices=f, values=g_harmonic3(v).astype(np.float32), points=p, nwalks=W,
                        nsteps=256, epsilon=1e-3, seed=j, exact=g_harmonic3(p))
        out.append(item)
    return out


def points_in_ball(n, radius, seed=2):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * rng.random(n) ** (1.0 / 3.0)
    return (d * r[:, None]).astype(np.float32)


def config4(n_tris=200_000, n_points=262_144, seed=0):
    """3D shape gradient: the seeded decimation of the subdivision-7 sphere
    (200k triangles), displaced by 1 + 0.15 n.  Vertex values carry g = x^2 -
    y^2 + z of each vertex's undisplaced position (the data painted on the
    sphere), so the target -- the solution on the undisplaced sphere -- is
    g itself.  Query points: uniform in the ball of radius 0.8 (inside both
    surfaces).  Loss L = mean (u - g)^2; gradient w.r.t. all vertex positions."""
    V, U, F, noise = decimated_star_mesh(n_tris, seed=seed)
    vals = g_harmonic3(U).astype(np.float32)
    p = points_in_ball(n_points, 0.8, seed=seed + 2)
    return dict(vertices=V, sphere_vertices=U, indices=F, values=vals, points=p, target=g_harmonic3(p),
                nwalks=16, nsteps=256, epsilon=1e-3, epsilon2=1e-3, clamping=0.1, seed=0)
