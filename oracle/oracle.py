"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.  Arrays are NumPy.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    return ctypes.CDLL(LIB_PATH)


lib = _load()
_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float


class _Scene(ctypes.Structure):
    _fields_ = [("dim", _i), ("n_vertices", _i), ("n_prims", _i), ("vertices", _vp), ("indices", _vp),
                ("values", _vp), ("normals", _vp), ("aabb_min", _f * 3), ("aabb_max", _f * 3),
                ("source_type", _i), ("source_params", _vp), ("n_source_params", _i), ("bvh", _vp)]


class _WalkParams(ctypes.Structure):
    _fields_ = [("nwalks", _i), ("nsteps", _i), ("epsilon", _f), ("double_sided", _i),
                ("prevent_fd_artifacts", _i), ("use_IS_for_greens", _i)]


class _GradParams(ctypes.Structure):
    _fields_ = [("nwalks", _i), ("nsteps", _i), ("epsilon", _f), ("epsilon2", _f), ("clamping", _f),
                ("double_sided", _i), ("control_variates", _i)]


class _Pcg(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint64), ("inc", ctypes.c_uint64)]


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


lib.or_pcg32_next_u32.restype = ctypes.c_uint32
lib.or_basis_exponent.restype = ctypes.c_int
lib.or_basis_exponent.argtypes = [ctypes.c_int, ctypes.c_int]
lib.or_pcg32_next_u32.argtypes = [ctypes.POINTER(_Pcg)]
lib.or_pcg32_next_1d.restype = _f
lib.or_pcg32_next_1d.argtypes = [ctypes.POINTER(_Pcg)]
lib.or_pcg32_next_f64.restype = ctypes.c_double
lib.or_pcg32_next_f64.argtypes = [ctypes.POINTER(_Pcg)]
lib.or_pcg32_seed.argtypes = [ctypes.POINTER(_Pcg), ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
lib.or_pcg32_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, _i, _vp]
lib.or_tea.argtypes = [ctypes.c_int32, ctypes.c_int32, _i, ctypes.POINTER(ctypes.c_int32),
                       ctypes.POINTER(ctypes.c_int32)]
lib.or_batch_elementary.argtypes = [_i, _vp, _i, _vp]
lib.or_scene_init.argtypes = [ctypes.POINTER(_Scene)]
lib.or_batch_closest.argtypes = [ctypes.POINTER(_Scene), _vp, _i, _vp, _vp]
lib.or_batch_sdf.argtypes = [ctypes.POINTER(_Scene), _vp, _i, _vp]
lib.or_batch_closest_point.argtypes = [ctypes.POINTER(_Scene), _vp, _i, _vp, _vp, _vp, _vp, _vp]
lib.or_batch_dirichlet.argtypes = [ctypes.POINTER(_Scene), _vp, _vp, _vp, _i, _vp]
lib.or_batch_lib.argtypes = [ctypes.POINTER(_Scene), _vp, _vp, _vp, _i, _f, _vp]
lib.or_batch_single_walk.argtypes = [ctypes.POINTER(_Scene), ctypes.POINTER(_WalkParams), _vp, _i, _vp, _vp,
                                     _vp, _vp, _vp, _vp, _vp, _vp]
lib.or_walk_reduce.argtypes = [ctypes.POINTER(_Scene), ctypes.POINTER(_WalkParams), _vp, _i, ctypes.c_int64, _i,
                               _vp, _vp, _vp, _vp, _vp, _vp]
lib.or_adjoint_values.argtypes = [ctypes.POINTER(_Scene), _i, _i, _vp, _vp, _vp, _vp]
lib.or_shape_gradient.argtypes = [ctypes.POINTER(_Scene), ctypes.POINTER(_GradParams), _vp, _i, ctypes.c_int64,
                                  _vp, _i, _vp, _vp]
lib.or_lbvh_build.argtypes = [ctypes.POINTER(_Scene), _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_i)]
lib.or_lbvh_build.restype = _i
lib.or_green_cdf.restype = ctypes.POINTER(_f)
lib.or_green_cdf.argtypes = [ctypes.POINTER(_i)]
lib.or_sample_green_radius.restype = _f
lib.or_sample_green_radius.argtypes = [_f, _f, ctypes.POINTER(_f), ctypes.POINTER(_i)]
lib.or_interior_sample.restype = _i
lib.or_interior_sample.argtypes = [_f, _f, _f, _i] + [ctypes.POINTER(_f)] * 4
lib.or_bvh_create.restype = _vp
lib.or_bvh_create.argtypes = [ctypes.POINTER(_Scene)]
lib.or_bvh_destroy.argtypes = [_vp]
lib.or_morton.restype = ctypes.c_uint32
lib.or_morton.argtypes = [_i, _vp]


# ---- random numbers --------------------------------------------------------
class Pcg32:
    def __init__(self, state=0, inc=1):
        self.s = _Pcg(state, inc)

    @staticmethod
    def seeded(index, initstate, initseq):
        r = Pcg32()
        lib.or_pcg32_seed(ctypes.byref(r.s), index & (2 ** 64 - 1), initstate & (2 ** 64 - 1),
                          initseq & (2 ** 64 - 1))
        return r

    def next_u32(self):
        return lib.or_pcg32_next_u32(ctypes.byref(self.s))

    def next_1d(self):
        return lib.or_pcg32_next_1d(ctypes.byref(self.s))

    def next_f64(self):
        return lib.or_pcg32_next_f64(ctypes.byref(self.s))


def pcg32_stream(index, initstate, initseq, n):
    out = np.zeros(n, dtype=np.uint32)
    lib.or_pcg32_stream(index & (2 ** 64 - 1), initstate & (2 ** 64 - 1), initseq & (2 ** 64 - 1), n, _p(out))
    return out


def tea(v0, v1, rounds=4):
    a, b = ctypes.c_int32(), ctypes.c_int32()
    lib.or_tea(int(np.int32(v0)), int(np.int32(v1)), rounds, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def elementary(fn, x):
    """fn: 'sin_turns', 'cos_turns', 'sin', 'log', 'exp'."""
    code = {"sin_turns": 0, "cos_turns": 1, "sin": 2, "log": 3, "exp": 4}[fn]
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    lib.or_batch_elementary(code, _p(x), x.size, _p(out))
    return out


def green_cdf():
    """The tabulated CDF of r log(1/r) (float32 array)."""
    n = _i()
    ptr = lib.or_green_cdf(ctypes.byref(n))
    return np.ctypeslib.as_array(ptr, shape=(n.value,)).copy()


def sample_green_radius(u, R):
    """(radius, radial pdf, bin) of include/wos.cuh:49-75 for each u."""
    u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1)
    rad = np.empty_like(u)
    pdf = np.empty_like(u)
    b = np.empty(u.shape, np.int32)
    pf, bi = _f(), _i()
    for j, x in enumerate(u):
        rad[j] = lib.or_sample_green_radius(float(x), float(R), ctypes.byref(pf), ctypes.byref(bi))
        pdf[j], b[j] = pf.value, bi.value
    return rad, pdf, b


def interior_sample(u0, u1, d, use_is):
    """(dx, dy, G, w, take) of one interior sample per (u0, u1)."""
    u0 = np.ascontiguousarray(u0, dtype=np.float32).reshape(-1)
    u1 = np.ascontiguousarray(u1, dtype=np.float32).reshape(-1)
    out = np.empty((u0.size, 4), np.float32)
    take = np.empty(u0.size, bool)
    v = [_f() for _ in range(4)]
    for j in range(u0.size):
        take[j] = bool(lib.or_interior_sample(float(u0[j]), float(u1[j]), float(d), int(bool(use_is)),
                                              *[ctypes.byref(x) for x in v]))
        out[j] = [x.value for x in v]
    return out[:, 0], out[:, 1], out[:, 2], out[:, 3], take


def morton(dim, unit):
    u = np.ascontiguousarray(unit, dtype=np.float32)
    return lib.or_morton(dim, _p(u))


# ---- scenes ------------------------------------------------------------------
class Scene:
    """use_bvh: answer closest-primitive queries through the CPU LBVH
    (or_bvh_create) instead of the reference's brute-force scan; results are
    identical, only the cost differs."""

    def __init__(self, vertices, indices, values, source_type=0, source_params=None, use_bvh=False):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float32)
        self.dim = self.vertices.shape[1]
        self.indices = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1, self.dim)
        self.values = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
        self.normals = np.zeros_like(self.vertices)
        self.source_params = (None if source_params is None else
                              np.ascontiguousarray(source_params, dtype=np.float32).reshape(-1))
        s = _Scene()
        s.dim = self.dim
        s.n_vertices = self.vertices.shape[0]
        s.n_prims = self.indices.shape[0]
        s.vertices = _p(self.vertices)
        s.indices = _p(self.indices)
        s.values = _p(self.values)
        s.normals = _p(self.normals)
        s.source_type = int(source_type)
        s.source_params = _p(self.source_params)
        s.n_source_params = 0 if self.source_params is None else self.source_params.size
        s.bvh = None
        lib.or_scene_init(ctypes.byref(s))
        self.s = s
        self._bvh = None
        if use_bvh:
            self.set_bvh(True)

    def set_bvh(self, enable):
        if enable and not self._bvh:
            self.s.bvh = None
            h = lib.or_bvh_create(ctypes.byref(self.s))
            if not h:
                raise RuntimeError("or_bvh_create failed")
            self._bvh = h
            self.s.bvh = h
        elif not enable and self._bvh:
            self.s.bvh = None
            lib.or_bvh_destroy(self._bvh)
            self._bvh = None

    def __del__(self):
        try:
            self.set_bvh(False)
        except Exception:   # interpreter shutdown
            pass

    @property
    def aabb(self):
        return np.array(self.s.aabb_min[: self.dim]), np.array(self.s.aabb_max[: self.dim])

    def _pts(self, p):
        return np.ascontiguousarray(p, dtype=np.float32).reshape(-1, self.dim)

    def closest_index(self, p):
        p = self._pts(p)
        idx = np.empty(p.shape[0], dtype=np.int32)
        d2 = np.empty(p.shape[0], dtype=np.float32)
        lib.or_batch_closest(ctypes.byref(self.s), _p(p), p.shape[0], _p(idx), _p(d2))
        return idx, d2

    def sdf(self, p):
        p = self._pts(p)
        out = np.empty(p.shape[0], dtype=np.float32)
        lib.or_batch_sdf(ctypes.byref(self.s), _p(p), p.shape[0], _p(out))
        return out

    def closest_point(self, p):
        p = self._pts(p)
        n, D = p.shape
        r = dict(valid=np.empty(n, np.int32), p=np.empty((n, D), np.float32), n=np.empty((n, D), np.float32),
                 t=np.empty(n if D == 2 else (n, 2), np.float32), prim=np.empty(n, np.int32))
        lib.or_batch_closest_point(ctypes.byref(self.s), _p(p), n, _p(r["valid"]), _p(r["p"]), _p(r["n"]),
                                   _p(r["t"]), _p(r["prim"]))
        return r

    def dirichlet(self, prim, t, valid=None):
        """Scene.dirichlet over records (include/scene.cuh:251-264)."""
        prim = np.ascontiguousarray(prim, dtype=np.int32).reshape(-1)
        t = np.ascontiguousarray(t, dtype=np.float32)
        v = None if valid is None else np.ascontiguousarray(valid, dtype=np.int32)
        out = np.empty(prim.size, np.float32)
        lib.or_batch_dirichlet(ctypes.byref(self.s), _p(v), _p(prim), _p(t), prim.size, _p(out))
        return out

    def largest_inscribed_ball(self, valid, p, n, eps=1e-3):
        p = self._pts(p)
        nn = self._pts(n)
        v = np.ascontiguousarray(valid, dtype=np.int32)
        out = np.empty(p.shape[0], np.float32)
        lib.or_batch_lib(ctypes.byref(self.s), _p(v), _p(p), _p(nn), p.shape[0], float(eps), _p(out))
        return out

    def single_walk(self, p, state, inc, nsteps=32, epsilon=1e-3, double_sided=False, prevent_fd_artifacts=False,
                    use_is=False):
        p = self._pts(p)
        n, D = p.shape
        st = np.ascontiguousarray(state, dtype=np.uint64).copy()
        ic = np.ascontiguousarray(inc, dtype=np.uint64).copy()
        w = _WalkParams(1, nsteps, epsilon, int(double_sided), int(prevent_fd_artifacts), int(bool(use_is)))
        r = dict(valid=np.empty(n, np.int32), p=np.empty((n, D), np.float32), n=np.empty((n, D), np.float32),
                 t=np.empty(n if D == 2 else (n, 2), np.float32), prim=np.empty(n, np.int32),
                 contrib=np.empty(n, np.float32))
        lib.or_batch_single_walk(ctypes.byref(self.s), ctypes.byref(w), _p(p), n, _p(st), _p(ic), _p(r["valid"]),
                                 _p(r["p"]), _p(r["n"]), _p(r["t"]), _p(r["prim"]), _p(r["contrib"]))
        r["state"], r["inc"] = st, ic
        return r

    def walk_reduce(self, points, nwalks, nsteps=32, epsilon=1e-3, seed=0, double_sided=False,
                    prevent_fd_artifacts=False, threads=0, walk_values=False, records=False, basis=False,
                    use_is=False):
        """or_walk_reduce: per-point FP64 mean and variance; on request the
        per-walk values, exit records and ``basis`` (P x K: per point the sum
        over its walks of d value / d coeff_k, in 64-bit fixed point)."""
        p = self._pts(points)
        P = p.shape[0]
        w = _WalkParams(int(nwalks), int(nsteps), float(epsilon), int(double_sided), int(prevent_fd_artifacts),
                        int(bool(use_is)))
        out = dict(mean=np.empty(P, np.float64), var=np.empty(P, np.float64))
        wv = np.empty(P * nwalks, np.float32) if walk_values else None
        rp = np.empty(P * nwalks, np.int32) if records else None
        rt = np.empty(P * nwalks * (self.dim - 1), np.float32) if records else None
        K = (self.source_params.size // 4) if (basis and self.s.source_type == 3) else 0
        bs = np.empty(P * K, np.float64) if K else None
        lib.or_walk_reduce(ctypes.byref(self.s), ctypes.byref(w), _p(p), P, int(seed), int(threads or os.cpu_count()),
                           _p(out["mean"]), _p(out["var"]), _p(wv), _p(rp), _p(rt), _p(bs))
        if walk_values:
            out["walk_value"] = wv
        if records:
            out["rec_prim"], out["rec_t"] = rp, rt
        if K:
            out["basis"] = bs.reshape(P, K)
        return out

    def adjoint_values(self, n_points, nwalks, rec_prim, rec_t, grad_out):
        g = np.zeros(self.vertices.shape[0], np.float64)
        go = np.ascontiguousarray(grad_out, dtype=np.float64)
        lib.or_adjoint_values(ctypes.byref(self.s), int(n_points), int(nwalks), _p(rec_prim), _p(rec_t), _p(go),
                              _p(g))
        return g

    def shape_gradient(self, points, nwalks, nsteps, epsilon, epsilon2, clamping, seed=0, grad_out=None,
                       double_sided=False, control_variates=True, threads=0, sample_grad_u=False):
        p = self._pts(points)
        P = p.shape[0]
        gp = _GradParams(int(nwalks), int(nsteps), float(epsilon), float(epsilon2), float(clamping),
                         int(double_sided), int(control_variates))
        gv = np.zeros(self.vertices.shape, np.float64)
        go = None if grad_out is None else np.ascontiguousarray(grad_out, dtype=np.float64)
        sg = np.zeros((P * nwalks, self.dim), np.float64) if sample_grad_u else None
        lib.or_shape_gradient(ctypes.byref(self.s), ctypes.byref(gp), _p(p), P, int(seed), _p(go),
                              int(threads or os.cpu_count()), _p(gv), _p(sg))
        return gv, sg

    def lbvh(self):
        n = self.indices.shape[0]
        codes = np.empty(n, np.uint32)
        prim = np.empty(n, np.int32)
        ch = np.empty(max(n - 1, 1) * 2, np.int32)
        pi = np.empty(max(n - 1, 1), np.int32)
        pl = np.empty(n, np.int32)
        depth = _i()
        rc = lib.or_lbvh_build(ctypes.byref(self.s), _p(codes), _p(prim), _p(ch), _p(pi), _p(pl),
                               ctypes.byref(depth))
        assert rc == 0
        return dict(codes=codes, prim=prim, children=ch[: 2 * (n - 1)].reshape(-1, 2),
                    parent_internal=pi[: n - 1], parent_leaf=pl, max_depth=depth.value)
