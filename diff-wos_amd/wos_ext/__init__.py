"""wos_ext -- the reference's native module surface over the MI355X C ABI.

Mirrors the pybind11 module of the reference (src/python/main.cpp:9-18,
src/python/scene.cu:10-88, src/python/wos.cu:5-25): ``Scene``, ``Scene3D``,
``ClosestPointRecord``, ``ClosestPointRecord3D``, ``AABB``, ``AABB3D``,
``WoS`` and ``WoS3D``.  Arrays are torch tensors on the current HIP device
(Dr.Jit arrays in the reference); points are ``[n, dim]`` row-major float32.

Every call goes to ``libwos_amd.so`` (include/wos_amd.h).  There is no CPU
path: importing this module fails if the library is missing, and calls fail
when no GPU is present.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import torch

__all__ = ["Scene", "Scene3D", "ClosestPointRecord", "ClosestPointRecord3D", "AABB", "AABB3D",
           "WoS", "WoS3D", "WoSError", "library_path", "lib"]

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libwos_amd.so")


class WoSError(RuntimeError):
    """A libwos_amd entry point returned an error status."""


def _load():
    if not os.path.exists(library_path):
        raise ImportError(
            f"{library_path} is missing: build it with `make -C diff-wos_amd/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    return ctypes.CDLL(library_path)


lib = _load()

_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float


class _WalkParams(ctypes.Structure):
    _fields_ = [("nwalks", _i), ("nsteps", _i), ("epsilon", _f), ("double_sided", _i),
                ("prevent_fd_artifacts", _i), ("use_IS_for_greens", _i)]


class _GradParams(ctypes.Structure):
    _fields_ = [("nwalks", _i), ("nsteps", _i), ("epsilon", _f), ("epsilon2", _f), ("clamping", _f),
                ("double_sided", _i), ("control_variates", _i)]


class _Record(ctypes.Structure):
    _fields_ = [("valid", _vp), ("p", _vp), ("n", _vp), ("t", _vp), ("prim_id", _vp), ("contrib", _vp)]


def _sig(name, *args):
    fn = getattr(lib, name)
    fn.argtypes = list(args)
    fn.restype = _i
    return fn


lib.wos_last_error.restype = ctypes.c_char_p
lib.wos_last_error.argtypes = []
_sig("wos_scene_create", _i, _vp, _i, _vp, _i, _vp, _i, _i, _vp, _i, _vp, ctypes.POINTER(_vp))
_sig("wos_scene_destroy", _vp)
_sig("wos_scene_set_values", _vp, _vp, _vp)
_sig("wos_scene_set_source", _vp, _i, _vp, _i, _vp)
_sig("wos_scene_set_use_bvh", _vp, _i)
_sig("wos_scene_get_normals", _vp, _vp, _vp)
_sig("wos_scene_get_info", _vp, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("wos_scene_get_bvh", _vp, _vp, _vp, _vp, _vp)
_sig("wos_scene_set_grid", _vp, _f, _i, _i, _vp)
_sig("wos_scene_get_grid", _vp, _vp, _vp, _vp, _vp)
_sig("wos_scene_set_start_grid", _vp, _i, _vp)
_sig("wos_scene_get_start_grid", _vp, _vp, _vp, _vp)
_sig("wos_closest_point_index", _vp, _vp, _i, _vp, _vp, _vp)
_sig("wos_sdf", _vp, _vp, _i, _vp, _vp)
_sig("wos_closest_point", _vp, _vp, _i, ctypes.POINTER(_Record), _vp)
_sig("wos_largest_inscribed_ball", _vp, ctypes.POINTER(_Record), _i, _f, _vp, _vp)
_sig("wos_dirichlet", _vp, ctypes.POINTER(_Record), _i, _vp, _vp)
_sig("wos_single_walk", _vp, ctypes.POINTER(_WalkParams), _vp, _i, _vp, _vp, ctypes.POINTER(_Record), _vp)
_sig("wos_walk_reduce", _vp, ctypes.POINTER(_WalkParams), _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("wos_walk_reduce_noalloc", _vp, ctypes.POINTER(_WalkParams), _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
     _vp)
_sig("wos_walk_reduce_counts", _vp, ctypes.POINTER(_WalkParams), _vp, _i, _i, _i, _vp, _vp, _vp)
class _BatchItem(ctypes.Structure):
    _fields_ = [("scene", _vp), ("params", _WalkParams), ("points", _vp), ("n_points", _i), ("seed", _i),
                ("mean", _vp), ("mean64", _vp), ("var64", _vp), ("walk_value", _vp), ("rec_prim", _vp),
                ("rec_t", _vp), ("basis", _vp), ("grad_out", _vp), ("grad_values", _vp), ("grad_coeffs", _vp),
                ("grad_params", _GradParams), ("grad_vertices", _vp)]


_sig("wos_batch_create", ctypes.POINTER(_BatchItem), _i, ctypes.POINTER(_vp))
_sig("wos_batch_run", _vp, _vp)
_sig("wos_batch_adjoint", _vp, _vp)
_sig("wos_batch_shape_gradient", _vp, _vp)
_sig("wos_batch_destroy", _vp)
_sig("wos_adjoint_source", _vp, ctypes.POINTER(_WalkParams), _vp, _i, _i, _vp, _vp, _vp)
_sig("wos_debug_fail_workspace", _i)
_sig("wos_shape_gradient_counts", _vp, ctypes.POINTER(_GradParams), _vp, _i, _i, _vp, _vp)
_sig("wos_set_tuning", _i, _i, _i, _i)
_sig("wos_get_tuning", _vp, _vp, _vp, _vp)
_sig("wos_set_spatial_order", _i)
_sig("wos_adjoint_values", _vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("wos_shape_gradient", _vp, ctypes.POINTER(_GradParams), _vp, _i, _i, _vp, _vp, _vp, _vp)
_sig("wos_shape_derivative", _vp, ctypes.POINTER(_GradParams), _vp, _i, _i, _vp, _vp, _vp, _vp, _vp)

EXPORTED = ("wos_last_error", "wos_scene_create", "wos_scene_destroy", "wos_scene_set_values",
            "wos_scene_set_source", "wos_scene_set_use_bvh", "wos_scene_get_normals", "wos_scene_get_info",
            "wos_scene_get_bvh", "wos_scene_set_grid", "wos_scene_get_grid",
            "wos_scene_set_start_grid", "wos_scene_get_start_grid", "wos_closest_point_index", "wos_sdf", "wos_closest_point",
            "wos_largest_inscribed_ball", "wos_dirichlet", "wos_single_walk", "wos_walk_reduce",
            "wos_walk_reduce_counts", "wos_adjoint_values", "wos_shape_gradient", "wos_shape_derivative",
            "wos_set_tuning",
            "wos_get_tuning", "wos_set_spatial_order", "wos_batch_create", "wos_batch_run", "wos_batch_adjoint",
            "wos_batch_destroy", "wos_adjoint_source", "wos_debug_fail_workspace", "wos_shape_gradient_counts",
            "wos_walk_reduce_noalloc", "wos_batch_shape_gradient")


def set_tuning(chunk=None, wgs_per_cu=None, top_levels=None, window=None):
    """Process-wide launch tuning (walks per queue claim, workgroups per CU,
    BVH levels in LDS, node visits per lane between refills; window 0 is the
    per-dimension default, 8 in 2D and 16 in 3D); unspecified values keep
    their current setting."""
    c, w, t, v = get_tuning()
    _check(lib.wos_set_tuning(int(chunk or c), int(wgs_per_cu or w), int(t if top_levels is None else top_levels),
                              int(v if window is None else window)))


def set_spatial_order(enable=True):
    """Visit query points in Morton order (default) or input order; results
    are identical, only cache behaviour changes."""
    _check(lib.wos_set_spatial_order(int(bool(enable))))


def get_tuning():
    c, w, t, v = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(lib.wos_get_tuning(ctypes.byref(c), ctypes.byref(w), ctypes.byref(t), ctypes.byref(v)))
    return c.value, w.value, t.value, v.value


def _check(rc):
    if rc != 0:
        msg = lib.wos_last_error()
        raise WoSError(msg.decode() if msg else "libwos_amd call failed")


def device():
    if not torch.cuda.is_available():
        raise WoSError("wos_ext needs a HIP device; none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    return _vp(torch.cuda.current_stream().cuda_stream)


def _dev(t, dtype, shape_last=None, name="array"):
    """Contiguous device tensor of dtype (detached: the library only reads it)."""
    if not isinstance(t, torch.Tensor):
        t = torch.as_tensor(t)
    t = t.detach()
    if t.dtype != dtype:
        t = t.to(dtype)
    if t.device.type != "cuda":
        t = t.to(device())
    t = t.contiguous()
    if shape_last is not None:
        t = t.reshape(-1, shape_last)
    return t


def _p(t):
    return _vp(t.data_ptr()) if t is not None else _vp(None)


@dataclass
class AABB:
    """AABB (include/scene.cuh:62-66) -- host copies of min / max."""
    min: torch.Tensor
    max: torch.Tensor


AABB3D = AABB


class ClosestPointRecord:
    """dr::ClosestPointRecord<2> (include/scene.cuh:18-37)."""

    def __init__(self, valid=None, p=None, n=None, t=None, prim_id=None, contrib=None):
        self.valid, self.p, self.n, self.t, self.prim_id, self.contrib = valid, p, n, t, prim_id, contrib

    @staticmethod
    def create(valid, p, n, t, prim_id, contrib):
        return ClosestPointRecord(valid, p, n, t, prim_id, contrib)

    def _struct(self, dim):
        rec = _Record()
        keep = []
        for name, dt, last in (("valid", torch.uint8, None), ("p", torch.float32, dim), ("n", torch.float32, dim),
                               ("prim_id", torch.int32, None), ("contrib", torch.float32, None)):
            v = getattr(self, name)
            if v is None:
                continue
            if name == "valid":
                v = v.to(torch.uint8) if isinstance(v, torch.Tensor) else v
            v = _dev(v, dt, last, name)
            keep.append(v)
            setattr(rec, name, v.data_ptr())
        tv = getattr(self, "t", None) if dim == 2 else getattr(self, "uv", None)
        if tv is not None:
            tv = _dev(tv, torch.float32, None if dim == 2 else 2)
            keep.append(tv)
            rec.t = tv.data_ptr()
        return rec, keep


class ClosestPointRecord3D(ClosestPointRecord):
    """dr::ClosestPointRecord<3> (include/scene.cuh:39-58): ``uv`` instead of ``t``."""

    def __init__(self, valid=None, p=None, n=None, uv=None, prim_id=None, contrib=None):
        self.valid, self.p, self.n, self.uv, self.prim_id, self.contrib = valid, p, n, uv, prim_id, contrib

    @staticmethod
    def create(valid, p, n, uv, prim_id, contrib):
        return ClosestPointRecord3D(valid, p, n, uv, prim_id, contrib)


def _alloc_record(n, dim):
    dv = device()
    valid = torch.empty(n, dtype=torch.uint8, device=dv)
    p = torch.empty(n, dim, dtype=torch.float32, device=dv)
    nn = torch.empty(n, dim, dtype=torch.float32, device=dv)
    t = torch.empty(n if dim == 2 else (n, 2), dtype=torch.float32, device=dv)
    prim = torch.empty(n, dtype=torch.int32, device=dv)
    contrib = torch.empty(n, dtype=torch.float32, device=dv)
    rec = _Record(valid.data_ptr(), p.data_ptr(), nn.data_ptr(), t.data_ptr(), prim.data_ptr(), contrib.data_ptr())
    if dim == 2:
        out = ClosestPointRecord(valid.bool(), p, nn, t, prim, contrib)
    else:
        out = ClosestPointRecord3D(valid.bool(), p, nn, t, prim, contrib)
    return rec, out, valid


def _finish_record(out, valid):
    out.valid = valid.bool()
    return out


class _SceneBase:
    """Scene<N> (include/scene.cuh:317-495) behind a ``wos_scene`` handle."""

    _dim = 2

    def __init__(self, vertices=None, indices=None, values=None, use_bvh=False, source_type=0,
                 source_params=None):
        self._h = _vp(None)
        if vertices is None:   # py::init<>() (src/python/scene.cu:29)
            return
        self._build(vertices, indices, values, use_bvh, source_type, source_params)

    def _build(self, vertices, indices, values, use_bvh, source_type, source_params):
        D = self._dim
        v = _dev(vertices, torch.float32, D, "vertices")
        ix = _dev(indices, torch.int32, D, "indices")
        val = _dev(values, torch.float32, None, "values").reshape(-1)
        if val.numel() != v.shape[0]:
            raise WoSError(f"values has {val.numel()} entries for {v.shape[0]} vertices")
        sp = None
        if source_type:
            if source_params is None:
                raise WoSError("source_params is required when source_type != 0")
            sp = _dev(source_params, torch.float32).reshape(-1)
        h = _vp()
        _check(lib.wos_scene_create(D, _p(v), v.shape[0], _p(ix), ix.shape[0], _p(val), int(bool(use_bvh)),
                                    int(source_type), _p(sp), 0 if sp is None else sp.numel(), _stream(),
                                    ctypes.byref(h)))
        self.__dict__["_h"] = h
        self.__dict__["_vertices"] = v
        self.__dict__["_indices"] = ix
        self.__dict__["_values"] = val
        self.__dict__["_use_bvh"] = bool(use_bvh)
        self.__dict__["_source_type"] = int(source_type)
        self.__dict__["_source_params"] = sp

    def __del__(self):
        h = self.__dict__.get("_h")
        if h is not None and h.value:
            try:
                lib.wos_scene_destroy(h)
            except Exception:   # interpreter shutdown: module globals already cleared
                pass
            self.__dict__["_h"] = None

    @property
    def handle(self):
        if not self._h.value:
            raise WoSError("scene is empty")
        return self._h

    @property
    def dim(self):
        return self._dim

    @property
    def c_source_params(self):
        return self.__dict__.get("_source_params")

    @property
    def source_type(self):
        return self.__dict__.get("_source_type", 0)

    # ---- def_readwrite fields (src/python/scene.cu:35-47) -------------------
    @property
    def c_vertices(self):
        return self._vertices

    @property
    def c_indices(self):
        return self._indices

    @property
    def c_values(self):
        return self._values

    def set_values(self, values):
        val = _dev(values, torch.float32).reshape(-1)
        _check(lib.wos_scene_set_values(self.handle, _p(val), _stream()))
        self.__dict__["_values"] = val

    def set_source(self, source_type, source_params):
        sp = None if not source_type else _dev(source_params, torch.float32).reshape(-1)
        _check(lib.wos_scene_set_source(self.handle, int(source_type), _p(sp), 0 if sp is None else sp.numel(),
                                        _stream()))
        self.__dict__["_source_type"] = int(source_type)
        self.__dict__["_source_params"] = sp

    @property
    def normals(self):
        out = torch.empty(self._vertices.shape, dtype=torch.float32, device=self._vertices.device)
        _check(lib.wos_scene_get_normals(self.handle, _p(out), _stream()))
        return out

    @property
    def aabb(self):
        lo = (ctypes.c_float * 3)()
        hi = (ctypes.c_float * 3)()
        _check(lib.wos_scene_get_info(self.handle, lo, hi, None, None, None, None))
        D = self._dim
        return AABB(torch.tensor(list(lo)[:D]), torch.tensor(list(hi)[:D]))

    @property
    def max_depth(self):
        d = ctypes.c_int()
        _check(lib.wos_scene_get_info(self.handle, None, None, None, None, None, ctypes.byref(d)))
        return d.value

    @property
    def use_bvh(self):
        return self._use_bvh

    @use_bvh.setter
    def use_bvh(self, flag):
        _check(lib.wos_scene_set_use_bvh(self.handle, int(bool(flag))))
        self.__dict__["_use_bvh"] = bool(flag)

    def set_grid(self, cell_factor=1.0, band_cells=16, max_cand=48):
        """Rebuilds the 2D near-boundary candidate grid (cell_factor <= 0
        removes it).  Query results are identical with and without it."""
        _check(lib.wos_scene_set_grid(self.handle, float(cell_factor), int(band_cells), int(max_cand), _stream()))

    def grid_info(self):
        nx, ny, nc = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint()
        cell = ctypes.c_float()
        _check(lib.wos_scene_get_grid(self.handle, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(cell),
                                      ctypes.byref(nc)))
        return {"nx": nx.value, "ny": ny.value, "cell": cell.value, "entries": nc.value}

    def set_start_grid(self, resolution=None):
        """Rebuilds the start-node grid with `resolution` cells along the
        longest side (default: two cells per mean segment length in 2D, at
        most 2048, as scenes are built; 96 in 3D, where scenes are built
        without one; 0 removes it).  Query results are identical with and
        without it."""
        if resolution is None:
            resolution = -1
        _check(lib.wos_scene_set_start_grid(self.handle, int(resolution), _stream()))

    def start_grid_info(self):
        nx, ny, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib.wos_scene_get_start_grid(self.handle, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz)))
        return {"nx": nx.value, "ny": ny.value, "nz": nz.value}

    def bvh_topology(self):
        n = self._indices.shape[0]
        dv = self._vertices.device
        prim = torch.empty(n, dtype=torch.int32, device=dv)
        codes = torch.empty(n, dtype=torch.int32, device=dv)
        ch = torch.empty(max(n - 1, 1), 2, dtype=torch.int32, device=dv)
        _check(lib.wos_scene_get_bvh(self.handle, _p(prim), _p(codes), _p(ch), _stream()))
        return prim, codes, ch[: max(n - 1, 0)]

    # ---- queries --------------------------------------------------------------
    def _points(self, p):
        return _dev(p, torch.float32, self._dim, "p")

    def closest_point_preliminary(self, p):
        """Scene::closestPointPreliminary (include/scene.cuh:497-537)."""
        q = self._points(p)
        out = torch.empty(q.shape[0], dtype=torch.int32, device=q.device)
        _check(lib.wos_closest_point_index(self.handle, _p(q), q.shape[0], _p(out), _vp(None), _stream()))
        return out

    def closest_point_index_d2(self, p):
        q = self._points(p)
        out = torch.empty(q.shape[0], dtype=torch.int32, device=q.device)
        d2 = torch.empty(q.shape[0], dtype=torch.float32, device=q.device)
        _check(lib.wos_closest_point_index(self.handle, _p(q), q.shape[0], _p(out), _p(d2), _stream()))
        return out, d2

    def sdf_native(self, p):
        """SceneDevice::sdf (include/scene.cuh:148-178)."""
        q = self._points(p)
        out = torch.empty(q.shape[0], dtype=torch.float32, device=q.device)
        _check(lib.wos_sdf(self.handle, _p(q), q.shape[0], _p(out), _stream()))
        return out

    def closest_point(self, p):
        """Scene::closestPoint (include/scene.cuh:539-657)."""
        q = self._points(p)
        rec, out, valid = _alloc_record(q.shape[0], self._dim)
        _check(lib.wos_closest_point(self.handle, _p(q), q.shape[0], ctypes.byref(rec), _stream()))
        return _finish_record(out, valid)

    def largest_inscribed_ball(self, its, epsilon=1e-3):
        """Scene::largestInscribedBall (include/scene.cuh:659-704)."""
        rec, keep = its._struct(self._dim)
        n = _dev(its.p, torch.float32, self._dim).shape[0]
        out = torch.empty(n, dtype=torch.float32, device=device())
        _check(lib.wos_largest_inscribed_ball(self.handle, ctypes.byref(rec), n, float(epsilon), _p(out), _stream()))
        del keep
        return out

    def dirichlet(self, cRec):
        """Scene::dirichlet (include/scene.cuh:455-468)."""
        rec, keep = cRec._struct(self._dim)
        n = _dev(cRec.prim_id, torch.int32).numel()
        out = torch.empty(n, dtype=torch.float32, device=device())
        _check(lib.wos_dirichlet(self.handle, ctypes.byref(rec), n, _p(out), _stream()))
        del keep
        return out


class Scene(_SceneBase):
    """wos_ext.Scene (src/python/scene.cu:28-47): 2D polyline scene."""
    _dim = 2

    @staticmethod
    def from_vertices(vertices, values):
        """Scene::fromVertices (include/scene.cuh:344-352), closing the polyline
        with segment (n-1, 0) as wos/scene.py:206-209 does."""
        v = _dev(vertices, torch.float32, 2)
        n = v.shape[0]
        i = torch.arange(n, dtype=torch.int32, device=v.device)
        return Scene(v, torch.stack([i, (i + 1) % n], dim=1), values)

    @property
    def source_type(self):
        return self._source_type

    @property
    def c_source_params(self):
        return self._source_params


class Scene3D(_SceneBase):
    """wos_ext.Scene3D (src/python/scene.cu:68-82): triangle-mesh scene."""
    _dim = 3

    def __init__(self, vertices=None, indices=None, values=None, use_bvh=False):
        super().__init__(vertices, indices, values, use_bvh)

    @staticmethod
    def from_vertices(vertices, values):
        raise WoSError("not implemented")   # include/scene.cuh:349-350


# ---- walks --------------------------------------------------------------------

class WoS:
    """wos_ext.WoS (src/python/wos.cu:5-25, include/wos.cuh:155-372)."""
    _dim = 2

    def __init__(self, nwalks, nsteps=32, epsilon=1e-3, double_sided=False, prevent_fd_artifacts=False,
                 use_IS_for_greens=True):
        self.nwalks = int(nwalks)
        self.nsteps = int(nsteps)
        self.epsilon = float(epsilon)
        self.double_sided = bool(double_sided)
        self.prevent_fd_artifacts = bool(prevent_fd_artifacts)
        self.use_IS_for_greens = bool(use_IS_for_greens)

    @staticmethod
    def create(nwalks, nsteps=32, epsilon=1e-3, double_sided=False, prevent_fd_artifacts=False,
               use_IS_for_greens=True):
        return WoS(nwalks, nsteps, epsilon, double_sided, prevent_fd_artifacts, use_IS_for_greens)

    def _params(self, nwalks=None):
        return _WalkParams(self.nwalks if nwalks is None else int(nwalks), self.nsteps, self.epsilon,
                           int(self.double_sided), int(self.prevent_fd_artifacts), int(self.use_IS_for_greens))

    def single_walk(self, p, scene, sampler):
        """WoS::singleWalk: one walk per row of p; sampler.state / sampler.inc
        (int64 tensors holding uint64) are advanced in place."""
        q = _dev(p, torch.float32, scene.dim)
        n = q.shape[0]
        if sampler.state.numel() != n:
            raise WoSError(f"sampler has {sampler.state.numel()} streams for {n} points")
        for name in ("state", "inc"):
            t = getattr(sampler, name)
            if t.dtype != torch.int64 or not t.is_cuda or not t.is_contiguous():
                setattr(sampler, name, _dev(t, torch.int64).reshape(-1))
        rec, out, valid = _alloc_record(n, scene.dim)
        wp = self._params()
        _check(lib.wos_single_walk(scene.handle, ctypes.byref(wp), _p(q), n, _p(sampler.state), _p(sampler.inc),
                                   ctypes.byref(rec), _stream()))
        return _finish_record(out, valid)

    def walk_reduce(self, p, scene, seed=0, *, nwalks=None, records=False, basis=False, stats=False, chunk=0,
                    noalloc=False):
        """Solver.walk_reduce (wos/solver.py:49-65) over this walk, fused into
        one persistent launch.  Returns a dict with ``mean`` (float32) and, on
        request, ``mean64``/``var64`` (stats), ``rec_prim``/``rec_t``
        (records), ``basis`` (float64, K x P: per point the sum over its
        walks of d value / d coeff_k, bit-reproducible) and ``walk_value``.  ``noalloc`` runs the launch
        through wos_walk_reduce_noalloc, whose scratch is a static per-GPU
        array (same results; the library allocates nothing)."""
        q = _dev(p, torch.float32, scene.dim)
        P = q.shape[0]
        W = self.nwalks if nwalks is None else int(nwalks)
        dv = q.device
        out = {"mean": torch.empty(P, dtype=torch.float32, device=dv)}
        out["walk_value"] = torch.empty(P * W, dtype=torch.float32, device=dv)
        if stats:
            out["mean64"] = torch.empty(P, dtype=torch.float64, device=dv)
            out["var64"] = torch.empty(P, dtype=torch.float64, device=dv)
        if records:
            out["rec_prim"] = torch.empty(P * W, dtype=torch.int32, device=dv)
            out["rec_t"] = torch.empty(P * W * (scene.dim - 1), dtype=torch.float32, device=dv)
        if basis:
            K = scene.c_source_params.numel() // 4 if scene.source_type == 3 else 0
            if K == 0:
                raise WoSError("basis output needs a Fourier source (source_type 3)")
            out["basis"] = torch.empty(K * P, dtype=torch.float64, device=dv)
        wp = self._params(W)
        g = out.get
        fn = lib.wos_walk_reduce_noalloc if noalloc else lib.wos_walk_reduce
        _check(fn(scene.handle, ctypes.byref(wp), _p(q), P, int(seed), int(chunk), _p(g("mean")), _p(g("mean64")),
                  _p(g("var64")), _p(g("walk_value")), _p(g("rec_prim")), _p(g("rec_t")), _p(g("basis")), _stream()))
        return out

    def walk_counts(self, p, scene, seed=0, *, nwalks=None, basis=True):
        """Instrumented walk_reduce: totals of AABB tests, primitive tests,
        closest-primitive queries and walk steps of one launch, and the mean
        fraction of a wave's lanes at work in its step phases (step_lanes)
        and node-visit iterations (query_lanes)."""
        q = _dev(p, torch.float32, scene.dim)
        P = q.shape[0]
        W = self.nwalks if nwalks is None else int(nwalks)
        wv = torch.empty(P * W, dtype=torch.float32, device=q.device)
        cnt = torch.zeros(8, dtype=torch.int64, device=q.device)
        wp = self._params(W)
        _check(lib.wos_walk_reduce_counts(scene.handle, ctypes.byref(wp), _p(q), P, int(seed), int(bool(basis)),
                                          _p(wv), _p(cnt), _stream()))
        c = [int(x) for x in cnt.cpu()]
        return dict(box_tests=c[0], prim_tests=c[1], queries=c[2], steps=c[3], walks=P * W,
                    step_lanes=c[4] / max(c[5], 1) / 64, query_lanes=c[6] / max(c[7], 1) / 64,
                    step_phases=c[5], query_iterations=c[7])

    def solve(self, p, scene, seed=0):
        """WoS::solve (include/wos.cuh:320-364): mean of nwalks walks per point.
        All nwalks walks are run (the reference drops nwalks mod 100)."""
        return self.walk_reduce(p, scene, seed)["mean"]


class WoS3D(WoS):
    """wos_ext.WoS3D (src/python/wos.cu:21)."""
    _dim = 3

    @staticmethod
    def create(nwalks, nsteps=32, epsilon=1e-3, double_sided=False, prevent_fd_artifacts=False,
               use_IS_for_greens=True):
        return WoS3D(nwalks, nsteps, epsilon, double_sided, prevent_fd_artifacts, use_IS_for_greens)


def adjoint_values(scene, n_points, nwalks, rec_prim, rec_t, grad_out, basis=None, want_values=True,
                   want_coeffs=False):
    """Reverse mode of walk_reduce w.r.t. vertex values / Fourier coefficients."""
    dv = device()
    go = _dev(grad_out, torch.float32).reshape(-1)
    gv = torch.zeros(scene.c_values.numel(), dtype=torch.float32, device=dv) if want_values else None
    K = scene.c_source_params.numel() // 4 if (want_coeffs and scene.source_type == 3) else 0
    gc = torch.zeros(K, dtype=torch.float32, device=dv) if K else None
    _check(lib.wos_adjoint_values(scene.handle, int(n_points), int(nwalks), _p(rec_prim), _p(rec_t),
                                  _p(basis if K else None), _p(go), _p(gv), _p(gc), _stream()))
    return gv, gc


def adjoint_source(scene, cwos, points, seed, grad_out, *, nwalks=None):
    """Reverse mode of ``cwos.walk_reduce(points, scene, seed)`` w.r.t. every
    source parameter (wos_adjoint_source): the walks are re-run from their
    keys.  Returns a float32 tensor shaped like the scene's source
    parameters (zeros for a scene without a source)."""
    q = _dev(points, torch.float32, scene.dim)
    P = q.shape[0]
    sp = scene.c_source_params
    n = 0 if (sp is None or not scene.source_type) else sp.numel()
    out = torch.zeros(max(n, 1), dtype=torch.float32, device=q.device)
    if n:
        go = _dev(grad_out, torch.float32).reshape(-1)
        if go.numel() != P:
            raise WoSError(f"grad_out has {go.numel()} entries for {P} points")
        wp = cwos._params(nwalks)
        _check(lib.wos_adjoint_source(scene.handle, ctypes.byref(wp), _p(q), P, int(seed), _p(go), _p(out),
                                      _stream()))
    return out[:n].reshape(sp.shape) if n else out[:0]


def debug_fail_workspace(n):
    """Testing hook: the next n traversal-spill requests fail."""
    _check(lib.wos_debug_fail_workspace(int(n)))


def shape_gradient(scene, points, seed, grad_out, *, nwalks, nsteps, epsilon, epsilon2, clamping,
                   double_sided=False, control_variates=True, sample_grad_u=False):
    """OursCUDA / Ours3D backward onto vertex positions: dL/d vertices (summed
    in 64-bit fixed point, so bit-reproducible) and, with sample_grad_u, the
    per-sample spatial gradient estimates."""
    q = _dev(points, torch.float32, scene.dim)
    P = q.shape[0]
    gp = _GradParams(int(nwalks), int(nsteps), float(epsilon), float(epsilon2), float(clamping),
                     int(bool(double_sided)), int(bool(control_variates)))
    gv = torch.zeros(scene.c_vertices.shape, dtype=torch.float32, device=q.device)
    go = None if grad_out is None else _dev(grad_out, torch.float32).reshape(-1)
    sg = torch.empty(P * int(nwalks), scene.dim, dtype=torch.float32, device=q.device) if sample_grad_u else None
    _check(lib.wos_shape_gradient(scene.handle, ctypes.byref(gp), _p(q), P, int(seed), _p(go),
                                  _p(gv if go is not None else None), _p(sg), _stream()))
    return gv, sg


def shape_gradient_counts(scene, points, seed, *, nwalks, nsteps, epsilon, epsilon2, clamping, double_sided=False,
                          control_variates=True):
    """Instrumented shape-gradient launch: totals of bound tests, primitive
    tests, closest-primitive queries and walk steps of its samples."""
    q = _dev(points, torch.float32, scene.dim)
    gp = _GradParams(int(nwalks), int(nsteps), float(epsilon), float(epsilon2), float(clamping),
                     int(bool(double_sided)), int(bool(control_variates)))
    cnt = torch.zeros(4, dtype=torch.int64, device=q.device)
    _check(lib.wos_shape_gradient_counts(scene.handle, ctypes.byref(gp), _p(q), q.shape[0], int(seed), _p(cnt),
                                         _stream()))
    c = [int(x) for x in cnt.cpu()]
    return dict(box_tests=c[0], prim_tests=c[1], queries=c[2], steps=c[3], samples=q.shape[0] * int(nwalks))


def shape_derivative(scene, points, seed, velocity, *, nwalks, nsteps, epsilon, epsilon2, clamping,
                     double_sided=False, control_variates=True, stats=False):
    """Forward mode of the OursCUDA / Ours3D integrand: per point, the
    derivative of its estimate when vertex v moves with velocity[v]
    (wos_shape_derivative).  Returns the float32 mean, and with stats=True
    also (mean64, var64)."""
    q = _dev(points, torch.float32, scene.dim)
    P = q.shape[0]
    vel = _dev(velocity, torch.float32, scene.dim)
    if vel.shape[0] != scene.c_vertices.shape[0]:
        raise ValueError(f"velocity needs one row per vertex ({scene.c_vertices.shape[0]}), got {vel.shape[0]}")
    gp = _GradParams(int(nwalks), int(nsteps), float(epsilon), float(epsilon2), float(clamping),
                     int(bool(double_sided)), int(bool(control_variates)))
    mean = torch.empty(P, dtype=torch.float32, device=q.device)
    m64 = torch.empty(P, dtype=torch.float64, device=q.device) if stats else None
    v64 = torch.empty(P, dtype=torch.float64, device=q.device) if stats else None
    _check(lib.wos_shape_derivative(scene.handle, ctypes.byref(gp), _p(q), P, int(seed), _p(vel), _p(mean),
                                    _p(m64), _p(v64), _stream()))
    return (mean, m64, v64) if stats else mean


class Batch:
    """Many (scene, walk settings, points, seed) items walked by one
    persistent launch per dimension (``wos_batch_run``).  Item j's outputs
    are ``outputs[j]`` (``mean``, ``mean64``, ``var64``, ``walk_value`` and,
    on request, ``rec_prim``/``rec_t``/``basis``); each equals what
    ``WoS.walk_reduce`` returns for that item alone.

    Reverse pass: an item with ``grad_values=True`` (and/or, for a Fourier
    source, ``grad_coeffs=True``) also gets ``grad_out`` (dL/d mean, filled
    by the caller after ``run()``) and ``grad_values`` / ``grad_coeffs``
    buffers that ``adjoint()`` adds into (``zero_grad()`` clears them); the
    result equals ``adjoint_values`` on that item's records.  An item with
    ``grad_vertices=True`` and ``grad_params`` (the OursCUDA / Ours3D
    settings: nwalks, nsteps, epsilon, epsilon2, clamping, double_sided,
    control_variates) gets a ``grad_vertices`` buffer that
    ``shape_gradient()`` adds dL/d vertex positions into from the same
    ``grad_out``, equal to ``shape_gradient`` on that item alone bit for bit
    (both sum in the same 64-bit fixed point).  ``run()``, ``zero_grad()``, ``adjoint()``
    and ``shape_gradient()`` allocate nothing, so a whole primal + loss +
    adjoint step can be captured with ``torch.cuda.graph``.  The
    batch holds its scenes' device buffers: changing an item's source, grids
    or use_bvh afterwards makes ``run()`` raise (create the batch again)."""

    def __init__(self, items):
        self._h = _vp(None)
        self.outputs, self._keep = [], []
        arr = (_BatchItem * len(items))()
        for j, it in enumerate(items):
            scene, cw = it["scene"], it["wos"]
            q = _dev(it["points"], torch.float32, scene.dim)
            P, W = q.shape[0], cw.nwalks
            dv = q.device
            want_v = bool(it.get("grad_values"))
            want_c = bool(it.get("grad_coeffs"))
            want_s = bool(it.get("grad_vertices"))
            gpar = dict(it.get("grad_params") or {})
            if want_s and "nwalks" not in gpar:
                raise WoSError(f"item {j}: grad_vertices needs grad_params with at least nwalks")
            gp = _GradParams(int(gpar.get("nwalks", 0)), int(gpar.get("nsteps", 0)), float(gpar.get("epsilon", 0.0)),
                             float(gpar.get("epsilon2", 0.0)), float(gpar.get("clamping", 0.0)),
                             int(bool(gpar.get("double_sided", False))), int(bool(gpar.get("control_variates", True))))
            o = {"mean": torch.empty(P, dtype=torch.float32, device=dv),
                 "mean64": torch.empty(P, dtype=torch.float64, device=dv),
                 "var64": torch.empty(P, dtype=torch.float64, device=dv),
                 "walk_value": torch.empty(P * W, dtype=torch.float32, device=dv)}
            if it.get("records") or want_v or want_c:
                o["rec_prim"] = torch.empty(P * W, dtype=torch.int32, device=dv)
                o["rec_t"] = torch.empty(P * W * (scene.dim - 1), dtype=torch.float32, device=dv)
            if it.get("basis") or want_c:
                if scene.source_type != 3:
                    raise WoSError(f"item {j}: basis / coefficient gradients need a Fourier source (source_type 3)")
                K = scene.c_source_params.numel() // 4
                o["basis"] = torch.empty(K * P, dtype=torch.float64, device=dv)
            if want_v or want_c or want_s:
                o["grad_out"] = torch.zeros(P, dtype=torch.float32, device=dv)
            if want_s:
                o["grad_vertices"] = torch.zeros(scene.c_vertices.shape, dtype=torch.float32, device=dv)
            if want_v:
                o["grad_values"] = torch.zeros(scene.c_values.numel(), dtype=torch.float32, device=dv)
            if want_c:
                o["grad_coeffs"] = torch.zeros(scene.c_source_params.numel() // 4, dtype=torch.float32, device=dv)
            self.outputs.append(o)
            self._keep += [q, scene]
            g = o.get
            arr[j] = _BatchItem(scene.handle, cw._params(), q.data_ptr(), P, int(it.get("seed", 0)),
                                g("mean").data_ptr(), g("mean64").data_ptr(), g("var64").data_ptr(),
                                g("walk_value").data_ptr(), _p(g("rec_prim")), _p(g("rec_t")), _p(g("basis")),
                                _p(g("grad_out")), _p(g("grad_values")), _p(g("grad_coeffs")), gp,
                                _p(g("grad_vertices")))
        h = _vp()
        torch.cuda.synchronize()
        _check(lib.wos_batch_create(arr, len(items), ctypes.byref(h)))
        self._h = h

    def run(self):
        _check(lib.wos_batch_run(self._h, _stream()))
        return self.outputs

    def zero_grad(self):
        """Clears every item's grad_values / grad_coeffs (fill kernels, so it
        can sit in a captured graph)."""
        for o in self.outputs:
            for k in ("grad_values", "grad_coeffs", "grad_vertices"):
                if k in o:
                    o[k].zero_()

    def adjoint(self):
        """Reverse pass of the last run: adds each item's dL/d values and dL/d
        coefficients, from its ``grad_out``, into its gradient buffers."""
        _check(lib.wos_batch_adjoint(self._h, _stream()))
        return self.outputs

    def shape_gradient(self):
        """Shape-gradient pass: adds each item's dL/d vertex positions, from
        its ``grad_out``, into its ``grad_vertices`` (one persistent launch
        per dimension)."""
        _check(lib.wos_batch_shape_gradient(self._h, _stream()))
        return self.outputs

    def autotune(self, key=None, reps=2):
        """Chooses the launch tuning for this batch (wos_ext.autotune);
        outputs are unchanged by it.  `key` names the workload in the
        tuning cache (default: the batch's item count and walk total)."""
        from .autotune import autotune
        if key is None:
            key = f"batch-{len(self.outputs)}-{sum(o['walk_value'].numel() for o in self.outputs)}"
        return autotune(self.run, key, reps=reps)

    def __del__(self):
        h = self.__dict__.get("_h")
        if h is not None and h.value:
            try:
                lib.wos_batch_destroy(h)
            except Exception:   # interpreter shutdown: module globals already cleared
                pass
            self.__dict__["_h"] = None
