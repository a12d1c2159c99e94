"""2D scenes (wos/scene.py): Detector, ClosestPointRecord, Polyline.

``Polyline`` subclasses ``wos_ext.Scene`` exactly as the reference subclasses
its pybind class (wos/scene.py:192); its queries run in the HIP kernels.
The tensors handed to the constructor are kept as given (``vertices``,
``values``, ``source_params``): when they require grad, the solvers'
autograd functions return gradients to them.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

import wos_ext
from wos.fwd import Array2, Array2i, Float, device
from wos.utils import DiscreteDistribution


@dataclass
class Detector:
    """wos/scene.py:9-42: pixel grid on [vmin, vmax] (y flipped)."""
    vmin: tuple = (-1.0, -1.0)
    vmax: tuple = (1.0, 1.0)
    res: tuple = (100, 100)

    def size(self):
        return float(np.prod(np.subtract(self.vmax, self.vmin)))

    def make_points(self):
        x = torch.linspace(self.vmin[0], self.vmax[0], self.res[0], device=device())
        y = torch.linspace(self.vmax[1], self.vmin[1], self.res[1], device=device())
        Y, X = torch.meshgrid(y, x, indexing="ij")
        return torch.stack([X.reshape(-1), Y.reshape(-1)], dim=1)

    def make_jittered_points(self, sampler, spp):
        p = self.make_points().repeat_interleave(spp, dim=0)
        d = torch.tensor(np.subtract(self.vmax, self.vmin) / np.asarray(self.res), dtype=torch.float32,
                         device=p.device)
        jx = sampler.next_float64().float() - 0.5
        jy = sampler.next_float64().float() - 0.5
        return p + torch.stack([jx, jy], dim=1) * d

    def index(self, p):
        """wos/scene.py:18-24: pixel of p, rows flipped.  The reference's
        row (res_y - uv.y) puts uv.y = 0 one row past the image; such samples
        are marked invalid here rather than scattered out of bounds."""
        vmin = torch.tensor(self.vmin, device=p.device)
        vmax = torch.tensor(self.vmax, device=p.device)
        uv = (p - vmin) / (vmax - vmin) * torch.tensor(self.res, device=p.device) + 0.5
        uv = uv.to(torch.int32)
        valid = (uv[:, 0] >= 0) & (uv[:, 0] < self.res[0]) & (uv[:, 1] >= 0) & (uv[:, 1] < self.res[1])
        idx = uv[:, 0] + (self.res[1] - uv[:, 1]) * self.res[0]
        valid = valid & (idx >= 0) & (idx < self.res[0] * self.res[1])
        return valid, torch.where(valid, idx, torch.zeros_like(idx))

    def save(self, image, filename):
        np.save(str(filename), np.asarray(image))


@dataclass
class ClosestPointRecord:
    """wos/scene.py:65-82."""
    valid: torch.Tensor = None
    p: torch.Tensor = None
    n: torch.Tensor = None
    t: torch.Tensor = None
    val: torch.Tensor = None
    prim_id: torch.Tensor = None
    J: torch.Tensor = None
    contrib: torch.Tensor = None

    def c_object(self):
        return wos_ext.ClosestPointRecord(self.valid, self.p, self.n, self.t, self.prim_id, self.contrib)


@dataclass
class BoundarySamplingRecord:
    """wos/scene.py:86-92: p attached (to the vertices), n detached, the
    one-sided Dirichlet value, primitive, parameter and pdf (per length or
    area)."""
    p: torch.Tensor = None
    n: torch.Tensor = None
    val: torch.Tensor = None
    prim_id: torch.Tensor = None
    t: torch.Tensor = None
    pdf: torch.Tensor = None


class NativeFields:
    """``vertices``, ``values`` and ``source_params`` of a scene class over
    the native scene.  In the reference they are the native scene's
    def_readwrite fields (src/python/scene.cu:35-47), so assigning one
    changes what the walks use: here assigning ``values`` or
    ``source_params`` copies the new tensor into the native scene
    (wos_scene_set_values / wos_scene_set_source), and ``vertices``, from
    which the BVH is built, cannot be reassigned (build a new scene).  The
    tensors are kept as given, so solvers' gradients reach them."""

    @property
    def vertices(self):
        return self.__dict__.get("_t_vertices")

    @vertices.setter
    def vertices(self, v):
        if self.__dict__.get("_t_vertices") is not None:
            raise wos_ext.WoSError("a scene's vertices are fixed when it is built (the BVH is built from them); "
                                   "build a new scene for new vertices")
        self.__dict__["_t_vertices"] = v

    @property
    def values(self):
        return self.__dict__.get("_t_values")

    @values.setter
    def values(self, v):
        if not isinstance(v, torch.Tensor):
            v = Float(v)
        self.__dict__["_t_values"] = v
        h = self.__dict__.get("_h")
        if h is not None and h.value:
            self.set_values(v)

    @property
    def source_params(self):
        return self.__dict__.get("_t_source_params")

    @source_params.setter
    def source_params(self, v):
        if v is not None and not isinstance(v, torch.Tensor):
            v = Float(v)
        self.__dict__["_t_source_params"] = v
        h = self.__dict__.get("_h")
        if h is not None and h.value and self.source_type:
            if v is None:
                raise wos_ext.WoSError("a scene with a source needs source_params")
            self.set_source(self.source_type, v)

    def swap_fields(self, vertices, values, source_params):
        """Puts tensors equal in value to the current fields in their place
        (e.g. detached copies that require grad) without touching the native
        scene; returns the previous ones for the matching restore."""
        d = self.__dict__
        saved = (d.get("_t_vertices"), d.get("_t_values"), d.get("_t_source_params"))
        d["_t_vertices"], d["_t_values"] = vertices, values
        if source_params is not None:
            d["_t_source_params"] = source_params
        return saved

    def restore_fields(self, saved):
        d = self.__dict__
        d["_t_vertices"], d["_t_values"], d["_t_source_params"] = saved


class Polyline(NativeFields, wos_ext.Scene):
    """wos/scene.py:192-475 over the native scene.  ``use_bvh`` defaults to
    True: the LBVH query returns the same primitive as the brute-force scan
    (equal squared distances resolve to the smallest index in both), so the
    flag only selects speed; ``use_bvh=False`` runs the reference's scan."""

    def __init__(self, vertices, indices, values, source_type=0, source_params=None, use_bvh=True):
        self.vertices = Array2(vertices) if not isinstance(vertices, torch.Tensor) else vertices
        self.indices = Array2i(indices)
        self.values = Float(values) if not isinstance(values, torch.Tensor) else values
        self.source_params = source_params
        if source_params is not None and not isinstance(source_params, torch.Tensor):
            self.source_params = Float(source_params)
        super().__init__(self.vertices, self.indices, self.values, use_bvh, source_type, self.source_params)
        self.init_dist(self.vertices, self.indices)

    def init_dist(self, vertices, indices):
        """wos/scene.py:275-279: segments sampled proportionally to length."""
        f = indices.long()
        self.edge_distr = DiscreteDistribution((vertices[f[:, 1]] - vertices[f[:, 0]]).detach().norm(dim=1))

    def sample_boundary(self, sampler):
        """wos/scene.py:435-449: a point uniform in arc length, attached to
        the vertices; n = (-ab.y, ab.x) normalised and detached."""
        rnd = sampler.next_float32()
        prim, t, prob = self.edge_distr.sample_reuse_pmf(rnd)
        f = self.indices[prim].long()
        a, b = self.vertices[f[:, 0]], self.vertices[f[:, 1]]
        ab = b - a
        val = torch.lerp(self.values[f[:, 0]], self.values[f[:, 1]], t)
        x = a + ab * t[:, None]
        n = torch.stack([-ab[:, 1], ab[:, 0]], dim=1).detach()
        n = n / n.norm(dim=1, keepdim=True)
        return BoundarySamplingRecord(p=x, n=n, val=val, prim_id=prim.int(), t=rnd, pdf=prob / ab.norm(dim=1))

    def J(self, prim_id):
        """wos/scene.py:409-414: l / detach(l) of the segment."""
        f = self.indices[prim_id.clamp(min=0).long()].long()
        l = (self.vertices[f[:, 1]] - self.vertices[f[:, 0]]).norm(dim=1)
        return l / l.detach()

    def geometric(self, x, y, n):
        """wos/scene.py:401-407: cos(angle(y - x, n)) / (2 pi |y - x|)."""
        d = y - x
        r = d.norm(dim=1)
        return ((d / r[:, None]) * n).sum(1) / (2 * math.pi * r)

    @property
    def source_type(self):
        return self._source_type

    @staticmethod
    def from_vertices(vertices, values):
        v = Array2(vertices)
        n = v.shape[0]
        i = torch.arange(n, dtype=torch.int32, device=v.device)
        return Polyline(v, torch.stack([i, (i + 1) % n], dim=1), values)

    @staticmethod
    def from_obj(filename, normalize=True, flip_orientation=False):
        """wos/scene.py:211-240 (`v`, `l`, `c` lines)."""
        vertices, indices, values = [], [], []
        with open(filename) as f:
            for line in f:
                line = line.strip()
                if line.startswith("v "):
                    vertices.append([float(x) for x in line[2:].split()][:2])
                elif line.startswith("l "):
                    l = [int(x) - 1 for x in line[2:].split()]
                    indices.append(l[::-1] if flip_orientation else l)
                elif line.startswith("c "):
                    values.append(float(line[2:].split()[0]))
        v = np.asarray(vertices, dtype=np.float64)
        if normalize:
            v = v - v.mean(axis=0)
            v = v / np.linalg.norm(v, axis=1).max()
        return Polyline(Array2(v.astype(np.float32)), Array2i(np.asarray(indices)), Float(np.asarray(values)))

    # ---- queries (the reference's Dr.Jit versions, wos/scene.py:281-357) -----
    def sdf(self, p):
        return self.sdf_native(p)

    def dirichlet(self, its):
        """wos/scene.py:353-357: lerp of the segment's vertex values at t (0
        for invalid records), attached to ``values``."""
        f = self.indices[its.prim_id.clamp(min=0).long()].long()
        out = torch.lerp(self.values[f[:, 0]], self.values[f[:, 1]], its.t.float())
        if its.valid is not None:
            out = torch.where(its.valid.bool(), out, torch.zeros_like(out))
        return out

    def closest_point(self, p):
        r = super().closest_point(p)
        rec = ClosestPointRecord(valid=r.valid, p=r.p, n=r.n, t=r.t, prim_id=r.prim_id, contrib=r.contrib)
        rec.val = self.dirichlet(rec)
        return rec

    # ---- queries attached to the vertex tensor (for autograd estimators) -----
    def _segment(self, p):
        """Closest segment of each point (selected by the kernels, detached)
        and the clamped projection parameter recomputed from ``vertices``."""
        idx = self.closest_point_preliminary(p.detach()).long()
        f = self.indices[idx.clamp(min=0)].long()
        a, b = self.vertices[f[:, 0]], self.vertices[f[:, 1]]
        ba = b - a
        h = (((p - a) * ba).sum(1) / (ba * ba).sum(1)).clamp(0.0, 1.0)
        return idx, a, ba, h

    def sdf_attached(self, p):
        """wos/scene.py:325-339 with the distance differentiable in p and the
        vertices (the sign, piecewise constant, comes from the kernels)."""
        idx, a, ba, h = self._segment(p)
        dist = (p - a - ba * h[:, None]).norm(dim=1)
        sgn = torch.where(self.sdf(p.detach()) >= 0, 1.0, -1.0)
        return sgn * dist

    def closest_point_attached(self, p):
        """wos/scene.py:303-323 with p and t differentiable."""
        idx, a, ba, h = self._segment(p)
        r = super().closest_point(p.detach())
        rec = ClosestPointRecord(valid=idx >= 0, p=a + ba * h[:, None], n=r.n, t=h, prim_id=idx.int(),
                                 contrib=r.contrib)
        rec.val = self.dirichlet(rec)
        return rec

    def source_function(self, p):
        """include/scene.cuh:278-301 and the Fourier type (torch, float32)."""
        q = self.source_params
        x, y = p[:, 0], p[:, 1]
        st = self.source_type
        if st == 1:
            return q[4] * torch.exp(-(x - q[0]) ** 2 / (2 * q[2] ** 2) - (y - q[1]) ** 2 / (2 * q[3] ** 2))
        if st == 2:
            return q[3] * torch.sin(q[0] * x + q[1] * y + q[2]) + q[4]
        if st == 3:
            c = q.reshape(-1, 4)
            return (c[:, 3] * torch.sin(x[:, None] * c[:, 0] + y[:, None] * c[:, 1] + c[:, 2])).sum(-1)
        return torch.zeros_like(x)

    def get_point(self, its):
        """wos/scene.py:416-433: hit point re-gathered from (attached) vertices."""
        f = self.indices[its.prim_id.clamp(min=0).long()].long()
        t = its.t.detach()
        p0, p1 = self.vertices[f[:, 0]], self.vertices[f[:, 1]]
        v0, v1 = self.values[f[:, 0]], self.values[f[:, 1]]
        p = p0 + (p1 - p0) * t[:, None]
        return ClosestPointRecord(valid=its.valid, p=p, n=its.n, t=t, val=v0 + (v1 - v0) * t,
                                  prim_id=its.prim_id, J=torch.ones_like(t))

    def tangent_derivative(self, its):
        """wos/scene.py:451-460."""
        f = self.indices[its.prim_id.clamp(min=0).long()].long()
        v0, v1 = self.vertices[f[:, 0]], self.vertices[f[:, 1]]
        dv = (self.values[f[:, 1]] - self.values[f[:, 0]]) / (v1 - v0).norm(dim=1)
        d_t = (v1 - v0) / (v1 - v0).norm(dim=1, keepdim=True)
        return dv * d_t[:, 0], dv * d_t[:, 1]

    def write_obj(self, filename):
        with open(filename, "w") as f:
            for v in self.vertices.detach().cpu().numpy():
                f.write("v %f %f\n" % (v[0], v[1]))
            for i in self.indices.cpu().numpy():
                f.write("l %d %d\n" % (i[0] + 1, i[1] + 1))
            for c in self.values.detach().cpu().numpy():
                f.write("c %f\n" % c)
