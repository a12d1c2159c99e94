"""3D scenes (wos/scene3d.py): Detector3D, ClosestPointRecord3D, Scene3D."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

import wos_ext
from wos.fwd import Array3, Array3i, Float, device
from wos.scene import BoundarySamplingRecord, Detector, NativeFields
from wos.utils import DiscreteDistribution


@dataclass
class Detector3D(Detector):
    """wos/scene3d.py:9-46: a z = const slab of pixels."""
    res: tuple = (256, 256)
    z: float = 0.05

    def __post_init__(self):
        self.l_z = (self.vmax[0] - self.vmin[0]) / self.res[0]

    def size(self):
        return float(np.prod(np.subtract(self.vmax, self.vmin))) * self.l_z

    def index(self, p):
        """wos/scene3d.py:35-43: the 2D pixel of (x, y) for points within
        half a pixel depth of the slab."""
        inside_z = (p[:, 2] - float(self.z)).abs() < (self.l_z / 2.0)
        valid, idx = super().index(p[:, :2])
        return valid & inside_z, idx

    def make_points(self):
        p2 = super().make_points()
        return torch.cat([p2, torch.full((p2.shape[0], 1), float(self.z), device=p2.device)], dim=1)

    def make_jittered_points(self, sampler, spp):
        p = self.make_points().repeat_interleave(spp, dim=0)
        d = torch.tensor(np.subtract(self.vmax, self.vmin) / np.asarray(self.res), dtype=torch.float32,
                         device=p.device)
        off = torch.stack([sampler.next_float64().float(), sampler.next_float64().float()], dim=1) * d - d / 2
        return p + torch.cat([off, torch.zeros_like(off[:, :1])], dim=1)


@dataclass
class ClosestPointRecord3D:
    """wos/scene3d.py:87-103."""
    valid: torch.Tensor = None
    p: torch.Tensor = None
    n: torch.Tensor = None
    uv: torch.Tensor = None
    val: torch.Tensor = None
    prim_id: torch.Tensor = None
    J: torch.Tensor = None
    contrib: torch.Tensor = None

    def c_object(self):
        return wos_ext.ClosestPointRecord3D(self.valid, self.p, self.n, self.uv, self.prim_id, self.contrib)


class Scene3D(NativeFields, wos_ext.Scene3D):
    """wos/scene3d.py:106-274 over the native triangle-mesh scene (LBVH on by
    default; the flag selects speed only, see wos.scene.Polyline)."""

    def __init__(self, vertices, indices, values, use_bvh=True):
        self.vertices = vertices if isinstance(vertices, torch.Tensor) else Array3(vertices)
        self.indices = Array3i(indices)
        self.values = values if isinstance(values, torch.Tensor) else Float(values)
        self.source_params = None
        super().__init__(self.vertices, self.indices, self.values, use_bvh)
        self.init_dist(self.vertices, self.indices)

    def init_dist(self, vertices, indices):
        """wos/scene3d.py:150-155: triangles sampled proportionally to area."""
        f = indices.long()
        a, b, c = vertices[f[:, 0]], vertices[f[:, 1]], vertices[f[:, 2]]
        self.edge_distr = DiscreteDistribution(torch.cross(b - a, c - a, dim=1).detach().norm(dim=1))

    def sample_boundary(self, sampler):
        """wos/scene3d.py:241-262: a triangle by area, a point uniform on it
        (u = sqrt(r1)(1 - r2), v = sqrt(r1) r2), attached to the vertices;
        n = -(b - a) x (c - a) normalised, detached."""
        rnd = sampler.next_float32()
        prim, _, prob = self.edge_distr.sample_reuse_pmf(rnd)
        f = self.indices[prim].long()
        a, b, c = self.vertices[f[:, 0]], self.vertices[f[:, 1]], self.vertices[f[:, 2]]
        va, vb, vc = self.values[f[:, 0]], self.values[f[:, 1]], self.values[f[:, 2]]
        cr = torch.cross(b - a, c - a, dim=1)
        area = cr.norm(dim=1)
        r1 = sampler.next_float32()
        r2 = sampler.next_float32()
        u = torch.sqrt(r1) * (1 - r2)
        v = torch.sqrt(r1) * r2
        p = b * u[:, None] + c * v[:, None] + a * (1 - u - v)[:, None]
        val = vb * u + vc * v + va * (1 - u - v)
        n = -(cr / area[:, None]).detach()
        return BoundarySamplingRecord(p=p, n=n, val=val, prim_id=prim.int(), t=torch.stack([u, v], dim=1),
                                      pdf=prob / area)

    @property
    def source_type(self):
        return 0

    @staticmethod
    def from_obj(filename, normalize=True, flip_orientation=False):
        """wos/scene3d.py:116-144 (`v`, `f`, `c` lines), keeping all three
        coordinates (the reference keeps only x, y at :127)."""
        vertices, indices, values = [], [], []
        with open(filename) as f:
            for line in f:
                line = line.strip()
                if line.startswith("v "):
                    vertices.append([float(x) for x in line[2:].split()][:3])
                elif line.startswith("f "):
                    l = [int(x.split("/")[0]) - 1 for x in line[2:].split()]
                    indices.append(l[::-1] if flip_orientation else l)
                elif line.startswith("c "):
                    values.append(float(line[2:].split()[0]))
        v = np.asarray(vertices, dtype=np.float64)
        if normalize:
            v = v - v.mean(axis=0)
            v = v / np.linalg.norm(v, axis=1).max()
        return Scene3D(Array3(v.astype(np.float32)), Array3i(np.asarray(indices)), Float(np.asarray(values)))

    def sdf(self, p):
        return self.sdf_native(p)

    def dirichlet(self, its):
        """wos/scene3d.py:216-223: b u + c v + a (1 - u - v) (0 for invalid
        records), attached to ``values``."""
        f = self.indices[its.prim_id.clamp(min=0).long()].long()
        va, vb, vc = self.values[f[:, 0]], self.values[f[:, 1]], self.values[f[:, 2]]
        u, v = its.uv[:, 0].float(), its.uv[:, 1].float()
        out = vb * u + vc * v + va * (1 - u - v)
        if its.valid is not None:
            out = torch.where(its.valid.bool(), out, torch.zeros_like(out))
        return out

    def closest_point(self, p):
        r = super().closest_point(p)
        rec = ClosestPointRecord3D(valid=r.valid, p=r.p, n=r.n, uv=r.uv, prim_id=r.prim_id, contrib=r.contrib)
        rec.val = self.dirichlet(rec)
        return rec

    # ---- queries attached to the vertex tensor (for autograd estimators) -----
    def _triangle(self, p):
        from wos.utils import closest_point_triangle
        idx = self.closest_point_preliminary(p.detach()).long()
        f = self.indices[idx.clamp(min=0)].long()
        a, b, c = self.vertices[f[:, 0]], self.vertices[f[:, 1]], self.vertices[f[:, 2]]
        pt, uv, d = closest_point_triangle(p, a, b, c)
        return idx, pt, uv, d

    def sdf_attached(self, p):
        """Signed distance differentiable in p and the vertices (sign from
        the kernels)."""
        idx, pt, uv, d = self._triangle(p)
        sgn = torch.where(self.sdf(p.detach()) >= 0, 1.0, -1.0)
        return sgn * d

    def closest_point_attached(self, p):
        """wos/scene3d.py:179-199 with p and uv differentiable."""
        idx, pt, uv, d = self._triangle(p)
        r = super().closest_point(p.detach())
        rec = ClosestPointRecord3D(valid=idx >= 0, p=pt, n=r.n, uv=uv, prim_id=idx.int(), contrib=r.contrib)
        rec.val = self.dirichlet(rec)
        return rec

    def get_point(self, its):
        """wos/scene3d.py:225-239."""
        f = self.indices[its.prim_id.clamp(min=0).long()].long()
        uv = its.uv.detach()
        a, b, c = self.vertices[f[:, 0]], self.vertices[f[:, 1]], self.vertices[f[:, 2]]
        u, v = uv[:, 0:1], uv[:, 1:2]
        p = b * u + c * v + a * (1 - u - v)
        va, vb, vc = self.values[f[:, 0]], self.values[f[:, 1]], self.values[f[:, 2]]
        val = vb * uv[:, 0] + vc * uv[:, 1] + va * (1 - uv[:, 0] - uv[:, 1])
        return ClosestPointRecord3D(valid=its.valid, p=p, n=its.n, uv=uv, val=val, prim_id=its.prim_id)

    def write_obj(self, filename):
        with open(filename, "w") as f:
            for v in self.vertices.detach().cpu().numpy():
                f.write(f"v {v[0]} {v[1]} {v[2]}\n")
            for i in self.indices.cpu().numpy():
                f.write(f"f {i[0] + 1} {i[1] + 1} {i[2] + 1}\n")
            for v in self.values.detach().cpu().numpy():
                f.write(f"c {v}\n")
