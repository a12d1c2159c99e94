"""3D Laplace walk-on-spheres (wos/wos3d.py)."""
from __future__ import annotations

from dataclasses import dataclass

import wos_ext
from wos.wos import WoS


@dataclass
class WoS3D(WoS):
    """wos/wos3d.py:9-156 (CUDA walk semantics)."""
    epsilon: float = 1e-3
    nwalks: int = 10
    nsteps: int = 32
    double_sided: bool = False

    def __post_init__(self):
        self.cwos = wos_ext.WoS3D(nwalks=self.nwalks, nsteps=self.nsteps, epsilon=self.epsilon,
                                  double_sided=self.double_sided, prevent_fd_artifacts=self.prevent_fd_artifacts)

    def tangent_derivative(self, its, scene, sampler=None):
        """wos/wos3d.py:137-156."""
        import torch
        f = scene.indices[its.prim_id.clamp(min=0).long()].long()
        a, b, c = scene.vertices[f[:, 0]], scene.vertices[f[:, 1]], scene.vertices[f[:, 2]]
        va, vb, vc = scene.values[f[:, 0]], scene.values[f[:, 1]], scene.values[f[:, 2]]
        ab, ac = b - a, c - a
        t = ab / ab.norm(dim=1, keepdim=True)
        fn = torch.cross(ab, ac, dim=1)
        fn = fn / fn.norm(dim=1, keepdim=True)
        n = torch.cross(fn, t, dim=1)
        n = n / n.norm(dim=1, keepdim=True)
        gt = (vb - va) / ab.norm(dim=1)
        gn = ((vc - va) - gt * (ac * t).sum(1)) / (ac * n).sum(1)
        g = gt[:, None] * t + gn[:, None] * n
        return g[:, 0], g[:, 1], g[:, 2]


@dataclass
class WoS3DCUDA(WoS3D):
    """wos/wos3d.py:159-178."""
