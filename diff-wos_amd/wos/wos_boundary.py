"""Boundary-sampling term of the shape derivative (wos/wos_boundary.py):
points sampled on the boundary by length / area, scattered into detector
pixels, with value 0 and, through the attached sample position, gradient
-g (x . n) w.r.t. the vertices -- the images of the reference's renderD."""
from __future__ import annotations

from dataclasses import dataclass

import torch

from wos.fwd import PCG32
from wos.solver import Solver


@dataclass
class WoSBoundary(Solver):
    """wos/wos_boundary.py:11-31."""

    def walk_detector(self, scene, detector, seed=0):
        npoints = detector.res[0] * detector.res[1]
        n = npoints * self.nwalks
        sampler = PCG32(n, seed)
        valid, idx, value = self.sample_boundary(scene, detector, sampler)
        out = torch.zeros(npoints, dtype=value.dtype, device=value.device)
        out = out.index_add(0, idx[valid].long(), value[valid])
        return out / detector.size() / self.nwalks

    def sample_boundary(self, scene, detector, sampler):
        b = scene.sample_boundary(sampler)
        val = self.boundary_term(b) / b.pdf
        valid, idx = detector.index(b.p.detach())
        return valid, idx, val

    def boundary_term(self, b):
        xn = (b.p * b.n.detach()).sum(1)
        return -b.val.detach() * (xn - xn.detach())
