"""3D shape-gradient integrators (wos/wos_grad_3d.py).  ``Ours3D`` runs in
the fused persistent kernel; ``Baseline3D`` (soft-edge walk differentiated
through the attached scene queries), ``Baseline23D`` (spatial gradient near
the exit) and ``WoSGradient3D`` are the reference's loop integrators on
torch autograd."""
from __future__ import annotations

from dataclasses import dataclass

import torch

from wos.solver import ShapeSolverMixin, Solver
from wos.wos3d import WoS3D, WoS3DCUDA
from wos.wos_grad import _inside, _velocity_term


@dataclass
class Ours3D(ShapeSolverMixin, Solver):
    """wos/wos_grad_3d.py:104-137."""
    epsilon: float = 1e-3
    nwalks: int = 10
    nsteps: int = 32
    double_sided: bool = False
    epsilon2: float = 1e-3
    clamping: float = 1e-1


@dataclass
class WoSGradient3D(Solver):
    """wos/wos_grad_3d.py:10-20: holds a native 3D walk."""
    epsilon: float = 1e-3
    nwalks: int = 10
    nsteps: int = 32
    double_sided: bool = False

    def __post_init__(self):
        self.wos = WoS3DCUDA(nwalks=self.nwalks, nsteps=self.nsteps, epsilon=self.epsilon,
                             double_sided=self.double_sided)


@dataclass
class Baseline3D(Solver):
    """wos/wos_grad_3d.py:23-57: soft-edge weights min(|d|, eps)/eps."""

    def rand_on_sphere(self, sampler):
        return WoS3D.rand_on_sphere(self, sampler)

    def weight(self, d):
        return torch.clamp(d.abs(), max=self.epsilon) / self.epsilon

    def single_walk(self, p, scene, sampler):
        d = scene.sdf_attached(p)
        active = torch.ones_like(d, dtype=torch.bool) if self.double_sided else d.detach() < 0
        result = torch.zeros_like(d)
        beta = active.to(d.dtype)
        d = d.abs()
        for _ in range(self.nsteps):
            w = self.weight(d)
            its = scene.closest_point_attached(p)
            result = result + torch.where(d.detach() > 0, beta * (1.0 - w) * scene.dirichlet(its),
                                          torch.zeros_like(d))
            beta = beta * w
            p = p + d[:, None] * self.rand_on_sphere(sampler)
            d = scene.sdf_attached(p).abs()
        return torch.where(active, result, torch.zeros_like(result))


@dataclass
class Baseline23D(Solver):
    """wos/wos_grad_3d.py:60-101."""
    l: float = 5e-3
    normal_derivative_only: bool = False

    def __post_init__(self):
        self.wos = WoS3DCUDA(nwalks=self.nwalks, nsteps=self.nsteps, epsilon=self.epsilon,
                             double_sided=self.double_sided)

    def single_walk(self, p, scene, sampler):
        inside = _inside(scene, p, self.double_sided)
        with torch.no_grad():
            its = self.wos.single_walk_preliminary(p, scene, sampler)
        its = scene.get_point(its)
        p_in_shell = its.p.detach() + its.n * self.l
        with torch.no_grad():
            g = self.wos.grad(p_in_shell, scene, sampler)
            if self.normal_derivative_only:
                gn = g[0] * its.n[:, 0] + g[1] * its.n[:, 1] + g[2] * its.n[:, 2]
                gt = self.wos.tangent_derivative(its, scene, sampler)
                g = tuple(gn * its.n[:, k] + gt[k] for k in range(3))
        return _velocity_term(its, list(g), inside & its.valid.bool())
