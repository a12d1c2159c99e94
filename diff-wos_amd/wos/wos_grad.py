"""2D shape-gradient integrators (wos/wos_grad.py).  Each returns a
zero-valued walk value whose gradient w.r.t. the scene's vertices is the
shape-derivative estimate, -grad u . V with V the boundary velocity of the
exit point (``scene.get_point`` re-gathers it from the attached vertices).

``OursCUDA`` runs in the fused persistent kernel (primary walk, largest
inscribed ball, two antithetic secondary walks, tangential derivative).
``Ours``, ``Baseline2`` and ``Baseline2CUDA`` are the reference's loop
integrators, differentiated by torch autograd through ``get_point``.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from wos.solver import ShapeSolverMixin, Solver
from wos.wos import WoS, WoSCUDA


def _velocity_term(its, grad_u, mask):
    v = its.p - its.p.detach()
    res = -sum(g * v[:, k] for k, g in enumerate(grad_u))
    return torch.where(mask, res, torch.zeros_like(res))


def _inside(scene, p, double_sided):
    with torch.no_grad():
        d = scene.sdf(p.detach())
    return torch.ones_like(d, dtype=torch.bool) if double_sided else d < 0


@dataclass
class OursCUDA(ShapeSolverMixin, Solver):
    """wos/wos_grad.py:113-151."""
    epsilon2: float = 1e-3
    clamping: float = 1e-1
    control_variates: bool = True


@dataclass
class Ours(Solver):
    """wos/wos_grad.py:57-86: the loop walk (projected exits) and the loop
    normal derivative, both at ``epsilon``."""

    def __post_init__(self):
        self.wos = WoS(nsteps=self.nsteps, nwalks=self.nwalks, epsilon=self.epsilon, double_sided=self.double_sided)

    def single_walk(self, p, scene, sampler):
        inside = _inside(scene, p, self.double_sided)
        with torch.no_grad():
            its = self.wos.single_walk_preliminary(p, scene, sampler)
        its = scene.get_point(its)
        with torch.no_grad():
            gn = self.wos.normal_derivative(its, scene, sampler)
            gt = self.wos.tangent_derivative(its, scene, sampler)
        return _velocity_term(its, [gn[0] + gt[0], gn[1] + gt[1]], inside & its.valid.bool())


@dataclass
class Baseline2(Solver):
    """wos/wos_grad.py:15-54: the native walk's exit nudged a distance l
    along the record normal, and the spatial gradient estimated there (all
    of it, or its normal part plus the tangential derivative)."""
    epsilon: float = 1e-3
    nwalks: int = 10
    nsteps: int = 32
    variance_reduction: bool = False
    l: float = 5e-3
    normal_derivative_only: bool = False

    def __post_init__(self):
        self.wos = WoSCUDA(nsteps=self.nsteps, nwalks=self.nwalks, epsilon=self.epsilon,
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
                gn = (g[0] * its.n[:, 0] + g[1] * its.n[:, 1])
                gt = self.wos.tangent_derivative(its, scene, sampler)
                g = (gn * its.n[:, 0] + gt[0], gn * its.n[:, 1] + gt[1])
        return _velocity_term(its, list(g), inside & its.valid.bool())


@dataclass
class Baseline2CUDA(Solver):
    """wos/wos_grad.py:89-110: the native walk, the spatial gradient 5e-3
    inside the exit point, no inside test (the reference's class reads a
    ``wos`` attribute it never sets; it is set here).  The reference steps
    to p - 5e-3 n, taking the record normal for the outward one (its comment
    at wos/wos_grad.py:101); exit records' normals point to the side the
    walk came from, so the point inside is p + 5e-3 n."""

    def __post_init__(self):
        self.wos = WoSCUDA(nsteps=self.nsteps, nwalks=self.nwalks, epsilon=self.epsilon,
                           double_sided=self.double_sided)

    def single_walk(self, p, scene, sampler):
        with torch.no_grad():
            its = self.wos.single_walk_preliminary(p, scene, sampler)
        its = scene.get_point(its)
        p_in_shell = its.p.detach() + its.n * 5e-3
        with torch.no_grad():
            g = self.wos.grad(p_in_shell, scene, sampler)
        return _velocity_term(its, list(g), its.valid.bool())
