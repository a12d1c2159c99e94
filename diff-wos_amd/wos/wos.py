"""2D Laplace walk-on-spheres solvers (wos/wos.py).

``WoS`` is the reference's loop estimator (wos/wos.py:10-133): every walk of
the batch advances one step per pass, with the scene queries answered by the
HIP kernels and the arithmetic in torch; walks that do not reach the
epsilon shell within ``nsteps`` are projected to their closest boundary
point.  ``WoSCUDA`` swaps the walk for the native one (wos/wos.py:136-155):
non-terminated walks are invalid, and ``walk`` runs the fused persistent
launch.  ``Baseline`` is the soft-edge estimator differentiated by autograd
through the scene (wos/wos.py:158-187).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch

import wos_ext
from wos.solver import Solver, ValueSolverMixin


def _sel(mask, a, b):
    return torch.where(mask, a, b)


@dataclass
class WoS(Solver):
    """wos/wos.py:10-133."""
    epsilon: float = 1e-3
    nwalks: int = 10
    nsteps: int = 32
    double_sided: bool = False

    def rand_in_disk(self, sampler):
        """wos/wos.py:21-24: r = sqrt(u0), angle = 2 pi u1."""
        r = torch.sqrt(sampler.next_float64())
        angle = sampler.next_float64() * (2.0 * math.pi)
        return torch.stack([r * torch.cos(angle), r * torch.sin(angle)], dim=1).float()

    def single_walk_preliminary(self, p, scene, sampler):
        """wos/wos.py:26-44.  The step uses the signed distance, as the
        reference does; the direction draw advances every walk's stream."""
        p = p.detach().float().clone()
        d = scene.sdf(p)
        active = torch.ones(p.shape[0], dtype=torch.bool, device=p.device)
        for _ in range(self.nsteps):
            active &= ~(d.abs() < self.epsilon)
            step = self.rand_on_circle(sampler)
            idx = active.nonzero().squeeze(1)
            if idx.numel():
                p[idx] = p[idx] + d[idx, None] * step[idx]
                d[idx] = scene.sdf(p[idx])
        return scene.closest_point(p)

    def single_walk(self, p, scene, sampler):
        """wos/wos.py:46-54: g at the exit point, 0 outside / invalid."""
        d = scene.sdf(p.detach())
        active = torch.ones_like(d, dtype=torch.bool) if self.double_sided else d < 0
        its = self.single_walk_preliminary(p, scene, sampler)
        its.valid = its.valid.bool() & active
        v = scene.dirichlet(its)
        return _sel(its.valid, v, torch.zeros_like(v))

    def u(self, p, scene, sampler):
        return self.single_walk(p, scene, sampler)

    def normal_derivative(self, its, scene, sampler, override_R=None, clamping=1e-1, control_variates=True):
        """wos/wos.py:59-107: off-centred Poisson-kernel estimator of du/dn on
        the largest inscribed ball, two antithetic samples."""
        n = -its.n.float()
        p = its.p.float()
        u_ref = scene.dirichlet(its)
        R = scene.largest_inscribed_ball(its.c_object()) if override_R is None else override_R
        if not isinstance(R, torch.Tensor):
            R = torch.full_like(u_ref, float(R))
        c = p - n * R[:, None]
        theta = self.sample_uniform(sampler).clamp(clamping, 2 * math.pi - clamping)
        grad = torch.zeros_like(u_ref)
        for i in range(2):
            if i == 1:
                theta = -theta
            p_dir = torch.stack([-n[:, 1], n[:, 0]], dim=1)
            y = c + R[:, None] * (n * torch.cos(theta)[:, None] + p_dir * torch.sin(theta)[:, None])
            d = scene.sdf(y).abs()
            u = _sel(d < self.epsilon, scene.dirichlet(scene.closest_point(y)), self.single_walk(y, scene, sampler))
            P = 1.0 / (torch.cos(theta) - 1.0)
            grad = grad + (P * (u - u_ref) / R if control_variates else P * u / R)
        grad = grad / 2.0
        return grad * n[:, 0], grad * n[:, 1]

    def tangent_derivative(self, its, scene, sampler=None):
        """wos/wos.py:109-129."""
        return scene.tangent_derivative(its)

    def grad(self, p, scene, sampler):
        """Spatial gradient at p from the first ball: the boundary term of
        wos/wos_with_source.py:109-152 (antithetic, control variate u(x)).
        The reference's 2D Laplace class has no ``grad`` although its
        Baseline2 integrators call one; this is that estimator without a
        source."""
        x = p.detach().float()
        R = scene.sdf(x)
        active = torch.ones_like(R, dtype=torch.bool) if self.double_sided else R < 0
        R = R.abs()
        active &= ~(R < self.epsilon)
        theta = self.sample_uniform(sampler)
        ret = torch.zeros_like(x)
        u_x = scene.dirichlet(scene.closest_point(x))
        for i in range(2):
            if i == 1:
                theta = -theta
            y = x + R[:, None] * torch.stack([torch.cos(theta), torch.sin(theta)], dim=1)
            G = 2.0 / R[:, None] * (y - x) / R[:, None]
            u = self.u(y, scene, sampler) - u_x
            u = _sel(active, u, torch.zeros_like(u))
            G = _sel(active[:, None], G, torch.zeros_like(G))
            ret = ret + G * u[:, None]
        ret = ret / 2.0
        return ret[:, 0], ret[:, 1]


@dataclass
class WoSCUDA(ValueSolverMixin, WoS):
    """wos/wos.py:136-155: the native walk.  ``walk`` / ``walk_reduce`` run
    the fused persistent launch (same streams as single walks on a keyed
    sampler); ``single_walk_preliminary`` runs one native walk per row on a
    caller's sampler."""
    prevent_fd_artifacts: bool = False

    def __post_init__(self):
        self.cwos = wos_ext.WoS(nwalks=self.nwalks, nsteps=self.nsteps, epsilon=self.epsilon,
                                double_sided=self.double_sided, prevent_fd_artifacts=self.prevent_fd_artifacts)

    def _cwos(self):
        self.cwos.nwalks, self.cwos.nsteps, self.cwos.epsilon = self.nwalks, self.nsteps, self.epsilon
        self.cwos.double_sided = self.double_sided
        self.cwos.prevent_fd_artifacts = self.prevent_fd_artifacts
        return self.cwos

    def single_walk_preliminary(self, p, scene, sampler):
        """wos/wos.py:152-155: one native walk per row, sampler advanced."""
        return self._cwos().single_walk(p.detach(), scene, sampler)


@dataclass
class Baseline(Solver):
    """wos/wos.py:158-187: soft-edge walk, every step weighted by
    min(|d|, eps)/eps, differentiated by autograd through the attached scene
    queries (distances and closest points recomputed from the scene's
    vertex tensor on the primitives the kernels select)."""
    epsilon: float = 0.005
    nwalks: int = 10
    nsteps: int = 32

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
            result = result + _sel(d.detach() > 0, beta * (1.0 - w) * scene.dirichlet(its), torch.zeros_like(d))
            beta = beta * w
            p = p + d[:, None] * self.rand_on_circle(sampler)
            d = scene.sdf_attached(p).abs()
        return _sel(active, result, torch.zeros_like(result))
