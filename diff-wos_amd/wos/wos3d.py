"""3D Laplace walk-on-spheres (wos/wos3d.py).

``WoS3D`` is the reference's loop estimator (wos/wos3d.py:9-156): |sdf|
steps with directions uniform on the sphere, and walks that never reach the
epsilon shell are evaluated at the closest point of the origin (the
reference's ``p_in_shell`` starts at 0); plus its spatial gradient, normal
derivative and tangential derivative.  ``WoS3DCUDA`` runs the native walk
(wos/wos3d.py:159-178).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch

import wos_ext
from wos.solver import Solver, ValueSolverMixin
from wos.wos import _sel


@dataclass
class WoS3D(Solver):
    """wos/wos3d.py:9-156."""
    epsilon: float = 1e-3
    nwalks: int = 10
    nsteps: int = 32
    double_sided: bool = False

    def rand_on_sphere(self, sampler):
        """wos/wos3d.py:16-22: z = 1 - 2 u0, azimuth 2 pi u1."""
        u0 = sampler.next_float64()
        u1 = sampler.next_float64()
        z = 1.0 - 2.0 * u0
        r = torch.sqrt(torch.clamp(1.0 - z * z, min=0.0))
        th = 2.0 * math.pi * u1
        return torch.stack([r * torch.cos(th), r * torch.sin(th), z], dim=1).float()

    def rand_on_sphere2(self, sampler):
        """wos/wos3d.py:24-32: (cos theta, phi)."""
        u0 = sampler.next_float64()
        u1 = sampler.next_float64()
        return (1.0 - 2.0 * u0).float(), (2.0 * math.pi * u1).float()

    def single_walk_preliminary(self, p, scene, sampler):
        """wos/wos3d.py:34-51."""
        p = p.detach().float().clone()
        n = p.shape[0]
        active = torch.ones(n, dtype=torch.bool, device=p.device)
        p_in_shell = torch.zeros_like(p)
        d = scene.sdf(p).abs()
        for _ in range(self.nsteps):
            in_shell = active & (d < self.epsilon)
            p_in_shell = torch.where(in_shell[:, None], p, p_in_shell)
            active &= ~in_shell
            step = self.rand_on_sphere(sampler)
            idx = active.nonzero().squeeze(1)
            if idx.numel():
                p[idx] = p[idx] + step[idx] * d[idx, None]
                d[idx] = scene.sdf(p[idx]).abs()
        return scene.closest_point(p_in_shell)

    def single_walk(self, p, scene, sampler):
        """wos/wos3d.py:53-61."""
        d = scene.sdf(p.detach())
        active = torch.ones_like(d, dtype=torch.bool) if self.double_sided else d < 0
        its = self.single_walk_preliminary(p, scene, sampler)
        its.valid = its.valid.bool() & active
        return scene.dirichlet(its)

    def u(self, p, scene, sampler):
        return self.single_walk(p, scene, sampler)

    def grad(self, p, scene, sampler):
        """wos/wos3d.py:66-95: grad u at p from the first ball (eq. 13 of
        Sawhney and Crane 2020), antithetic, control variate u(x)."""
        x = p.detach().float()
        R = scene.sdf(x)
        active = torch.ones_like(R, dtype=torch.bool) if self.double_sided else R < 0
        R = R.abs()
        active &= ~(R < self.epsilon)
        ret = torch.zeros_like(x)
        u_x = scene.dirichlet(scene.closest_point(x))
        direction = self.rand_on_sphere(sampler)
        for i in range(2):
            if i == 1:
                direction = -direction
            y = x + R[:, None] * direction
            G = 3.0 / R[:, None] * (y - x) / R[:, None]
            u = self.u(y, scene, sampler) - u_x
            u = _sel(active, u, torch.zeros_like(u))
            G = _sel(active[:, None], G, torch.zeros_like(G))
            ret = ret + G * u[:, None]
        ret = ret / 2.0
        return ret[:, 0], ret[:, 1], ret[:, 2]

    def normal_derivative(self, its, scene, sampler, clamping=1e-1):
        """wos/wos3d.py:97-135."""
        n = -its.n.float()
        p = its.p.float()
        u_ref = scene.dirichlet(its)
        R = scene.largest_inscribed_ball(its.c_object())
        c = p - n * R[:, None]
        cos_t, phi = self.rand_on_sphere2(sampler)
        cos_t = cos_t.clamp(-1.0, 1.0 - clamping)
        up = torch.where((n[:, 2].abs() < 0.9)[:, None],
                         torch.tensor([0.0, 0.0, 1.0], device=n.device).expand_as(n),
                         torch.tensor([1.0, 0.0, 0.0], device=n.device).expand_as(n))
        p_dir = torch.cross(n, up, dim=1)
        p_dir = p_dir / p_dir.norm(dim=1, keepdim=True)
        q_dir = torch.cross(n, p_dir, dim=1)
        grad = torch.zeros_like(u_ref)
        valid = its.valid.bool()
        for i in range(2):
            if i == 1:
                phi = phi + math.pi
            r = torch.sqrt(torch.clamp(1.0 - cos_t * cos_t, min=0.0))
            y = c + R[:, None] * (cos_t[:, None] * n + r[:, None] * (torch.cos(phi)[:, None] * p_dir +
                                                                      torch.sin(phi)[:, None] * q_dir))
            d = scene.sdf(y).abs()
            # an invalid record's queries are masked off in the reference: u = 0
            u_walk = self.u(y, scene, sampler)
            u = _sel(d < self.epsilon, scene.dirichlet(scene.closest_point(y)), u_walk)
            u = _sel(valid, u, torch.zeros_like(u))
            P = -1.0 / (math.sqrt(2.0) * torch.pow(1.0 - cos_t, 1.5))
            grad = grad + P * (u - u_ref) / R
        grad = grad / 2.0
        return grad * n[:, 0], grad * n[:, 1], grad * n[:, 2]

    def tangent_derivative(self, its, scene, sampler=None):
        """wos/wos3d.py:137-156: in-plane gradient of the linear interpolant."""
        f = scene.indices[its.prim_id.clamp(min=0).long()].long()
        a, b, c = scene.vertices[f[:, 0]], scene.vertices[f[:, 1]], scene.vertices[f[:, 2]]
        va, vb, vc = scene.values[f[:, 0]], scene.values[f[:, 1]], scene.values[f[:, 2]]
        ab, ac = b - a, c - a
        t = ab / ab.norm(dim=1, keepdim=True)
        fn = torch.cross(ab, ac, dim=1)
        fn = fn / fn.norm(dim=1, keepdim=True)
        nn = torch.cross(fn, t, dim=1)
        nn = nn / nn.norm(dim=1, keepdim=True)
        gt = (vb - va) / ab.norm(dim=1)
        gn = ((vc - va) - gt * (ac * t).sum(1)) / (ac * nn).sum(1)
        g = gt[:, None] * t + gn[:, None] * nn
        return g[:, 0], g[:, 1], g[:, 2]


@dataclass
class WoS3DCUDA(ValueSolverMixin, WoS3D):
    """wos/wos3d.py:159-178: the native walk (non-terminated walks invalid)."""
    prevent_fd_artifacts: bool = False

    def __post_init__(self):
        self.cwos = wos_ext.WoS3D(nwalks=self.nwalks, nsteps=self.nsteps, epsilon=self.epsilon,
                                  double_sided=self.double_sided, prevent_fd_artifacts=self.prevent_fd_artifacts)

    def _cwos(self):
        self.cwos.nwalks, self.cwos.nsteps, self.cwos.epsilon = self.nwalks, self.nsteps, self.epsilon
        self.cwos.double_sided = self.double_sided
        self.cwos.prevent_fd_artifacts = self.prevent_fd_artifacts
        return self.cwos

    def single_walk_preliminary(self, p, scene, sampler):
        return self._cwos().single_walk(p.detach(), scene, sampler)
