"""2D Poisson walk-on-spheres (wos/wos_with_source.py), -Lap u = f.

``WoSWithSource`` is the reference's loop estimator (wos/wos_with_source.py:
8-232): each step adds f(y) G(y) / pdf(y) for one point y of the current
ball, drawn uniformly or -- ``use_IS_for_greens`` -- by the radial density
r log(1/r) tabulated at 10^4 midpoints (the pmf of the reference's
``G_distr``); plus its spatial-gradient estimator (``grad``) and the normal
derivative with the interior term.  ``WoSWithSourceCUDA`` runs the native
walk, which samples the ball uniformly (include/wos.cuh:115-139; the
importance-sampled branch of the CUDA walk is compiled out there too).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch

import wos_ext
from wos.solver import ValueSolverMixin
from wos.utils import DiscreteDistribution
from wos.wos import WoS, _sel


@dataclass
class WoSWithSource(WoS):
    """wos/wos_with_source.py:8-232."""
    use_IS_for_greens: bool = False
    control_variates: bool = True
    source_epsilon: float = 0.0

    def __post_init__(self):
        N = 10000
        step = 1.0 / N
        r = torch.linspace(0.0, 1.0 - step, N, dtype=torch.float32) + 0.5 * step
        G = torch.log(1.0 / r) / (2.0 * math.pi)
        self.step = step
        self._pmf = r * G
        self.G_distr = None   # built on the walk's device on first use

    def _distr(self, dev):
        if self.G_distr is None or self.G_distr.pmf.device != dev:
            self.G_distr = DiscreteDistribution(self._pmf.to(dev))
        return self.G_distr

    def eval_G(self, p, R, r_clamp=1e-4):
        """G(0, p) of the disk of radius R (wos/wos_with_source.py:29-31)."""
        r = p.norm(dim=1)
        return torch.log(R / torch.clamp(r, min=r_clamp)) / (2.0 * math.pi)

    def eval_P(self, center, y, s, R):
        """Poisson kernel of the disk (wos/wos_with_source.py:38-45)."""
        y = y - center
        s = s - center
        r = y.norm(dim=1)
        c = (y * s).sum(1) / s.norm(dim=1)
        denom = R * R - 2 * R * c + r * r
        return (R * R - r * r) / torch.clamp(denom, min=1e-5) / (2.0 * math.pi * R)

    def sample_G(self, R, sampler):
        """wos/wos_with_source.py:47-54: (offset, pdf) of one point of the
        disk of radius R drawn with density proportional to G."""
        rnd = sampler.next_float32()
        idx, t, pmf = self._distr(rnd.device).sample_reuse_pmf(rnd)
        radius = (t + idx.float()) * self.step
        pdf = pmf / (R * self.step)
        phi = 2.0 * math.pi * sampler.next_float32()
        pdf = pdf / (2.0 * math.pi)
        return torch.stack([radius * R * torch.cos(phi), radius * R * torch.sin(phi)], dim=1), pdf

    def single_walk_preliminary(self, p, scene, sampler):
        """wos/wos_with_source.py:56-97."""
        p = p.detach().float().clone()
        n = p.shape[0]
        y = p.clone()
        d = scene.sdf(p)
        result = torch.zeros(n, dtype=torch.float32, device=p.device)
        active = torch.ones(n, dtype=torch.bool, device=p.device)
        zero = torch.zeros_like(result)
        for _ in range(self.nsteps):
            active &= ~(d.abs() < self.epsilon)
            if self.use_IS_for_greens:
                dy, pdf = self.sample_G(d.abs(), sampler)
                y = _sel(active[:, None], dy, y)
                weight = dy.norm(dim=1)
            else:
                dy = self.rand_in_disk(sampler)
                pdf = 1.0 / (math.pi * d * d)
                y = _sel(active[:, None], dy * d[:, None], y)
                weight = torch.ones_like(result)
            G = self.eval_G(y, d.abs())
            G = _sel(active, G, zero)
            G = _sel(torch.isfinite(G), G, zero)
            pdf = _sel(active, pdf, zero)
            active &= pdf > 1e-5
            y = _sel(active[:, None], y + p, y)
            f = _sel(active, scene.source_function(y), zero)
            result = _sel(active, result + f * G * weight / pdf, result)
            step = self.rand_on_circle(sampler)
            idx = active.nonzero().squeeze(1)
            if idx.numel():
                p[idx] = p[idx] + d[idx, None] * step[idx]
                d[idx] = scene.sdf(p[idx])
        its = scene.closest_point(p)
        its.contrib = result
        return its

    def single_walk(self, p, scene, sampler):
        """wos/wos_with_source.py:99-107: g(exit) + source contribution."""
        d = scene.sdf(p.detach())
        active = torch.ones_like(d, dtype=torch.bool) if self.double_sided else d < 0
        its = self.single_walk_preliminary(p, scene, sampler)
        its.valid = its.valid.bool() & active
        v = scene.dirichlet(its) + its.contrib
        return _sel(its.valid, v, torch.zeros_like(v))

    def grad(self, p, scene, sampler):
        """wos/wos_with_source.py:109-152: spatial gradient at p -- boundary
        term on the first ball (antithetic, control variate u(x)) plus the
        interior term of the source."""
        x = p.detach().float()
        R = scene.sdf(x)
        active = torch.ones_like(R, dtype=torch.bool) if self.double_sided else R < 0
        R = R.abs()
        active &= ~(R < self.epsilon)
        theta = self.sample_uniform(sampler)
        ret = [torch.zeros_like(R), torch.zeros_like(R)]
        u_x = scene.dirichlet(scene.closest_point(x))
        for i in range(2):
            if i == 1:
                theta = -theta
            y = x + R[:, None] * torch.stack([torch.cos(theta), torch.sin(theta)], dim=1)
            G = 2.0 / R[:, None] * (y - x) / R[:, None]
            u = self.u(y, scene, sampler) - u_x
            u = _sel(active, u, torch.zeros_like(u))
            G = _sel(active[:, None], G, torch.zeros_like(G))
            ret[0] = ret[0] + G[:, 0] * u
            ret[1] = ret[1] + G[:, 1] * u
        ret = [r / 2.0 for r in ret]
        y = x + R[:, None] * self.rand_in_disk(sampler)
        r = (y - x).norm(dim=1)
        grad_G = (y - x) / (2.0 * math.pi) * (1.0 / (r * r) - 1.0 / (R * R))[:, None]
        f = scene.source_function(y) * math.pi * R * R
        return ret[0] + f * grad_G[:, 0], ret[1] + f * grad_G[:, 1]

    def normal_derivative(self, its, scene, sampler, override_R=None, clamping=1e-1, antithetic=True,
                          control_variates=True, ball_ratio=1.0, source_epsilon=0.0):
        """wos/wos_with_source.py:154-232: boundary term on the largest ball
        plus the source's interior term (with its control variate and the
        finite-difference fallback within source_epsilon of the boundary)."""
        n = -its.n.float()
        p = its.p.float()
        u_ref = scene.dirichlet(its)
        R = scene.largest_inscribed_ball(its.c_object()) if override_R is None else override_R
        if not isinstance(R, torch.Tensor):
            R = torch.full_like(u_ref, float(R))
        R = R * ball_ratio
        c = p - n * R[:, None]
        theta = self.sample_uniform(sampler).clamp(clamping, 2 * math.pi - clamping)
        grad = torch.zeros_like(u_ref)
        nn = 2 if antithetic else 1
        for i in range(nn):
            if i == 1:
                theta = -theta
            p_dir = torch.stack([-n[:, 1], n[:, 0]], dim=1)
            y = c + R[:, None] * (n * torch.cos(theta)[:, None] + p_dir * torch.sin(theta)[:, None])
            d = scene.sdf(y).abs()
            u = _sel(d < self.epsilon, scene.dirichlet(scene.closest_point(y)), self.single_walk(y, scene, sampler))
            P = 1.0 / (torch.cos(theta) - 1.0)
            grad = grad + (P * (u - u_ref) / R if control_variates else P * u / R)
        grad = grad / nn
        y = c + R[:, None] * self.rand_in_disk(sampler)
        f = scene.source_function(y)
        P = self.eval_P(c, y, p, R)
        if control_variates:
            f_s = scene.source_function(p)
            interior = (f - f_s) * P * math.pi * R * R + f_s * R / 2.0
        else:
            interior = f * P * math.pi * R * R
        rel = (y - p).norm(dim=1) / R
        delta = 1e-4
        df = (scene.source_function(p + delta * its.n.float()) - scene.source_function(p)) / delta
        interior = _sel(rel < source_epsilon, df / math.pi, interior)
        grad = grad - interior
        return grad * n[:, 0], grad * n[:, 1]


@dataclass
class WoSWithSourceCUDA(ValueSolverMixin, WoSWithSource):
    """wos/wos_with_source.py:235-265: the native walk with its source
    contribution."""
    prevent_fd_artifacts: bool = False

    def __post_init__(self):
        super().__post_init__()
        self.cwos = wos_ext.WoS(nwalks=self.nwalks, nsteps=self.nsteps, epsilon=self.epsilon,
                                double_sided=self.double_sided, prevent_fd_artifacts=self.prevent_fd_artifacts,
                                use_IS_for_greens=self.use_IS_for_greens)

    def _cwos(self):
        self.cwos.nwalks, self.cwos.nsteps, self.cwos.epsilon = self.nwalks, self.nsteps, self.epsilon
        self.cwos.double_sided = self.double_sided
        self.cwos.prevent_fd_artifacts = self.prevent_fd_artifacts
        return self.cwos

    def single_walk_preliminary(self, p, scene, sampler):
        return self._cwos().single_walk(p.detach(), scene, sampler)

    def single_walk(self, p, scene, sampler):
        """wos/wos_with_source.py:257-265."""
        d = scene.sdf(p.detach())
        active = torch.ones_like(d, dtype=torch.bool) if self.double_sided else d < 0
        its = self.single_walk_preliminary(p, scene, sampler)
        its.valid = its.valid.bool() & active
        return scene.dirichlet(its) + _sel(its.valid, its.contrib, torch.zeros_like(its.contrib))
