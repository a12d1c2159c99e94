"""Solver base and CustomSolver (wos/solver.py:1-125) on torch autograd.

``Solver.walk(p, scene, seed)`` returns the per-point mean of ``nwalks``
walks (wos/solver.py:49-65) as a float32 tensor.  Walk k = i*W + w of point
i draws from the PCG32 stream keyed by TEA(seed, i) and k, so a second call
with the same seed retraces the same paths -- the property the reference's
CustomSolver backward relies on (wos/solver.py:109-114).

Gradients: a solver whose value depends on the scene parameters (the walk
value solvers) returns a tensor attached to ``scene.values`` and, for a
Fourier source, ``scene.source_params``; the adjoint reuses the exit records
stored by the forward launch.  Shape-gradient solvers (OursCUDA, Ours3D)
return zeros attached to ``scene.vertices``, as the reference's integrators
return zero-valued, gradient-carrying results.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from enum import Enum

import torch

import wos_ext
from wos.fwd import PCG32
from wos.utils import sample_tea_32


class ControlVarianceType(Enum):
    NoControlVariate = 0
    RunningControlVariate = 1
    BoundaryControlVariate = 2
    BoundaryAndRunningControlVariate = 3


def _requires(t):
    return isinstance(t, torch.Tensor) and t.requires_grad


def _has_tangent(t):
    """True for a forward-mode dual tensor (torch.autograd.forward_ad)."""
    if not isinstance(t, torch.Tensor):
        return False
    import torch.autograd.forward_ad as fwAD
    return fwAD.unpack_dual(t).tangent is not None


def _source_tensor(scene):
    sp = getattr(scene, "source_params", None)
    return sp if isinstance(sp, torch.Tensor) else None


class _ValueWalk(torch.autograd.Function):
    """mean_i = 1/W sum_w value(walk i,w); reverse mode through the exit
    records stored by the forward launch (vertex values) and, for the source
    parameters, by re-running the walks from their keys (every parameter of
    a Gaussian, sinusoid or Fourier source; wos_ext.adjoint_source)."""

    @staticmethod
    def forward(ctx, cwos, nwalks, points, scene, seed, values, source_params):
        need_v = ctx.needs_input_grad[5]
        need_s = ctx.needs_input_grad[6] and bool(scene.source_type)
        out = cwos.walk_reduce(points, scene, seed, nwalks=nwalks, records=need_v)
        ctx.scene, ctx.cwos, ctx.nwalks, ctx.seed, ctx.P = scene, cwos, nwalks, seed, out["mean"].shape[0]
        ctx.points = points.detach() if need_s else None
        ctx.need_v, ctx.need_s = need_v, need_s
        ctx.sp_shape = None if source_params is None else source_params.shape
        ctx.out = {k: out[k] for k in ("rec_prim", "rec_t") if k in out}
        return out["mean"]

    @staticmethod
    def backward(ctx, g):
        o = ctx.out
        g = g.contiguous()
        gv = gsp = None
        if ctx.need_v:
            gv, _ = wos_ext.adjoint_values(ctx.scene, ctx.P, ctx.nwalks, o["rec_prim"], o["rec_t"], g,
                                           want_values=True, want_coeffs=False)
        if ctx.need_s:
            gsp = wos_ext.adjoint_source(ctx.scene, ctx.cwos, ctx.points, ctx.seed, g,
                                         nwalks=ctx.nwalks).reshape(ctx.sp_shape)
        return None, None, None, None, None, gv, gsp


class _ShapeWalk(torch.autograd.Function):
    """Zero-valued output whose gradient w.r.t. the vertices is the
    OursCUDA / Ours3D shape-gradient estimate, walks recomputed from the key.
    Forward mode (jvp) gives the derivative along a vertex velocity field
    from the same samples -- what the reference's renderD takes with
    dr.forward_to (diff_solve/runner.py:75-109)."""

    @staticmethod
    def forward(ctx, solver, points, scene, seed, vertices, nwalks):
        ctx.solver, ctx.points, ctx.scene, ctx.seed, ctx.nwalks = solver, points.detach(), scene, seed, nwalks
        return torch.zeros(points.shape[0], dtype=torch.float32, device=points.device)

    @staticmethod
    def backward(ctx, g):
        s = ctx.solver
        gv, _ = wos_ext.shape_gradient(ctx.scene, ctx.points, ctx.seed, g.contiguous(), nwalks=ctx.nwalks,
                                       nsteps=s.nsteps, epsilon=s.epsilon, epsilon2=s.epsilon2,
                                       clamping=s.clamping, double_sided=s.double_sided,
                                       control_variates=getattr(s, "control_variates", True))
        return None, None, None, None, gv, None

    @staticmethod
    def jvp(ctx, d_solver, d_points, d_scene, d_seed, d_vertices, d_nwalks):
        if d_vertices is None:
            return torch.zeros(ctx.points.shape[0], dtype=torch.float32, device=ctx.points.device)
        return ctx.solver.walk_derivative(ctx.points, ctx.scene, d_vertices, seed=ctx.seed, nwalks=ctx.nwalks)


@dataclass
class Solver:
    epsilon: float = 0.005   # controls condition to stop recursion
    nwalks: int = 20         # number of samples per point queried
    nsteps: int = 4          # maximum depth of a sampled path
    double_sided: bool = False
    control_variance: ControlVarianceType = ControlVarianceType.NoControlVariate
    antithetic: bool = False

    # ---- wos/solver.py:41-65 ------------------------------------------------
    def solve(self, p, scene, seed=0):
        return self.walk_reduce(p, scene, seed)

    def walk(self, p, scene, seed=0):
        return self.walk_reduce(p, scene, seed)

    def walk_reduce(self, p, scene, seed=0):
        return self._walk(p, scene, seed, self.nwalks)

    def walk_detector(self, scene, detector, seed=0):
        """wos/solver.py:67-84 (pixel jitter drawn per walk before walking).
        The jitter streams are the reference's TEA(seed, k) keying; walks run
        on the jittered points with the per-point mean over nwalks."""
        p = detector.make_points()
        n = p.shape[0] * self.nwalks
        idx = torch.arange(n, dtype=torch.int64, device=p.device)
        v0, v1 = sample_tea_32(seed, idx)
        sampler = PCG32(n, v0.to(torch.int64), v1.to(torch.int64))
        jp = detector.make_jittered_points(sampler, self.nwalks)
        vals = self._walk(jp, scene, seed, 1)
        return vals.reshape(-1, self.nwalks).mean(dim=1)

    def single_walk(self, p, scene, sampler):
        raise NotImplementedError

    # ---- sampling helpers (wos/solver.py:25-42); draws advance every lane ----
    def rand_on_circle(self, sampler):
        angle = sampler.next_float64() * (2.0 * math.pi)
        return torch.stack([torch.cos(angle), torch.sin(angle)], dim=1).float()

    def rand_in_half_circle(self, n, sampler):
        d = self.rand_on_circle(sampler)
        return torch.where(((n * d).sum(1) > 0)[:, None], d, -d)

    def sample_uniform(self, sampler):
        return (sampler.next_float64() * (2.0 * math.pi)).float()

    def pdf_uniform(self, theta):
        return 1.0 / (2.0 * math.pi)

    # ---- the generic walk_reduce (wos/solver.py:49-65) ------------------------
    def _walk(self, p, scene, seed, nwalks):
        """Walk k = i*W + w of point i on the PCG32 stream keyed by TEA(seed,
        i) and k, with this solver's single_walk over all P*W walks at once
        (torch, one pass per step); the per-point mean stays attached to
        whatever the walk values are attached to."""
        P = p.shape[0]
        n = P * nwalks
        idx = torch.arange(n, dtype=torch.int64, device=p.device)
        if nwalks > 1:
            idx = idx // nwalks
        v0, v1 = sample_tea_32(seed, idx)
        sampler = PCG32(n, v0.to(torch.int64), v1.to(torch.int64))
        value = self.single_walk(p.repeat_interleave(nwalks, dim=0), scene, sampler)
        out = torch.zeros(P, dtype=value.dtype, device=p.device).index_add(0, idx, value)
        return out / nwalks if nwalks > 1 else out

    def _walk_with(self, p, scene, seed, nwalks, vertices, values, source_params):
        """_walk with the scene's vertex / value / source tensors replaced by
        the given ones for the duration of the call (CustomSolver backward)."""
        saved = scene.swap_fields(vertices, values, source_params)
        try:
            return self._walk(p, scene, seed, nwalks)
        finally:
            scene.restore_fields(saved)


class _CustomWalk(torch.autograd.Function):
    """dr.custom(_CustomSolver) (wos/solver.py:90-114): detached forward
    value of fwd_solver, gradient of bwd_solver under the same seed."""

    @staticmethod
    def forward(ctx, fwd, bwd, points, scene, seed, vertices, values, source_params):
        with torch.no_grad():
            val = fwd._walk(points, scene, seed, fwd.nwalks)
        ctx.bwd, ctx.points, ctx.scene, ctx.seed = bwd, points.detach(), scene, seed
        ctx.sp = source_params
        return val.detach()

    @staticmethod
    def backward(ctx, g):
        bwd, scene = ctx.bwd, ctx.scene
        with torch.enable_grad():
            vt = scene.vertices.detach().requires_grad_(ctx.needs_input_grad[5])
            vv = scene.values.detach().requires_grad_(ctx.needs_input_grad[6])
            sp = ctx.sp.detach().requires_grad_(ctx.needs_input_grad[7]) if ctx.sp is not None else None
            out = bwd._walk_with(ctx.points, scene, ctx.seed, bwd.nwalks, vt, vv, sp)
            ins = [x for x in (vt, vv, sp) if x is not None and x.requires_grad]
            if not ins or not out.requires_grad:
                return None, None, None, None, None, None, None, None
            grads = torch.autograd.grad(out, ins, g, allow_unused=True)
        it = iter(grads)
        gvt = next(it) if vt.requires_grad else None
        gvv = next(it) if vv.requires_grad else None
        gsp = next(it) if (sp is not None and sp.requires_grad) else None
        return None, None, None, None, None, gvt, gvv, gsp


@dataclass
class CustomSolver:
    """wos/solver.py:117-125."""
    fwd_solver: Solver = None
    bwd_solver: Solver = None

    def walk(self, p, scene, seed=0):
        return _CustomWalk.apply(self.fwd_solver, self.bwd_solver, p, scene, seed, scene.vertices, scene.values,
                                 _source_tensor(scene))


class ValueSolverMixin:
    """Solvers whose walk value is g(exit) (+ source contribution)."""

    def _cwos(self):
        raise NotImplementedError

    def _walk(self, p, scene, seed, nwalks):
        return self._walk_with(p, scene, seed, nwalks, scene.vertices, scene.values, _source_tensor(scene))

    def _walk_with(self, p, scene, seed, nwalks, vertices, values, source_params):
        need = _requires(values) or _requires(source_params)
        if not (torch.is_grad_enabled() and need):
            return self._cwos().walk_reduce(p, scene, seed, nwalks=nwalks)["mean"]
        return _ValueWalk.apply(self._cwos(), nwalks, p, scene, seed, values, source_params)

    def walk_stats(self, p, scene, seed=0):
        """Per-point FP64 mean and variance of the walk values."""
        o = self._cwos().walk_reduce(p, scene, seed, nwalks=self.nwalks, stats=True)
        return o["mean64"], o["var64"]


class ShapeSolverMixin:
    """Solvers that return zero with the shape derivative as gradient."""

    def _walk(self, p, scene, seed, nwalks):
        return self._walk_with(p, scene, seed, nwalks, scene.vertices, scene.values, _source_tensor(scene))

    def _walk_with(self, p, scene, seed, nwalks, vertices, values, source_params):
        if not ((torch.is_grad_enabled() and _requires(vertices)) or _has_tangent(vertices)):
            return torch.zeros(p.shape[0], dtype=torch.float32, device=p.device)
        return _ShapeWalk.apply(self, p, scene, seed, vertices, int(nwalks))

    def walk_derivative(self, p, scene, velocity, seed=0, stats=False, nwalks=None):
        """Per point, the derivative of this estimator when vertex v moves
        with velocity[v] (forward mode, wos_ext.shape_derivative); the
        samples are those of the backward pass with the same seed."""
        W = self.nwalks if nwalks is None else int(nwalks)
        return wos_ext.shape_derivative(scene, p, seed, velocity, nwalks=W, nsteps=self.nsteps,
                                        epsilon=self.epsilon, epsilon2=self.epsilon2, clamping=self.clamping,
                                        double_sided=self.double_sided,
                                        control_variates=getattr(self, "control_variates", True), stats=stats)

    def sample_grad_u(self, p, scene, seed=0):
        """Per-sample spatial gradient estimates (n_points*nwalks, dim)."""
        _, sg = wos_ext.shape_gradient(scene, p, seed, None, nwalks=self.nwalks, nsteps=self.nsteps,
                                       epsilon=self.epsilon, epsilon2=self.epsilon2, clamping=self.clamping,
                                       double_sided=self.double_sided,
                                       control_variates=getattr(self, "control_variates", True),
                                       sample_grad_u=True)
        return sg
