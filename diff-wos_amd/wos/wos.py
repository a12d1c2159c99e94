"""2D Laplace walk-on-spheres solvers (wos/wos.py).

``WoS`` and ``WoSCUDA`` both run the CUDA walk semantics of the reference
(include/wos.cuh:97-145): non-terminated walks are invalid and contribute 0.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

import wos_ext
from wos.solver import Solver, ValueSolverMixin


@dataclass
class WoS(ValueSolverMixin, Solver):
    """wos/wos.py:10-133."""
    epsilon: float = 1e-3
    nwalks: int = 10
    nsteps: int = 32
    double_sided: bool = False
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
        """wos/wos.py:152-155: one CUDA walk per row, sampler advanced."""
        return self._cwos().single_walk(p, scene, sampler)

    def single_walk(self, p, scene, sampler):
        """wos/wos.py:171-179: value of one walk per row (0 if invalid)."""
        its = self.single_walk_preliminary(p, scene, sampler)
        v = scene.dirichlet(its)
        return torch.where(its.valid, v, torch.zeros_like(v))

    def u(self, p, scene, sampler):
        return self.single_walk(p, scene, sampler)

    def tangent_derivative(self, its, scene, sampler=None):
        return scene.tangent_derivative(its)


@dataclass
class WoSCUDA(WoS):
    """wos/wos.py:136-155."""
