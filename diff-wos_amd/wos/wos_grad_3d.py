"""3D shape-gradient integrator (wos/wos_grad_3d.py)."""
from __future__ import annotations

from dataclasses import dataclass

from wos.solver import ShapeSolverMixin, Solver


@dataclass
class Ours3D(ShapeSolverMixin, Solver):
    """wos/wos_grad_3d.py:282-315."""
    epsilon: float = 1e-3
    nwalks: int = 10
    nsteps: int = 32
    double_sided: bool = False
    epsilon2: float = 1e-3
    clamping: float = 1e-1
