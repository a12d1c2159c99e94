"""2D shape-gradient integrators (wos/wos_grad.py).  ``OursCUDA`` (and
``Ours``, the same estimator) return zero with d/dvertices attached: primary
walk, largest inscribed ball, two antithetic secondary walks for the normal
derivative, tangential derivative, -grad u . V (wos/wos_grad.py:113-151)."""
from __future__ import annotations

from dataclasses import dataclass

from wos.solver import ShapeSolverMixin, Solver


@dataclass
class OursCUDA(ShapeSolverMixin, Solver):
    """wos/wos_grad.py:425-463."""
    epsilon2: float = 1e-3
    clamping: float = 1e-1
    control_variates: bool = True


@dataclass
class Ours(OursCUDA):
    """wos/wos_grad.py:57-86 (same estimator, run by the fused kernel)."""
