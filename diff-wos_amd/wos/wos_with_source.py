"""2D Poisson walk-on-spheres (wos/wos_with_source.py): -Lap u = f with the
interior term sampled uniformly in each ball (include/wos.cuh:115-139)."""
from __future__ import annotations

from dataclasses import dataclass

import torch

from wos.wos import WoS


@dataclass
class WoSWithSource(WoS):
    """wos/wos_with_source.py:8-232.  ``use_IS_for_greens`` is accepted; the
    CUDA walk samples the ball uniformly (include/wos.cuh:121-134)."""
    use_IS_for_greens: bool = False
    control_variates: bool = True
    source_epsilon: float = 0.0

    def single_walk(self, p, scene, sampler):
        """wos/wos_with_source.py:257-265: g(exit) + contrib (0 if invalid)."""
        its = self.single_walk_preliminary(p, scene, sampler)
        v = scene.dirichlet(its) + its.contrib
        return torch.where(its.valid, v, torch.zeros_like(v))


@dataclass
class WoSWithSourceCUDA(WoSWithSource):
    """wos/wos_with_source.py:235-265."""
