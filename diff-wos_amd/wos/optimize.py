"""Inverse-solve loop (inv_solve/optimize.py:16-121) on torch autograd.

``TrainRunner`` renders a target once with ``model_target``, then for
``niters`` iterations renders ``model``, takes an L1 or L2 image loss
(optionally masked to pixels where both images are non-negligible),
back-propagates through the solvers' autograd functions and lets the model
step its optimiser.  ``Writer`` keeps each logged scalar's history in a
text file and writes images as OpenEXR (wos.io); the reference's
TensorBoard writer is replaced by these files.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

from wos.io import to_numpy, write_exr


@dataclass
class Writer:
    out_dir: str = "./"
    history: dict = field(default_factory=dict)

    def _dump(self, name):
        np.savetxt(Path(self.out_dir) / f"{name}.txt", np.array(self.history[name]))

    def add_scalar(self, name, value, step):
        self.history.setdefault(name, []).append(float(to_numpy(value)))
        self._dump(name)

    def add_tensor(self, name, value, step):
        self.history.setdefault(name, []).append(np.squeeze(to_numpy(value)))
        self._dump(name)

    def add_image(self, name, value, step, vmin=-2.0, vmax=2.0, cmap=None):
        path = Path(self.out_dir) / name
        path.mkdir(parents=True, exist_ok=True)
        write_exr(str(path / f"{step:04d}.exr"), to_numpy(value))


@dataclass
class Model:
    """A differentiable renderer with parameters and an optimiser."""

    def forward(self, i=0):
        raise NotImplementedError

    def render(self):
        return self.forward()

    def parameters(self):
        raise NotImplementedError

    def step(self):
        raise NotImplementedError

    def report(self, i, writer=None):
        pass


@dataclass
class TrainRunner:
    name: str = None
    out_dir: str = "./"
    model_target: Model = None
    model: Model = None
    pts: torch.Tensor = None
    niters: int = 1000
    is_l1: bool = True
    is_mask: bool = True
    verbose: bool = False

    def __post_init__(self):
        with torch.no_grad():
            self.target_image = self.model_target.forward().detach()
        self.out_dir = Path(self.out_dir)
        self.out_dir.mkdir(parents=True, exist_ok=True)
        self.writer = Writer(str(self.out_dir))
        with torch.no_grad():
            image = self.model_target.render()
        write_exr(str(self.out_dir / "target.exr"), to_numpy(image).reshape(-1, 1))

    def loss(self, image, target_image):
        image = image.reshape(-1)
        target_image = target_image.reshape(-1)
        if self.is_mask:
            mask = (image.detach().abs() > 0.001) & (target_image.abs() > 0.001)
            image = torch.where(mask, image, torch.zeros_like(image))
            target_image = torch.where(mask, target_image, torch.zeros_like(target_image))
        if self.is_l1:
            return (image - target_image).abs().sum()
        return ((image - target_image) ** 2).sum()

    def run(self):
        for i in range(self.niters):
            image = self.model.forward(i)
            loss = self.loss(image, self.target_image)
            loss.backward()
            with torch.no_grad():
                param_error = (self.model.parameters() - self.model_target.parameters()).abs().mean()
            if self.verbose:
                print(f"iteration {i}: loss {loss.item():.6g}, param error {param_error.item():.6g}")
            self.writer.add_scalar("loss", loss.detach(), i)
            self.writer.add_scalar("param_error", param_error, i)
            self.model.report(i, self.writer)
            self.model.step()
        return self.writer.history
