"""Image tasks over a parametrised scene (diff_solve/runner.py:10-154).

``TestRunner.make_scene(delta)`` returns the scene deformed by the scalar
``delta``; ``renderC`` averages primal images over passes, ``renderD`` the
derivative images d image / d delta, and ``renderFD`` the one-sided finite
difference (image(delta) - image(0)) / delta with common seeds, as the
reference computes it.  The reference's renderD seeds delta's tangent and reads Dr.Jit's
forward-mode gradient of the image; here delta is a torch forward-AD dual
number: the shape solvers (OursCUDA, Ours3D) answer its tangent with the
native forward-mode shape derivative (wos_shape_derivative), the boundary
term (WoSBoundary) through torch's own forward AD.  Images are written as
OpenEXR by wos.io.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import List

import numpy as np
import torch
import torch.autograd.forward_ad as fwAD

from wos.io import to_numpy, write_exr
from wos.scene import Detector
from wos.solver import Solver


@dataclass
class Task:
    solver: Solver = None
    out_file: str = None
    npasses: int = None


@dataclass
class RenderC(Task):
    out_file: str = "renderC.exr"


@dataclass
class RenderD(Task):
    out_file: str = "renderD.exr"


@dataclass
class RenderFD(Task):
    out_file: str = "renderFD.exr"


@dataclass
class TestRunner:
    """diff_solve/runner.py:33-154."""
    detector: Detector = field(default_factory=Detector)
    tasks: List[Task] = field(default_factory=list)
    out_dir: str = "./out"
    delta: float = 1e-2
    npasses: int = 1
    exclude_boundary: bool = False
    wos_boundary: Solver = None
    pixel_sampling: bool = True
    verbose: bool = False

    __test__ = False   # not a pytest test class despite the reference's name

    def __post_init__(self):
        self.p = self.detector.make_points()
        Path(self.out_dir).mkdir(parents=True, exist_ok=True)

    def make_scene(self, delta=0.0):
        """The scene transformed by delta (a float or a dual tensor)."""
        raise NotImplementedError

    def _image(self, solver, scene, seed):
        if self.pixel_sampling:
            return solver.walk_detector(scene, self.detector, seed=seed)
        return solver.walk(self.p, scene, seed=seed)

    def _save(self, image, out_file):
        write_exr(str(Path(self.out_dir) / out_file), np.asarray(to_numpy(image), np.float32).reshape(self.detector.res))

    def _mask_boundary(self, image, scene):
        if not self.exclude_boundary:
            return image
        near = scene.sdf(self.p).abs() < self.delta + 1e-3
        return torch.where(near, torch.zeros_like(image), image)

    def renderC(self, solver, out_file="renderC.exr", npasses=None):
        """diff_solve/runner.py:54-73."""
        npasses = self.npasses if npasses is None else npasses
        scene = self.make_scene()
        image = None
        for i in range(npasses):
            if self.verbose:
                print("pass:", i)
            with torch.no_grad():
                img = self._image(solver, scene, i).detach()
            image = img if image is None else image + img
            self._save(image / (i + 1), out_file)
        image = self._mask_boundary(image / npasses, scene)
        self._save(image, out_file)
        return to_numpy(image).reshape(self.detector.res)

    def renderD(self, solver, out_file="renderD.exr", npasses=None):
        """diff_solve/runner.py:75-109: forward-mode derivative of the image
        with respect to delta at 0 (tangent 1)."""
        npasses = self.npasses if npasses is None else npasses
        d_image = None
        scene = None
        for i in range(npasses):
            if self.verbose:
                print("pass:", i)
            with fwAD.dual_level():
                delta = fwAD.make_dual(torch.zeros((), device=self.p.device), torch.ones((), device=self.p.device))
                scene = self.make_scene(delta)
                image = self._image(solver, scene, i)
                if self.wos_boundary is not None:
                    image = image + self.wos_boundary.walk_detector(scene, self.detector, seed=i)
                tangent = fwAD.unpack_dual(image).tangent
            if tangent is None:
                tangent = torch.zeros_like(fwAD.unpack_dual(image).primal)
            tangent = tangent.detach()
            d_image = tangent if d_image is None else d_image + tangent
            self._save(self._mask_boundary(d_image / (i + 1), self.make_scene()), out_file)
        d_image = self._mask_boundary(d_image / npasses, self.make_scene())
        self._save(d_image, out_file)
        return to_numpy(d_image).reshape(self.detector.res)

    def renderFD(self, solver, out_file="renderFD.exr", npasses=None):
        """diff_solve/runner.py:111-136: (image(delta) - image(0)) / delta,
        both scenes walked with the same seeds."""
        npasses = self.npasses if npasses is None else npasses
        scene1, scene2 = self.make_scene(), self.make_scene(self.delta)
        d_image = None
        for i in range(npasses):
            if self.verbose:
                print("pass:", i)
            with torch.no_grad():
                d = (self._image(solver, scene2, i) - self._image(solver, scene1, i)).detach() / self.delta
            d_image = d if d_image is None else d_image + d
            self._save(self._mask_boundary(d_image / (i + 1), scene1), out_file)
        d_image = self._mask_boundary(d_image / npasses, scene1)
        self._save(d_image, out_file)
        return to_numpy(d_image).reshape(self.detector.res)

    def run_task(self, task):
        if isinstance(task, RenderC):
            return self.renderC(task.solver, task.out_file, task.npasses)
        if isinstance(task, RenderD):
            return self.renderD(task.solver, task.out_file, task.npasses)
        if isinstance(task, RenderFD):
            return self.renderFD(task.solver, task.out_file, task.npasses)
        raise NotImplementedError(type(task).__name__)

    def run_task_idx(self, idx):
        return self.run_task(self.tasks[idx])

    def run(self):
        return [self.run_task(t) for t in self.tasks]
