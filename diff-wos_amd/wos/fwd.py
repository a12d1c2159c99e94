"""Type aliases of wos/fwd.py:1-21 as torch constructors.

The reference aliases Dr.Jit CUDA types (Float, Int, Array2, ...); here each
alias builds a torch tensor on the current HIP device: scalars arrays are 1-D,
``Array2``/``Array3`` are ``[n, 2]``/``[n, 3]`` float32, ``Array2i``/
``Array3i`` int32.  ``PCG32`` keeps Dr.Jit's per-element (state, inc)
layout in int64 tensors holding the uint64 bits.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

basedir = Path(__file__).parent.parent

PCG32_DEFAULT_STATE = 0x853c49e6748fea9b
PCG32_DEFAULT_STREAM = 0xda3e39cb94b95bdb
PCG32_MULT = 0x5851f42d4c957f2d


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("the wos package needs a HIP device")
    return torch.device("cuda", torch.cuda.current_device())


def _as(x, dtype, cols=None):
    t = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x))
    if t.device.type != "cuda":
        t = t.to(device())
    if t.dtype != dtype:
        t = t.to(dtype)
    if cols is not None:
        t = t.reshape(-1, cols)
    return t


def Float(x):
    t = _as(x, torch.float32)
    return t.reshape(-1) if t.dim() == 0 else t


def Float64(x):
    return _as(x, torch.float64)


def Int(x):
    t = _as(x, torch.int32)
    return t.reshape(-1) if t.dim() == 0 else t


def Bool(x):
    return _as(x, torch.bool)


def Array2(*xs):
    if len(xs) == 2:
        return torch.stack([Float(xs[0]).reshape(-1), Float(xs[1]).reshape(-1)], dim=1)
    return _as(xs[0], torch.float32, 2)


def Array3(*xs):
    if len(xs) == 3:
        return torch.stack([Float(x).reshape(-1) for x in xs], dim=1)
    return _as(xs[0], torch.float32, 3)


def Array2i(x):
    return _as(x, torch.int32, 2)


def Array3i(x):
    return _as(x, torch.int32, 3)


def _u64(x):
    """uint64 bit pattern as int64 (two's complement)."""
    x = int(x) & (2 ** 64 - 1)
    return x - 2 ** 64 if x >= 2 ** 63 else x


def _lsr(x, k):
    """Logical right shift of int64 tensors holding uint64 bits."""
    return (x >> k) & ((1 << (64 - k)) - 1)


class PCG32:
    """Dr.Jit PCG32 (the struct include/sampler2.cuh:11-56 restates):
    ``PCG32(size, initstate, initseq)`` seeds element k with sequence index k."""

    def __init__(self, size=1, initstate=PCG32_DEFAULT_STATE, initseq=PCG32_DEFAULT_STREAM, dev=None):
        dv = device() if dev is None else torch.device(dev)
        idx = torch.arange(int(size), dtype=torch.int64, device=dv)
        st = initstate if isinstance(initstate, torch.Tensor) else torch.full((int(size),), _u64(initstate),
                                                                               dtype=torch.int64, device=dv)
        sq = initseq if isinstance(initseq, torch.Tensor) else torch.full((int(size),), _u64(initseq),
                                                                           dtype=torch.int64, device=dv)
        # Int32 -> UInt64 conversion sign-extends (int64 cast of int32)
        st = st.to(torch.int64)
        sq = sq.to(torch.int64)
        self.state = torch.zeros(int(size), dtype=torch.int64, device=dv)
        self.inc = ((sq + idx) << 1) | 1
        self.next_uint32()
        self.state = self.state + st
        self.next_uint32()

    def next_uint32(self):
        old = self.state
        self.state = old * _u64(PCG32_MULT) + self.inc
        xs = _lsr(_lsr(old, 18) ^ old, 27) & 0xffffffff
        rot = _lsr(old, 59)
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xffffffff

    def next_float32(self):
        u = (self.next_uint32() >> 9) | 0x3f800000
        return u.to(torch.int32).view(torch.float32) - 1.0

    def next_float64(self):
        lo = self.next_uint32()
        hi = self.next_uint32()
        v = _lsr(lo | (hi << 32), 12) | 0x3ff0000000000000
        return v.view(torch.float64) - 1.0
