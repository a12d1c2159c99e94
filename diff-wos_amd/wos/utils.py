"""wos/utils.py helpers on torch: TEA seeding, interpolation, triangle
closest point, rotations and concatenation."""
from __future__ import annotations

import math

import torch

from wos.fwd import Array2, Float


def lerp(a, b, t):
    return a * (1 - t) + b * t


def interpolate(a, b, c, uv):
    """wos/utils.py:13-14: b*u + c*v + a*(1-u-v); uv is [n, 2]."""
    u = uv[..., 0:1] if a.dim() > 1 else uv[..., 0]
    v = uv[..., 1:2] if a.dim() > 1 else uv[..., 1]
    return b * u + c * v + a * (1 - u - v)


def _s32(x):
    """int64 tensor of 32-bit patterns -> signed value."""
    x = x & 0xffffffff
    return x - ((x >> 31) << 32)


def sample_tea_32(v0, v1, round=4):
    """wos/utils.py:250-256 with Int32 semantics (wrapping add/shl, arithmetic
    shr).  v0/v1: python ints or int tensors; returns int32 tensors."""
    dev = None
    for x in (v0, v1):
        if isinstance(x, torch.Tensor):
            dev = x.device
    a = torch.as_tensor(v0, dtype=torch.int64, device=dev) & 0xffffffff
    b = torch.as_tensor(v1, dtype=torch.int64, device=dev) & 0xffffffff
    a, b = torch.broadcast_tensors(a, b)
    a, b = a.clone(), b.clone()
    s = 0
    M = 0xffffffff
    for _ in range(round):
        s = (s + 0x9e3779b9) & M
        a = (a + ((((b << 4) + 0xa341316c) & M) ^ ((b + s) & M) ^ (((_s32(b) >> 5) + 0xc8013ea4) & M))) & M
        b = (b + ((((a << 4) + 0xad90777d) & M) ^ ((a + s) & M) ^ (((_s32(a) >> 5) + 0x7e95761e) & M))) & M
    return _s32(a).to(torch.int32), _s32(b).to(torch.int32)


def closest_point_triangle(p, a, b, c):
    """wos/utils.py:108-173 (Ericson): returns (pt, uv, d) with uv the
    barycentrics of (b, c)."""
    ab, ac = b - a, c - a
    ax = p - a
    d1 = (ab * ax).sum(-1)
    d2 = (ac * ax).sum(-1)
    bx = p - b
    d3 = (ab * bx).sum(-1)
    d4 = (ac * bx).sum(-1)
    cx = p - c
    d5 = (ab * cx).sum(-1)
    d6 = (ac * cx).sum(-1)
    vc = d1 * d4 - d3 * d2
    vb = d5 * d2 - d1 * d6
    va = d3 * d6 - d5 * d4
    zero = torch.zeros_like(d1)
    one = torch.ones_like(d1)
    # face region
    denom = 1.0 / (va + vb + vc)
    v = vb * denom
    w = vc * denom
    wa, wb = 1 - v - w, v
    pt = a + ab * v[..., None] + ac * w[..., None]
    conds = [
        ((va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0), lambda: (zero, 1 - (d4 - d3) / ((d4 - d3) + (d5 - d6)))),
        ((vb <= 0) & (d2 >= 0) & (d6 <= 0), lambda: (1 - d2 / (d2 - d6), zero)),
        ((vc <= 0) & (d1 >= 0) & (d3 <= 0), lambda: (1 - d1 / (d1 - d3), d1 / (d1 - d3))),
        ((d6 >= 0) & (d5 <= d6), lambda: (zero, zero)),
        ((d3 >= 0) & (d4 <= d3), lambda: (zero, one)),
        ((d1 <= 0) & (d2 <= 0), lambda: (one, zero)),
    ]
    for cond, f in conds:   # later entries take precedence (Ericson's order)
        xa, xb = f()
        wa = torch.where(cond, xa, wa)
        wb = torch.where(cond, xb, wb)
    wc = 1 - wa - wb
    pt = a * wa[..., None] + b * wb[..., None] + c * wc[..., None]
    d = (p - pt).norm(dim=-1)
    return pt, torch.stack([wb, wc], dim=-1), d


def rotate(v, angle):
    """wos/utils.py:192-195: rotate [n, 2] points by angle (radians)."""
    ang = torch.as_tensor(angle, dtype=v.dtype, device=v.device).reshape(())
    c, s = torch.cos(ang), torch.sin(ang)
    return torch.stack([c * v[:, 0] - s * v[:, 1], s * v[:, 0] + c * v[:, 1]], dim=1)


def rotate_euler(v, angles):
    ax, ay, az = [torch.as_tensor(x, dtype=v.dtype, device=v.device) for x in angles]
    cx, sx, cy, sy, cz, sz = torch.cos(ax), torch.sin(ax), torch.cos(ay), torch.sin(ay), torch.cos(az), torch.sin(az)
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    y, z = cx * y - sx * z, sx * y + cx * z
    x, z = cy * x + sy * z, -sy * x + cy * z
    x, y = cz * x - sz * y, sz * x + cz * y
    return torch.stack([x, y, z], dim=1)


def concat(a, b):
    """wos/utils.py:217-224: concatenate along the element axis."""
    return torch.cat([a, b], dim=0)


class DiscreteDistribution:
    """The sampling of mi.DiscreteDistribution used here: CDF over an
    unnormalised pmf; sample_reuse_pmf(u) returns the index whose CDF
    interval holds u * sum, the position of u inside that interval rescaled
    to [0, 1), and the normalised pmf of the index.  (Mitsuba's own tie and
    edge handling is not pinned here.)"""

    def __init__(self, pmf):
        pmf = pmf.double()
        self.pmf = pmf
        self.cdf = torch.cumsum(pmf, 0)
        self.sum = float(self.cdf[-1])

    def sample_reuse_pmf(self, u):
        v = u.double() * self.sum
        idx = torch.searchsorted(self.cdf, v, right=True).clamp(max=self.pmf.numel() - 1)
        lo = torch.where(idx > 0, self.cdf[(idx - 1).clamp(min=0)], torch.zeros_like(v))
        t = ((v - lo) / self.pmf[idx]).clamp(0.0, 1.0)
        return idx, t.float(), (self.pmf[idx] / self.sum).float()
