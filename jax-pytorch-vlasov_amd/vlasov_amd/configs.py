"""Initial conditions of the benchmark workloads (closed forms, built on the device).

* ``landau_1d``: (1 + eps cos(k x)) exp(-v^2/2) / sqrt(2 pi)        (1D1V Landau damping)
* ``landau_2d``: (1 + eps (cos kx + cos ky)) exp(-(vx^2+vy^2)/2) / (2 pi)   (2D2V Landau damping)
* ``two_stream_2d``: (1 + eps (cos kx + cos ky)) 1/2 [M(vx-u) + M(vx+u)] M(vy),
  M(v) = exp(-v^2/2)/sqrt(2 pi)                                       (2D2V two-stream)
"""
from __future__ import annotations

import math

import torch


def maxwellian(v: torch.Tensor, u: float = 0.0) -> torch.Tensor:
    return torch.exp(-0.5 * (v - u) ** 2) / math.sqrt(2.0 * math.pi)


def landau_1d(grid, eps: float = 0.01, k: float = 0.5) -> torch.Tensor:
    return (1.0 + eps * torch.cos(k * grid.x))[:, None] * maxwellian(grid.v)[None, :]


def landau_2d(grid, eps: float = 0.01, k: float = 0.5) -> torch.Tensor:
    pert = 1.0 + eps * (torch.cos(k * grid.x)[:, None] + torch.cos(k * grid.y)[None, :])
    fv = maxwellian(grid.vx)[:, None] * maxwellian(grid.vy)[None, :]
    return (pert[:, :, None, None] * fv[None, None, :, :]).contiguous()


def two_stream_2d(grid, eps: float = 0.001, k: float = 0.2, u: float = 2.4) -> torch.Tensor:
    pert = 1.0 + eps * (torch.cos(k * grid.x)[:, None] + torch.cos(k * grid.y)[None, :])
    fvx = 0.5 * (maxwellian(grid.vx, u) + maxwellian(grid.vx, -u))
    fv = fvx[:, None] * maxwellian(grid.vy)[None, :]
    return (pert[:, :, None, None] * fv[None, None, :, :]).contiguous()
