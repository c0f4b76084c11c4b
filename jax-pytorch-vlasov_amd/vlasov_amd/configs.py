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


def ensemble_specs(n: int = 512, seed: int = 0) -> list:
    """Parameters of the batched 1D1V ensemble (numpy PCG64, seeded):
    nx ~ U{32..256}, nv ~ U{64..512}, k ~ U[0.2, 0.6] (L = 2 pi / k),
    amplitude ~ log-U[1e-3, 0.5], beam drift u ~ U[0, 3] (u = 0 is Landau),
    V_max = max(6, u + 6).  Initial condition (see ``ensemble_f0``):
    (1 + A cos(k x)) 1/2 [M(v - u) + M(v + u)]."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(rng.integers(32, 257))
        nv = int(rng.integers(64, 513))
        k = float(rng.uniform(0.2, 0.6))
        amp = float(np.exp(rng.uniform(np.log(1e-3), np.log(0.5))))
        u = float(rng.uniform(0.0, 3.0))
        out.append({"nx": nx, "nv": nv, "k": k, "L": 2.0 * math.pi / k, "amplitude": amp, "drift": u,
                    "V_max": max(6.0, u + 6.0)})
    return out


def ensemble_f0(grid, k: float, amplitude: float, drift: float) -> torch.Tensor:
    """(1 + A cos(k x)) 1/2 [M(v - u) + M(v + u)] on a 1D1V grid."""
    fv = 0.5 * (maxwellian(grid.v, drift) + maxwellian(grid.v, -drift))
    return ((1.0 + amplitude * torch.cos(k * grid.x))[:, None] * fv[None, :]).contiguous()
