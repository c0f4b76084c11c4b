"""CPU oracle of the Kokkos-semantics sweep (NumPy, float64).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` alone, as the checker.

Restates ``lagrange_interp_5_uniform`` and ``advect_1d_periodic`` /
``advect_1d_dirichlet`` of the reference's Kokkos 2D2V solver
(``simulations/vlasov2d_2v/kokkos/src/vlasov_poisson.hpp`` = K2, lines
244-339), with the departure point ``grid[i] - vel * dt`` of
``advect_1d_core`` (K2:463-464, K2:592-593), operation for operation
(every product and sum rounded on its own, no fused multiply-add).  The
Kokkos library itself is not in this image, so the reference cannot be
built or run; how each function is pinned instead
(``tests/test_oracle_kokkos.py``):

* ``kokkos_weights``: equal to the Lagrange product basis at nodes -2..3
  (a restatement from the definition, not from K2) to 1e-15 on [0, 1);
  x = 0 gives exactly the centre tap; exact on polynomials of degree <= 5;
  partition of unity.
* ``advect_axis_kokkos``: whole-cell displacements are exact rolls
  (periodic, positive and negative, also beyond one box length) and exact
  zero-filled shifts (Dirichlet); departure points outside the velocity box
  give 0; on smooth periodic data it agrees with the PyTorch-form oracle
  sweep (``vlasov_oracle.advect_axis``) to 1e-13.
"""
from __future__ import annotations

import numpy as np


def kokkos_weights(x):
    """The six Horner-form basis values of lagrange_interp_5_uniform
    (K2:244-291) at x (array), in tap order -2..3."""
    x = np.asarray(x, dtype=np.float64)
    b0 = x * (1.0 / 20.0 + x * (-1.0 / 24.0 + x * (-1.0 / 24.0 + x * (1.0 / 24.0 + x * (-1.0 / 120.0)))))
    b1 = x * (-1.0 / 2.0 + x * (2.0 / 3.0 + x * (-1.0 / 24.0 + x * (-1.0 / 6.0 + x * (1.0 / 24.0)))))
    b2 = 1.0 + x * (-1.0 / 3.0 + x * (-5.0 / 4.0 + x * (5.0 / 12.0 + x * (1.0 / 4.0 + x * (-1.0 / 12.0)))))
    b3 = x * (1.0 + x * (2.0 / 3.0 + x * (-7.0 / 12.0 + x * (-1.0 / 6.0 + x * (1.0 / 12.0)))))
    b4 = x * (-1.0 / 4.0 + x * (-1.0 / 24.0 + x * (7.0 / 24.0 + x * (1.0 / 24.0 + x * (-1.0 / 24.0)))))
    x2 = x * x
    b5 = x * (1.0 / 30.0 + x2 * (-1.0 / 24.0 + x2 * (1.0 / 120.0)))
    return (b0, b1, b2, b3, b4, b5)


def _interp(y, x):
    b = kokkos_weights(x)
    return b[0] * y[0] + b[1] * y[1] + b[2] * y[2] + b[3] * y[3] + b[4] * y[4] + b[5] * y[5]


def advect_rows_kokkos(rows, vel, dt, grid, L, xmin, inv_dx, periodic):
    """Sweep of rows [B, n] with per-row velocity vel [B] (K2:293-339)."""
    rows = np.asarray(rows, dtype=np.float64)
    B, n = rows.shape
    grid = np.asarray(grid, dtype=np.float64)
    vel = np.asarray(vel, dtype=np.float64).reshape(B, 1)
    dp = grid[None, :] - vel * dt
    bidx = np.arange(B)[:, None]
    if periodic:
        m = np.fmod(dp - xmin, L)
        il = np.floor(m * inv_dx).astype(np.int64)
        x = m * inv_dx - il.astype(np.float64)
        first = il - 2
        y = []
        for k in range(6):
            first = np.where(first < 0, first + n, first)  # K2:305-310, re-tested for every tap
            idx = first + k
            idx = np.where(idx >= n, idx - n, idx)
            idx = idx % n  # where K2 would read outside the line, wrap
            y.append(rows[bidx, idx])
        return _interp(y, x)
    s = dp - xmin
    inside = (s >= 0.0) & (s <= L)
    ss = np.where(inside, s, 0.0)
    il = np.floor(ss * inv_dx).astype(np.int64)
    x = ss * inv_dx - il.astype(np.float64)
    y = []
    for k in range(6):
        idx = il + k - 2
        ok = (idx >= 0) & (idx < n)
        y.append(np.where(ok, rows[bidx, np.clip(idx, 0, n - 1)], 0.0))
    return np.where(inside, _interp(y, x), 0.0)


def advect_axis_kokkos(f, velocities, axis, dt, grid, L, xmin, inv_dx, periodic):
    """advect_1d_core<axis> (K2:417-597) on a 4D array; ``velocities``
    broadcasts to the batch shape (all dims but ``axis``, in order)."""
    f = np.asarray(f, dtype=np.float64)
    perm = list(range(f.ndim))
    perm.append(perm.pop(axis))
    fp = np.transpose(f, perm)
    batch_shape = fp.shape[:-1]
    n = fp.shape[-1]
    flat = np.ascontiguousarray(fp).reshape(-1, n)
    vel = np.broadcast_to(np.asarray(velocities, dtype=np.float64), batch_shape).reshape(-1)
    out = advect_rows_kokkos(flat, vel, dt, grid, L, xmin, inv_dx, periodic).reshape(fp.shape)
    inv = [0] * f.ndim
    for i, p in enumerate(perm):
        inv[p] = i
    return np.ascontiguousarray(np.transpose(out, inv))
