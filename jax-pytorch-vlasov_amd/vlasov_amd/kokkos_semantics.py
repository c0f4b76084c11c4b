"""Kokkos-semantics sweeps and step, for cross-checking against Kokkos runs.

The rest of the package follows the reference's PyTorch solvers.  The
reference's Kokkos implementation (simulations/vlasov2d_2v/kokkos/src/
vlasov_poisson.hpp, K2 below) computes the same scheme with different
boundary and rounding semantics: departure points in physical coordinates,
fmod wrapping by the box length, out-of-domain zeroing for the velocity
axes, Horner-form weights (K2:244-339).  ``advect_1d_kokkos`` runs that
sweep through the library's ``vp_advect_kokkos_f64``; ``vlasov_step_kokkos``
composes the Kokkos solver-0 step (K2:724-767) from it and the library's
field solve (the same delta_rho = rho - 1 spectral solve as K2:670-718).
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import check, lib, require_gpu, stream_handle
from . import vlasov2D2V as _v2


def advect_1d_kokkos(f, velocities, axis, dt, grid_points, L, xmin, dx, periodic):
    """advect_1d_core<axis> (K2:417-597) on a 4D float64 f(x, y, vx, vy).

    ``velocities`` broadcasts to the batch shape (all dims except ``axis``),
    like ``vlasov2D2V.advect_1d``; ``grid_points`` are the coordinates of the
    swept axis; ``L``, ``xmin``, ``dx`` as K2 passes them (box length, 0 and
    dx for x and y; 2 V_max, -V_max and dv for the velocity axes).
    """
    if f.ndim != 4:
        raise ValueError("advect_1d_kokkos expects a 4D f(x, y, vx, vy)")
    dev = require_gpu(f, velocities, grid_points)
    if f.dtype != torch.float64:
        raise TypeError("advect_1d_kokkos computes in float64")
    axis = int(axis) % 4
    f_c = f.contiguous()
    batch_shape = tuple(s for i, s in enumerate(f_c.shape) if i != axis)
    vel = velocities.to(device=dev, dtype=torch.float64).expand(batch_shape)
    strides = list(vel.stride())
    strides.insert(axis, 0)
    grid = grid_points.to(device=dev, dtype=torch.float64).contiguous()
    if grid.numel() != f_c.shape[axis]:
        raise ValueError("grid_points must have one coordinate per point of the swept axis")
    out = torch.empty_like(f_c)
    rc = lib().vp_advect_kokkos_f64(f_c.data_ptr(), out.data_ptr(), _lib.i64x4(f_c.shape), axis, grid.data_ptr(),
                                    vel.data_ptr(), _lib.i64x4(strides), float(dt), float(L), float(xmin),
                                    1.0 / float(dx), int(bool(periodic)), stream_handle(dev))
    check(rc, "vp_advect_kokkos")
    return out


def vlasov_step_kokkos(f, dt, G, physics_mode=True):
    """The Kokkos solver-0 step (K2:724-767): X, Y (dt/2); rho and Poisson;
    VX with Ex, VY with Ey (dt); Y, X (dt/2).  G is a ``vlasov2D2V.Grid``.
    Returns (f, Ex, Ey, phi)."""
    half = dt / 2.0
    vx = G.vx.view(-1, 1)  # batch (y, vx, vy) of the x sweep: velocity vx[ivx]
    vy = G.vy              # batch (x, vx, vy) of the y sweep: velocity vy[ivy]
    f = advect_1d_kokkos(f, vx, 0, half, G.x, G.Lx, 0.0, G.dx, True)
    f = advect_1d_kokkos(f, vy, 1, half, G.y, G.Ly, 0.0, G.dy, True)
    if physics_mode:
        Ex, Ey, _, phi = _v2.solve_poisson_real(f, G.kx, G.ky, G.inv_ksq, G.dvx, G.dvy)
    else:
        Ex = torch.zeros((G.Nx, G.Ny), dtype=torch.float64, device=f.device)
        Ey, phi = torch.zeros_like(Ex), torch.zeros_like(Ex)
    f = advect_1d_kokkos(f, Ex.view(G.Nx, G.Ny, 1), 2, dt, G.vx, 2.0 * G.Vx_max, -G.Vx_max, G.dvx, False)
    f = advect_1d_kokkos(f, Ey.view(G.Nx, G.Ny, 1), 3, dt, G.vy, 2.0 * G.Vy_max, -G.Vy_max, G.dvy, False)
    f = advect_1d_kokkos(f, vy, 1, half, G.y, G.Ly, 0.0, G.dy, True)
    f = advect_1d_kokkos(f, vx, 0, half, G.x, G.Lx, 0.0, G.dx, True)
    return f, Ex, Ey, phi
