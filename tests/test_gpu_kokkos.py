"""GPU parity of the Kokkos-semantics sweep and step against their oracle.

The sweep follows the oracle's operation order with no fused multiply-add,
so it must be bit-identical; the step adds the field solve, whose sums and
transforms round differently (<= 1e-12 relative over the step).
"""
import math

import numpy as np
import pytest
import torch

from oracle import kokkos_oracle as ko
from oracle import vlasov_oracle as vo

pytestmark = pytest.mark.gpu

SHAPE = (16, 8, 12, 10)


def _axis_setup(axis, rng):
    """(grid points, L, xmin, dx, periodic, velocities over the batch) of a sweep."""
    n = SHAPE[axis]
    if axis < 2:
        L = 4 * math.pi if axis == 0 else 2 * math.pi
        dx = L / n
        grid = np.arange(n) * dx
        xmin, periodic = 0.0, True
    else:
        vmax = 6.0
        L, dx = 2 * vmax, 2 * vmax / n
        grid = -vmax + np.arange(n) * dx
        xmin, periodic = -vmax, False
    batch = tuple(s for i, s in enumerate(SHAPE) if i != axis)
    vel = rng.uniform(-8.0, 8.0, size=batch)
    return grid, L, xmin, dx, periodic, vel


@pytest.mark.parametrize("axis", [0, 1, 2, 3])
@pytest.mark.parametrize("dt", [0.05, 0.9, -2.3])
def test_kokkos_sweep_is_bit_identical_to_oracle(device, axis, dt):
    from vlasov_amd import kokkos_semantics as ks
    rng = np.random.default_rng(axis * 7 + abs(int(dt * 10)))
    f = rng.random(SHAPE)
    grid, L, xmin, dx, periodic, vel = _axis_setup(axis, rng)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    out = ks.advect_1d_kokkos(t(f), t(vel), axis, dt, t(grid), L, xmin, dx, periodic)
    ref = ko.advect_axis_kokkos(f, vel, axis, dt, grid, L, xmin, 1.0 / dx, periodic)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    # the same sweep with the velocity broadcast along a batch dim
    vb = vel[..., :1]
    out = ks.advect_1d_kokkos(t(f), t(vb), axis, dt, t(grid), L, xmin, dx, periodic)
    ref = ko.advect_axis_kokkos(f, vb, axis, dt, grid, L, xmin, 1.0 / dx, periodic)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)


def test_kokkos_step_matches_oracle_composition(device):
    from vlasov_amd import kokkos_semantics as ks
    from vlasov_amd import vlasov2D2V as v2
    Nx, Ny, Nvx, Nvy = 16, 8, 24, 20
    G = v2.Grid.create(Nx, Ny, Nvx, Nvy, 10 * math.pi, 10 * math.pi, 6.0, 6.0, device=device)
    g = vo.grid_2d(Nx, Ny, Nvx, Nvy, 10 * math.pi, 10 * math.pi, 6.0, 6.0)
    x, y = g.x[:, None, None, None], g.y[None, :, None, None]
    vx, vy = g.vx[None, None, :, None], g.vy[None, None, None, :]
    f0 = (1 + 0.05 * np.cos(0.2 * x) + 0.03 * np.cos(0.2 * y)) * np.exp(-0.5 * (vx ** 2 + vy ** 2)) / (2 * math.pi)
    dt = 0.1
    out, Ex, Ey, phi = ks.vlasov_step_kokkos(torch.from_numpy(f0).to(device), dt, G, True)

    # the grid is an input of the sweep: the oracle takes the library's points
    # (torch.linspace and numpy.linspace differ in the last bits)
    gx, gy, gvx, gvy = (a.cpu().numpy() for a in (G.x, G.y, G.vx, G.vy))
    f = ko.advect_axis_kokkos(f0, gvx[:, None], 0, dt / 2, gx, G.Lx, 0.0, 1.0 / G.dx, True)
    f = ko.advect_axis_kokkos(f, gvy, 1, dt / 2, gy, G.Ly, 0.0, 1.0 / G.dy, True)
    rEx, rEy, _, _ = vo.compute_poisson_real(vo.compute_rho(f, g.dvx, g.dvy), g.kx, g.ky, g.inv_ksq)
    Exg, Eyg = Ex.cpu().numpy(), Ey.cpu().numpy()
    assert np.max(np.abs(Exg - rEx)) <= 1e-13 * np.max(np.abs(rEx))
    assert np.max(np.abs(Eyg - rEy)) <= 1e-13 * np.max(np.abs(rEy))
    # continue the oracle from the library's field, so the sweeps compare exactly
    f = ko.advect_axis_kokkos(f, Exg[:, :, None], 2, dt, gvx, 2 * G.Vx_max, -G.Vx_max, 1.0 / G.dvx, False)
    f = ko.advect_axis_kokkos(f, Eyg[:, :, None], 3, dt, gvy, 2 * G.Vy_max, -G.Vy_max, 1.0 / G.dvy, False)
    f = ko.advect_axis_kokkos(f, gvy, 1, dt / 2, gy, G.Ly, 0.0, 1.0 / G.dy, True)
    f = ko.advect_axis_kokkos(f, gvx[:, None], 0, dt / 2, gx, G.Lx, 0.0, 1.0 / G.dx, True)
    np.testing.assert_array_equal(out.cpu().numpy(), f)
    # and it is the same scheme as the library's PyTorch-form step, to interpolation accuracy
    ref_step, *_ = v2.step_with_field(torch.from_numpy(f0).to(device), dt, G.vx, G.vy, G.kx, G.ky, G.inv_ksq,
                                      G.dx, G.dy, G.dvx, G.dvy, True)
    assert float((out - ref_step).abs().max() / ref_step.abs().max()) <= 1e-6
