"""Pins the oracle's diagnostics (field energy, mass, kinetic energy, L2 norm).

The reference defines no diagnostics kernel (its nearest quantities are the
integral of phi plotted against exp(-0.153 t) in
simulations/vlasov1d_1v/postscript/postscript.py:297-300 and the mass sum of
simulations/vlasov1d_1v/jax/tests/test_vlasov_step.py:133-137), so these
functions are pinned by closed forms chosen so that a dropped 1/2, a
missing cell size, dx and dy swapped, vx and vy swapped or v in place of v^2
each fail:

* discrete orthogonality on the periodic grid x_j = j L / N: for 0 < m < N/2,
  sum_j sin^2(2 pi m j / N) = sum_j cos^2(2 pi m j / N) = N / 2 exactly;
* sums of v^2 over the endpoint-inclusive velocity grid
  v_j = -V + j 2V/(N-1): sum_j v_j^2 = N V^2 (N+1) / (3 (N-1));
* a single occupied cell;
* Maxwellians on wide boxes (the trapezoid rule on a smooth function that
  vanishes at the box ends converges faster than any power of dv).
"""
import math

import numpy as np
import pytest

from oracle import vlasov_oracle as vo


@pytest.mark.parametrize("m", [1, 3, 7])
def test_field_energy_1d_single_mode(m):
    N, L, A = 32, 4 * math.pi, 0.37
    x = np.arange(N) * (L / N)
    E = A * np.sin(2 * math.pi * m * x / L)
    assert vo.field_energy_1d(E, L / N) == pytest.approx(A * A * L / 4, rel=1e-14)
    E = A * np.cos(2 * math.pi * m * x / L)
    assert vo.field_energy_1d(E, L / N) == pytest.approx(A * A * L / 4, rel=1e-14)


def test_field_energy_1d_constant_and_scaling():
    N, L = 24, 10 * math.pi
    E = np.full(N, -1.5)
    assert vo.field_energy_1d(E, L / N) == pytest.approx(0.5 * 2.25 * L, rel=1e-15)
    # quadratic in E, linear in dx
    rng = np.random.default_rng(0)
    E = rng.normal(size=N)
    w = vo.field_energy_1d(E, 0.3)
    assert vo.field_energy_1d(3 * E, 0.3) == pytest.approx(9 * w, rel=1e-14)
    assert vo.field_energy_1d(E, 0.6) == pytest.approx(2 * w, rel=1e-14)


@pytest.mark.parametrize("mx,my", [(1, 1), (2, 3), (5, 1)])
def test_field_energy_2d_single_modes(mx, my):
    Nx, Ny, Lx, Ly = 16, 24, 4 * math.pi, 3 * math.pi
    dx, dy = Lx / Nx, Ly / Ny
    x = np.arange(Nx)[:, None] * dx
    y = np.arange(Ny)[None, :] * dy
    kx, ky = 2 * math.pi * mx / Lx, 2 * math.pi * my / Ly
    A, B = 0.8, -0.3
    Ex = A * np.sin(kx * x) * np.cos(ky * y)
    Ey = B * np.cos(kx * x) * np.sin(ky * y)
    want = (A * A + B * B) * Lx * Ly / 8
    assert vo.field_energy_2d(Ex, Ey, dx, dy) == pytest.approx(want, rel=1e-14)
    # one component, constant along y: 1/2 A^2 (Nx/2) Ny dx dy = A^2 Lx Ly / 4
    Ex1 = A * np.sin(kx * x) * np.ones_like(y)
    assert vo.field_energy_2d(Ex1, np.zeros_like(Ex1), dx, dy) == pytest.approx(A * A * Lx * Ly / 4, rel=1e-14)
    assert vo.field_energy_2d(np.zeros_like(Ex1), Ex1, dx, dy) == pytest.approx(A * A * Lx * Ly / 4, rel=1e-14)


def _sum_v2(N, V):
    return N * V * V * (N + 1) / (3 * (N - 1))


def test_sum_v2_closed_form():
    for N, V in [(5, 2.0), (64, 6.0), (127, 8.0)]:
        v = np.linspace(-V, V, N)
        assert float(np.sum(v * v)) == pytest.approx(_sum_v2(N, V), rel=1e-13)


def test_moments_1d_uniform_and_single_cell():
    g = vo.grid_1d(16, 33, 4 * math.pi, 6.0)
    c = 0.25
    f = np.full((g.Nx, g.Nv), c)
    mass, kin, l2 = vo.moments_1d(f, g.v, g.dx, g.dv)
    assert mass == pytest.approx(c * g.Nx * g.Nv * g.dx * g.dv, rel=1e-14)
    assert mass == pytest.approx(c * g.L * 2 * g.V_max * g.Nv / (g.Nv - 1), rel=1e-14)
    assert kin == pytest.approx(0.5 * c * g.Nx * g.dx * _sum_v2(g.Nv, g.V_max) * g.dv, rel=1e-13)
    assert l2 == pytest.approx(c * c * g.Nx * g.Nv * g.dx * g.dv, rel=1e-14)
    f = np.zeros((g.Nx, g.Nv))
    f[3, 29] = 2.0
    mass, kin, l2 = vo.moments_1d(f, g.v, g.dx, g.dv)
    assert mass == pytest.approx(2.0 * g.dx * g.dv, rel=1e-15)
    assert kin == pytest.approx(0.5 * 2.0 * g.v[29] ** 2 * g.dx * g.dv, rel=1e-15)
    assert l2 == pytest.approx(4.0 * g.dx * g.dv, rel=1e-15)


def test_moments_1d_maxwellian_quadrature():
    g = vo.grid_1d(8, 241, 2 * math.pi, 12.0)
    T = 1.7
    fM = np.exp(-0.5 * g.v ** 2 / T) / math.sqrt(2 * math.pi * T)
    f = np.broadcast_to(fM[None, :], (g.Nx, g.Nv))
    mass, kin, l2 = vo.moments_1d(f, g.v, g.dx, g.dv)
    assert mass == pytest.approx(g.L, rel=1e-12)
    assert kin == pytest.approx(0.5 * T * g.L, rel=1e-12)
    assert l2 == pytest.approx(g.L / (2 * math.sqrt(math.pi * T)), rel=1e-12)


def test_moments_2d_uniform_and_single_cell():
    g = vo.grid_2d(8, 6, 17, 13, 4 * math.pi, 3 * math.pi, 6.0, 4.0)
    c = 0.5
    f = np.full((g.Nx, g.Ny, g.Nvx, g.Nvy), c)
    cell = g.dx * g.dy * g.dvx * g.dvy
    n = g.Nx * g.Ny
    mass, kin, l2 = vo.moments_2d(f, g.vx, g.vy, g.dx, g.dy, g.dvx, g.dvy)
    assert mass == pytest.approx(c * n * g.Nvx * g.Nvy * cell, rel=1e-14)
    want = 0.5 * c * n * (_sum_v2(g.Nvx, g.Vx_max) * g.Nvy + _sum_v2(g.Nvy, g.Vy_max) * g.Nvx) * cell
    assert kin == pytest.approx(want, rel=1e-13)
    assert l2 == pytest.approx(c * c * n * g.Nvx * g.Nvy * cell, rel=1e-14)
    f = np.zeros_like(f)
    f[5, 2, 15, 1] = 3.0
    mass, kin, l2 = vo.moments_2d(f, g.vx, g.vy, g.dx, g.dy, g.dvx, g.dvy)
    assert mass == pytest.approx(3.0 * cell, rel=1e-15)
    assert kin == pytest.approx(0.5 * 3.0 * (g.vx[15] ** 2 + g.vy[1] ** 2) * cell, rel=1e-15)
    assert l2 == pytest.approx(9.0 * cell, rel=1e-15)


def test_moments_2d_drifting_maxwellian_quadrature():
    g = vo.grid_2d(4, 4, 201, 161, 2 * math.pi, 3 * math.pi, 12.0, 10.0)
    u, Tx, Ty = 1.3, 1.2, 0.8
    mx = np.exp(-0.5 * (g.vx - u) ** 2 / Tx) / math.sqrt(2 * math.pi * Tx)
    my = np.exp(-0.5 * g.vy ** 2 / Ty) / math.sqrt(2 * math.pi * Ty)
    f = np.broadcast_to((mx[:, None] * my[None, :])[None, None], (g.Nx, g.Ny, g.Nvx, g.Nvy))
    area = g.Lx * g.Ly
    mass, kin, l2 = vo.moments_2d(f, g.vx, g.vy, g.dx, g.dy, g.dvx, g.dvy)
    assert mass == pytest.approx(area, rel=1e-12)
    assert kin == pytest.approx(0.5 * (Tx + u * u + Ty) * area, rel=1e-12)
    assert l2 == pytest.approx(area / (4 * math.pi * math.sqrt(Tx * Ty)), rel=1e-12)
