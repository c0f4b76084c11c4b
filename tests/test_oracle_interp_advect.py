"""Pins the oracle's interpolation and advection against closed forms.

Known-answer cases follow the reference tests
(simulations/vlasov1d_1v/pytorch/tests/test_interpolation.py:4-65,
test_advection.py:50-195 and the 2D2V counterparts); the brute-force dense
interpolation matrix and permutation checks catch transposed operands, wrong
tap offsets and sign errors that the polynomial checks alone would not.
"""
import math

import numpy as np
import pytest

from oracle import vlasov_oracle as vo

ALPHAS = np.array([0.1, 0.25, 0.5, 0.75, 0.9])


@pytest.mark.parametrize("deg", range(6))
def test_interp_exact_on_polynomials(deg):
    nodes = np.arange(-2, 4, dtype=np.float64)
    vals = np.broadcast_to(nodes ** deg, (ALPHAS.size, 6))
    out = vo.lagrange5_interp_sum(ALPHAS, vals)
    np.testing.assert_allclose(out, ALPHAS ** deg, atol=1e-12, rtol=0)


def test_interp_degree6_is_not_exact():
    # the basis is exactly degree 5: x^6 must leave a nonzero residual
    nodes = np.arange(-2, 4, dtype=np.float64)
    out = vo.lagrange5_interp_sum(ALPHAS, np.broadcast_to(nodes ** 6, (ALPHAS.size, 6)))
    assert np.max(np.abs(out - ALPHAS ** 6)) > 1e-3


def test_interp_alpha_zero_is_centre_tap():
    vals = np.array([[7.0, 11.0, 42.0, -1.0, 3.0, 9.0]])
    assert vo.lagrange5_interp_sum(np.array([0.0]), vals)[0] == 42.0


def test_partition_of_unity():
    a = np.random.default_rng(0).random(1000)
    w = vo.lagrange5_weights(a)
    np.testing.assert_allclose(sum(w), 1.0, atol=1e-14)


def test_power_form_matches_product_form():
    a = np.random.default_rng(1).random(1000)
    for wp, wq in zip(vo.lagrange5_weights(a), vo.lagrange5_product_form(a)):
        np.testing.assert_allclose(wp, wq, atol=2e-15)


def _dense_shift_matrix(N, d, periodic):
    """Exact-basis interpolation matrix of one row shifted by d cells."""
    M = np.zeros((N, N))
    for i in range(N):
        dep = i - d
        il = math.floor(dep)
        a = dep - il
        w = vo.lagrange5_product_form(np.array([a]))
        for k, o in enumerate(range(-2, 4)):
            j = il + o
            if periodic:
                M[i, j % N] += w[k][0]
            elif 0 <= j < N:
                M[i, j] += w[k][0]
    return M


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("d", [0.3, -0.7, 1.6, -2.4, 5.25])
def test_advect_rows_matches_dense_matrix(periodic, d):
    rng = np.random.default_rng(2)
    N = 13
    f = rng.random((3, N))
    dt, dx = 0.1, 0.2
    vel = np.full(3, d * dx / dt)
    out = vo.advect_rows(f, vel, dt, dx, periodic=periodic)
    dd = (vel[0] * dt) / dx
    M = _dense_shift_matrix(N, dd, periodic)
    np.testing.assert_allclose(out, f @ M.T, atol=1e-13)


def test_integer_shift_is_permutation():
    rng = np.random.default_rng(3)
    f = rng.random((4, 16))
    dt, dx = 0.5, 0.25
    for s in (-3, -1, 0, 2, 5):
        out = vo.advect_rows(f, np.full(4, s * dx / dt), dt, dx, periodic=True)
        np.testing.assert_array_equal(out, np.roll(f, s, axis=1))


def test_periodic_drift_conserves_sum():
    rng = np.random.default_rng(4)
    f = rng.random((5, 32))
    vel = rng.normal(size=5)
    out = vo.advect_rows(f, vel, 0.37, 0.1, periodic=True)
    np.testing.assert_allclose(out.sum(axis=1), f.sum(axis=1), rtol=1e-13)


def test_per_point_velocity_rows():
    rng = np.random.default_rng(5)
    f = rng.random((2, 10))
    vel = rng.normal(size=(2, 10))
    out = vo.advect_rows(f, vel, 0.1, 0.3, periodic=False)
    for b in range(2):
        for i in range(10):
            one = vo.advect_rows(f[b:b + 1], np.full((1, 10), vel[b, i]), 0.1, 0.3, periodic=False)
            assert out[b, i] == one[0, i]


def test_gaussian_drift_periodic_1d():
    g = vo.grid_1d(16, 16, 2.0 * 3.14159, 4.0)
    x0, sigma = g.L / 2, g.L / 10
    f = np.exp(-0.5 * ((g.x - x0) / sigma) ** 2)[:, None].repeat(g.Nv, 1)
    out = vo.advect_rows(f.T, np.full(g.Nv, 1.0), 0.1, g.dx, periodic=True).T
    xd = g.x[:, None] - (x0 + 0.1)
    xd = xd - g.L * np.round(xd / g.L)
    np.testing.assert_allclose(out, np.exp(-0.5 * (xd / sigma) ** 2).repeat(g.Nv, 1), atol=1e-2)


@pytest.mark.parametrize("shift_factor", [0.5, 1.0, 2.7, 4.2])
def test_gaussian_kick_dirichlet_1d(shift_factor):
    g = vo.grid_1d(16, 16, 2.0 * 3.14159, 4.0)
    sigma = g.V_max / 4
    f = np.exp(-0.5 * (g.v / sigma) ** 2)[None, :].repeat(g.Nx, 0)
    a = shift_factor * g.dv / 0.1
    out = vo.advect_rows(f, np.full(g.Nx, a), 0.1, g.dv, periodic=False)
    dep = g.v - a * 0.1
    exp = np.where((dep >= -g.V_max) & (dep <= g.V_max), np.exp(-0.5 * (dep / sigma) ** 2), 0.0)
    np.testing.assert_allclose(out, exp[None, :].repeat(g.Nx, 0), atol=1e-2)
    out = vo.advect_rows(f, np.full(g.Nx, 1.0), 3.0 * g.V_max, g.dv, periodic=False)
    np.testing.assert_allclose(out, 0.0, atol=1e-6)


@pytest.mark.parametrize("axis", [0, 1, 2, 3])
def test_advect_axis_equals_rows_on_moved_axis(axis):
    rng = np.random.default_rng(6)
    shape = (5, 6, 7, 8)
    f = rng.random(shape)
    batch = tuple(s for i, s in enumerate(shape) if i != axis)
    vel = rng.normal(size=batch)
    out = vo.advect_axis(f, vel, axis, 0.1, 0.05, periodic=axis < 2)
    moved = np.moveaxis(f, axis, -1).reshape(-1, shape[axis])
    ref = vo.advect_rows(moved, vel.reshape(-1), 0.1, 0.05, periodic=axis < 2)
    np.testing.assert_array_equal(np.moveaxis(out, axis, -1).reshape(-1, shape[axis]), ref)
    # a pure x-dependence must travel only along x
    if axis == 0:
        prof = np.sin(2 * np.pi * np.arange(5) / 5)[:, None, None, None] * np.ones(shape)
        o2 = vo.advect_axis(prof, np.zeros(batch), 0, 0.1, 0.05, periodic=True)
        np.testing.assert_array_equal(o2, prof)


@pytest.mark.parametrize("periodic", [True, False])
def test_line_form_equals_direct_form_on_broadcast_velocities(periodic):
    """advect_axis (weights on the compact velocity shape) == advect_axis_direct
    (per-element weights on [B, N, 6] taps) for the velocity layouts of the
    steps: drift (vx[None, :, None] over axis 0, vy[None, None, :] over axis
    1), kick (E[:, :, None] over axes 2 and 3), 1D rows; displacements up to
    many cells and beyond the line (Dirichlet out of domain), both signs."""
    rng = np.random.default_rng(7)
    shape = (6, 5, 9, 8)
    f = rng.random(shape)
    dx = 0.05
    big = np.concatenate([rng.normal(size=20) * 0.3, [0.0, 1e-18, -1e-18, 2.5, -2.5, 11.0, -11.0, 40.0]]) * dx / 0.1
    cases = [
        (0, rng.choice(big, size=9)[None, :, None]),
        (1, rng.choice(big, size=8)[None, None, :]),
        (2, rng.choice(big, size=(6, 5))[:, :, None]),
        (3, rng.choice(big, size=(6, 5))[:, :, None]),
        (0, rng.choice(big, size=(5, 9, 8))),
    ]
    for axis, vel in cases:
        a = vo.advect_axis(f, vel, axis, 0.1, dx, periodic=periodic)
        b = vo.advect_axis_direct(f, vel, axis, 0.1, dx, periodic=periodic)
        np.testing.assert_array_equal(a, b)
    f2 = rng.random((7, 11))
    vel = rng.choice(big, size=7)
    np.testing.assert_array_equal(vo.advect_axis(f2, vel, 1, 0.1, dx, periodic),
                                  vo.advect_rows(f2, vel, 0.1, dx, periodic))


@pytest.mark.parametrize("periodic", [True, False])
def test_threaded_blocks_equal_direct_form(monkeypatch, periodic):
    """The blocked, threaded sweep of large arrays computes every element by
    the same operations: forced on for a small array, it equals the direct
    form for every velocity layout (blocks along the velocity's own axis and
    along a broadcast one)."""
    monkeypatch.setattr(vo, "CHUNK_MIN_ELEMS", 1)
    rng = np.random.default_rng(8)
    shape = (7, 9, 6, 10)
    f = rng.random(shape)
    dx = 0.05
    for axis, vel in [(0, rng.normal(size=6)[None, :, None] * 3), (1, rng.normal(size=10)[None, None, :] * 3),
                      (2, rng.normal(size=(7, 9))[:, :, None] * 3), (3, rng.normal(size=(7, 9))[:, :, None] * 3),
                      (1, rng.normal(size=(7, 6, 10)) * 3)]:
        np.testing.assert_array_equal(vo.advect_axis(f, vel, axis, 0.1, dx, periodic),
                                      vo.advect_axis_direct(f, vel, axis, 0.1, dx, periodic))
