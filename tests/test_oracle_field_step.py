"""Pins the oracle's field solves and Strang steps.

Closed forms follow the reference tests
(simulations/vlasov1d_1v/pytorch/tests/test_poisson.py:6-64,
simulations/vlasov2d_2v/pytorch/tests/test_poisson.py:6-75).  The dense-DFT
checks pin the rfft/irfft conventions (dropped imaginary parts of the DC and
Nyquist bins), and the Landau run pins the whole step against linear theory.
"""
import math

import numpy as np
import pytest

from oracle import vlasov_oracle as vo


@pytest.mark.parametrize("m", [1, 2, 3])
def test_poisson_1d_single_mode(m):
    g = vo.grid_1d(16, 16, 2.0 * 3.14159, 4.0)
    eps, k = 0.01, m * 2 * math.pi / g.L
    fM = np.exp(-0.5 * g.v ** 2) / math.sqrt(2 * math.pi)
    f = (1 + eps * np.cos(k * g.x))[:, None] * fM[None, :]
    E, rho, phi = vo.solve_poisson_real_1d(f, g.inv_kx, g.inv_kx_sq, g.dv)
    np.testing.assert_allclose(rho, f.sum(1) * g.dv - 1.0, atol=1e-14)
    np.testing.assert_allclose(phi, eps / k ** 2 * np.cos(k * g.x), atol=1e-5)
    np.testing.assert_allclose(E, eps / k * np.sin(k * g.x), atol=1e-5)


@pytest.mark.parametrize("m", [1, 2, 3])
def test_poisson_2d_single_mode(m):
    g = vo.grid_2d(16, 16, 16, 16, 2 * 3.14159265358979, 2 * 3.14159265358979, 4.0, 4.0)
    eps = 0.01
    kx0, ky0 = m * 2 * math.pi / g.Lx, m * 2 * math.pi / g.Ly
    pert = eps * np.cos(kx0 * g.x)[:, None] * np.cos(ky0 * g.y)[None, :]
    Ex, Ey, drho, phi = vo.compute_poisson_real(pert, g.kx, g.ky, g.inv_ksq)
    ks = kx0 ** 2 + ky0 ** 2
    np.testing.assert_allclose(drho, pert - 1, atol=1e-15)
    np.testing.assert_allclose(phi, eps / ks * np.cos(kx0 * g.x)[:, None] * np.cos(ky0 * g.y)[None, :], atol=1e-5)
    np.testing.assert_allclose(Ex, eps * kx0 / ks * np.sin(kx0 * g.x)[:, None] * np.cos(ky0 * g.y)[None, :], atol=1e-5)
    np.testing.assert_allclose(Ey, eps * ky0 / ks * np.cos(kx0 * g.x)[:, None] * np.sin(ky0 * g.y)[None, :], atol=1e-5)
    fM = np.exp(-0.5 * (g.vx[:, None] ** 2 + g.vy[None, :] ** 2)) / math.sqrt(2 * math.pi)
    f = (1 + pert)[:, :, None, None] * fM[None, None]
    np.testing.assert_allclose(vo.compute_rho(f, g.dvx, g.dvy), f.sum((2, 3)) * g.dvx * g.dvy, rtol=1e-13)


def _dense_irfft(X, n):
    """x_j = (1/n)[Re X0 + 2 sum_{0<k<n/2} Re(X_k e^{2 pi i k j/n}) + Re X_{n/2} (-1)^j]."""
    j = np.arange(n)
    out = np.full(n, X[0].real, dtype=np.float64)
    for k in range(1, (n + 1) // 2):
        out += 2 * (X[k] * np.exp(2j * np.pi * k * j / n)).real
    if n % 2 == 0:
        out += X[n // 2].real * (-1.0) ** j
    return out / n


@pytest.mark.parametrize("n", [8, 9, 16])
def test_irfft_drops_imaginary_dc_and_nyquist(n):
    rng = np.random.default_rng(7)
    X = rng.normal(size=n // 2 + 1) + 1j * rng.normal(size=n // 2 + 1)
    np.testing.assert_allclose(np.fft.irfft(X, n=n), _dense_irfft(X, n), atol=1e-13)


def test_irfft2_is_ifft_then_c2r():
    rng = np.random.default_rng(8)
    Nx, Ny = 8, 6
    X = rng.normal(size=(Nx, Ny // 2 + 1)) + 1j * rng.normal(size=(Nx, Ny // 2 + 1))
    col = np.fft.ifft(X, axis=0)
    ref = np.stack([_dense_irfft(col[i], Ny) for i in range(Nx)])
    np.testing.assert_allclose(np.fft.irfft2(X, s=(Nx, Ny)), ref, atol=1e-13)


def test_rfft_matches_dense_dft():
    rng = np.random.default_rng(9)
    x = rng.normal(size=12)
    k = np.arange(7)[:, None]
    j = np.arange(12)[None, :]
    np.testing.assert_allclose(np.fft.rfft(x), (x[None, :] * np.exp(-2j * np.pi * k * j / 12)).sum(1), atol=1e-12)


def _landau_1d(Nx=64, Nv=128, steps=0, dt=0.1):
    g = vo.grid_1d(Nx, Nv, 4 * math.pi, 6.0)
    fM = np.exp(-0.5 * g.v ** 2) / math.sqrt(2 * math.pi)
    f = (1 + 0.01 * np.cos(0.5 * g.x))[:, None] * fM[None, :]
    return g, f


def test_step_1d_dt_zero_is_identity():
    g, f = _landau_1d()
    out, *_ = vo.vlasov_step_1d(f, 0.0, g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv)
    np.testing.assert_array_equal(out, f)


def test_step_1d_conserves_mass():
    g, f = _landau_1d()
    m0 = f.sum()
    for _ in range(5):
        f, *_ = vo.vlasov_step_1d(f, 0.1, g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv)
    assert abs(f.sum() - m0) / m0 < 1e-6


def test_step_1d_free_streaming_is_strang_composition():
    g, f = _landau_1d()
    out, *_ = vo.vlasov_step_1d(f, 0.2, g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv, physics_mode=False)
    h = vo.advect_rows(f.T, g.v, 0.1, g.dx, periodic=True)
    h = vo.advect_rows(h, g.v, 0.1, g.dx, periodic=True).T
    np.testing.assert_array_equal(out, h)


def test_landau_damping_rate_1d():
    """Field energy decays at 2*gamma with gamma from the dispersion relation."""
    om = vo.landau_root(0.5)
    assert abs(om.real - 1.4156) < 1e-3 and abs(-om.imag - 0.1533) < 1e-3
    g, f = _landau_1d(Nx=32, Nv=128)
    dt, W = 0.1, []
    for _ in range(300):
        f, rho, phi, E = vo.vlasov_step_1d(f, dt, g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv)
        W.append(vo.field_energy_1d(E, g.dx))
    W = np.log(np.array(W))
    t = dt * (np.arange(300) + 0.5)
    peaks = [i for i in range(1, 299) if W[i] > W[i - 1] and W[i] > W[i + 1] and t[i] < 25]
    slope = np.polyfit(t[peaks], W[peaks], 1)[0]
    assert abs(-slope / 2 - 0.1533) < 0.01, slope


def test_step_2d_matches_separable_1d():
    """(x, vx)-only data, constant in (y, vy): the 2D step reduces to the 1D step."""
    g = vo.grid_2d(8, 4, 12, 6, 4 * math.pi, 4 * math.pi, 6.0, 6.0)
    g1 = vo.grid_1d(8, 12, 4 * math.pi, 6.0)
    fv = np.exp(-0.5 * g.vx ** 2) / math.sqrt(2 * math.pi)
    f1 = (1 + 0.05 * np.cos(0.5 * g.x))[:, None] * fv[None, :]
    c = 1.0 / (g.Nvy * g.dvy)
    f2 = np.ascontiguousarray(np.broadcast_to(f1[:, None, :, None] * c, (8, 4, 12, 6)))
    out2, drho2, phi2, Ex, Ey = vo.vlasov_step_2d(f2, 0.1, g, physics_mode=True)
    out1, rho1, phi1, E1 = vo.vlasov_step_1d(f1, 0.1, g1.v, g1.inv_kx, g1.inv_kx_sq, g1.dx, g1.dv)
    np.testing.assert_allclose(Ey, 0.0, atol=1e-15)
    np.testing.assert_allclose(Ex, np.broadcast_to(E1[:, None], Ex.shape), rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(drho2, np.broadcast_to(rho1[:, None], drho2.shape), atol=1e-14)
    np.testing.assert_allclose(out2 / c, np.broadcast_to(out1[:, None, :, None], out2.shape), rtol=1e-12, atol=1e-15)
