"""The PyTorch-CPU restatement (the timed CPU baseline) against the NumPy oracle.

Sweeps: array-equal (same element-wise operations; the row blocks of the
torch form forced small so that several blocks run).  Steps: the sums and
FFTs differ (torch.sum / torch.fft against numpy.sum / pocketfft), so f,
fields and charge density agree to 1e-13.
"""
import math

import numpy as np
import pytest
import torch

from oracle import torch_cpu as tc
from oracle import vlasov_oracle as vo


@pytest.mark.parametrize("periodic", [True, False])
def test_sweeps_equal_oracle(monkeypatch, periodic):
    monkeypatch.setattr(tc, "ROW_BLOCK_ELEMS", 6 * 7 * 5)
    rng = np.random.default_rng(0)
    shape = (6, 5, 9, 7)
    f = rng.random(shape)
    big = np.concatenate([rng.normal(size=12) * 0.3, [0.0, 2.5, -2.5, 11.0, -40.0]]) * 0.05 / 0.1
    for axis, vel in [(0, rng.choice(big, size=9)[None, :, None]), (1, rng.choice(big, size=7)[None, None, :]),
                      (2, rng.choice(big, size=(6, 5))[:, :, None]), (3, rng.choice(big, size=(6, 5))[:, :, None])]:
        a = tc.advect_axis(torch.from_numpy(f), torch.from_numpy(vel), axis, 0.1, 0.05, periodic).numpy()
        b = vo.advect_axis(f, vel, axis, 0.1, 0.05, periodic)
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("physics_mode", [False, True])
def test_step_2d_matches_oracle(physics_mode):
    shape = (8, 16, 12, 10)
    g = vo.grid_2d(*shape, 4 * math.pi, 3 * math.pi, 6.0, 5.0)
    rng = np.random.default_rng(1)
    f = (1 + 0.1 * rng.random(shape)) * np.exp(-0.5 * (g.vx[:, None] ** 2 + g.vy[None, :] ** 2))[None, None]
    ref = vo.vlasov_step_2d(f, 0.1, g, physics_mode)
    T = torch.from_numpy
    out = tc.vlasov_step_2d(T(f), 0.1, T(g.vx), T(g.vy), T(g.kx), T(g.ky), T(g.inv_ksq), g.dx, g.dy, g.dvx, g.dvy,
                            physics_mode)
    if not physics_mode:
        np.testing.assert_array_equal(out[0].numpy(), ref[0])
        return
    for o, r in zip(out, ref):
        o = o.numpy()
        assert np.max(np.abs(o - r)) <= 1e-13 * max(np.max(np.abs(r)), 1.0)


def test_step_1d_matches_oracle():
    g = vo.grid_1d(32, 48, 4 * math.pi, 6.0)
    f = (1 + 0.05 * np.cos(0.5 * g.x))[:, None] * np.exp(-0.5 * g.v ** 2)[None, :] / math.sqrt(2 * math.pi)
    fo, ft = f, torch.from_numpy(f)
    T = torch.from_numpy
    for _ in range(5):
        fo, rho_o, phi_o, E_o = vo.vlasov_step_1d(fo, 0.1, g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv)
        ft, rho_t, phi_t, E_t = tc.vlasov_step_1d(ft, 0.1, T(g.v), T(g.inv_kx), T(g.inv_kx_sq), g.dx, g.dv)
    assert np.max(np.abs(ft.numpy() - fo)) <= 1e-13 * np.max(fo)
    for a, b in ((rho_t, rho_o), (phi_t, phi_o), (E_t, E_o)):
        assert np.max(np.abs(a.numpy() - b)) <= 1e-13 * np.max(np.abs(b))
