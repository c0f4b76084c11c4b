"""CPU checks that pin the oracle's solver-1/2 (reflected stencil) forms.

* integer displacements are exact permutations (rolls), as for solver 0;
* the reflected form agrees with the departure form (solver 0) to
  roundoff: the reference's own tests allow 1e-10 (T2/test_advection.py:526);
* with |E dt/dv| < S the windowed kick equals the unwindowed one exactly,
  beyond it the dropped taps change the result (silently, as in the
  reference);
* advect_x_static with precompute_x_drift's data equals the reflected
  sweep bit for bit; build_v_tap_weights rows sum to 1 (partition of
  unity) while the window holds.
"""
import math

import numpy as np

from oracle import vlasov_oracle as vo


def _smooth(n, m, seed=0):
    rng = np.random.default_rng(seed)
    x = np.arange(n)[:, None] / n
    return np.sin(2 * math.pi * x + rng.random((1, m))) + 1.5


def test_integer_shift_is_roll():
    f = np.random.default_rng(1).random((8, 16))
    out = vo.advect_rows_reflected(f, np.full(8, 3.0), 1.0, 1.0, periodic=True)
    np.testing.assert_array_equal(out, np.roll(f, 3, axis=1))
    out = vo.advect_rows_reflected(f, np.full(8, -2.0), 1.0, 1.0, periodic=False)
    ref = np.zeros_like(f)
    ref[:, :-2] = f[:, 2:]
    np.testing.assert_array_equal(out, ref)


def test_reflected_matches_departure_form():
    f = _smooth(64, 10)
    v = np.linspace(-3.3, 3.3, 10)
    a = vo.advect_rows_reflected(f.T, v, 0.37, 0.11, periodic=True)
    b = vo.advect_rows(f.T, v, 0.37, 0.11, periodic=True)
    assert np.max(np.abs(a - b)) <= 1e-12
    E = np.linspace(-1, 1, 10) * 0.3
    a = vo.advect_rows_reflected(f.T, E, 0.2, 0.1, periodic=False)
    b = vo.advect_rows(f.T, E, 0.2, 0.1, periodic=False)
    assert np.max(np.abs(a - b)) <= 1e-12


def test_window_exact_inside_and_truncating_outside():
    f = _smooth(40, 6)
    E_in = np.array([-0.95, -0.4, 0.0, 0.3, 0.99, 0.7])  # |d| < 1 with dt = dv = 1
    a = vo.advect_rows_reflected(f.T, E_in, 1.0, 1.0, periodic=False, window=1)
    b = vo.advect_rows_reflected(f.T, E_in, 1.0, 1.0, periodic=False, window=-1)
    np.testing.assert_array_equal(a, b)
    E_out = np.array([2.5, -3.2, 0.0, 0.1, 0.2, 0.3])
    a = vo.advect_rows_reflected(f.T, E_out, 1.0, 1.0, periodic=False, window=1)
    b = vo.advect_rows_reflected(f.T, E_out, 1.0, 1.0, periodic=False, window=-1)
    assert np.max(np.abs(a[:2] - b[:2])) > 1e-3
    np.testing.assert_array_equal(a[2:], b[2:])


def test_static_x_drift_equals_reflected_sweep():
    f = _smooth(32, 24, seed=2)
    v = np.linspace(-6, 6, 24)
    blocks, W = vo.precompute_x_drift(v, 0.05, 4 * math.pi / 32)
    assert sum(b - a for a, b, _ in blocks) == 24
    a = vo.advect_x_static(f, W, blocks)
    b = vo.advect_rows_reflected(f.T, v, 0.05, 4 * math.pi / 32, periodic=True).T
    np.testing.assert_array_equal(a, b)


def test_tap_weights_partition_of_unity():
    E = np.linspace(-0.9, 0.9, 11)
    Wt = vo.build_v_tap_weights(E, 0.5, 0.5, 1)
    assert Wt.shape == (7, 11)
    np.testing.assert_allclose(Wt.sum(axis=0), 1.0, atol=1e-14)
