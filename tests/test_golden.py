"""The committed fixtures of tests/golden/ (written by tests/golden/make_golden.py
from the float64 oracle) against the oracle and against the physics they
must show.

* the oracle re-derives them (600-step 1D Landau run in full, the first two
  2D steps) to 1e-12 of the largest value (NumPy's FFT and sum may round
  differently across versions);
* the 1D W_E history decays at the Landau rate: 2 gamma with gamma = 0.1533
  at k = 0.5 (root of the dispersion relation), within 3 %;
* mass is conserved step by step in 2D to the Dirichlet-boundary loss.
"""
import math
import pathlib
import sys

import numpy as np
import pytest

from oracle import vlasov_oracle as vo

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))

import make_golden as mg  # noqa: E402


def _load(name):
    with np.load(GOLDEN / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_landau1d_fixture_rederived_by_oracle():
    fx = _load("landau1d_128.npz")
    got = mg.landau1d()
    W, Wg = fx["W_E"], got["W_E"]
    assert W.shape == (mg.LANDAU1D["steps"] + 1,)
    assert np.max(np.abs(Wg - W)) / np.max(W) <= 1e-12
    for k in ("f10", "f_last"):
        assert np.max(np.abs(got[k] - fx[k])) / np.max(np.abs(fx[k])) <= 1e-12
    np.testing.assert_allclose(got["sum_last"], fx["sum_last"], rtol=1e-12)


def test_landau1d_fixture_damps_at_landau_rate():
    W = _load("landau1d_128.npz")["W_E"]
    dt = mg.LANDAU1D["dt"]
    t = np.arange(len(W)) * dt
    # local maxima of W_E between t = 5 and t = 35 (linear phase)
    i = np.array([j for j in range(1, len(W) - 1) if W[j] > W[j - 1] and W[j] >= W[j + 1]])
    i = i[(t[i] > 5) & (t[i] < 35)]
    assert len(i) >= 4
    slope = np.polyfit(t[i], np.log(W[i]), 1)[0]
    gamma = -vo.landau_root(0.5).imag
    assert abs(-slope / 2 - gamma) / gamma <= 0.03


def test_landau2d_fixture_rederived_by_oracle():
    fx = _load("landau2d_32.npz")
    got = mg.landau2d(steps=2)
    assert np.max(np.abs(got["W_E"] - fx["W_E"][:3])) / np.max(fx["W_E"][:3]) <= 1e-12
    np.testing.assert_allclose(got["sums"], fx["sums"][:2], rtol=1e-12)
    assert np.max(np.abs(got["f2_sample"] - fx["f2_sample"])) / np.max(fx["f2_sample"]) <= 1e-12


def test_landau2d_fixture_conserves_mass():
    fx = _load("landau2d_32.npz")
    g, f0 = mg.landau2d_initial()
    m0 = f0.sum()
    masses = fx["sums"][:, 0]
    assert np.max(np.abs(masses - m0)) / m0 <= 1e-9
    assert fx["W_E"].shape == (mg.LANDAU2D["steps"] + 1,)
