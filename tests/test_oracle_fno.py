"""CPU checks that pin the surrogate's float64 oracle (oracle/fno_oracle.py).

* The truncated inverse DFT G undoes the forward DFT F on signals made of
  the kept modes (closed form), and G F is a projection (idempotent).
* Every gradient of ``backward`` matches central finite differences of the
  forward loss (relative 1e-6), for every parameter tensor.
* With zero spectral weights, the layer reduces to a pointwise MLP that a
  plain matrix product restates.
"""
import math

import numpy as np
import pytest

from oracle import fno_oracle as fo


def test_dft_mats_invert_band_limited_signals():
    T, M = 40, 7
    Fc, Fs, Gc, Gs = fo.dft_mats(T, M)
    rng = np.random.default_rng(0)
    t = np.arange(T)
    a = rng.standard_normal(M)
    b = rng.standard_normal(M)
    b[0] = 0.0
    x = sum(a[m] * np.cos(2 * math.pi * m * t / T) + b[m] * np.sin(2 * math.pi * m * t / T) for m in range(M))
    rec = Gc @ (Fc @ x) + Gs @ (Fs @ x)
    np.testing.assert_allclose(rec, x, atol=1e-12)
    P = Gc @ Fc + Gs @ Fs
    np.testing.assert_allclose(P @ P, P, atol=1e-12)
    # a mode above the cut is removed
    hi = np.cos(2 * math.pi * (M + 2) * t / T)
    np.testing.assert_allclose(Gc @ (Fc @ hi) + Gs @ (Fs @ hi), 0.0, atol=1e-12)


def test_backward_matches_finite_differences():
    B, T, cin, C, M, Hd, L = 2, 12, 3, 4, 3, 5, 2
    p = fo.init_params(cin, C, M, Hd, L, seed=1)
    for k in p:  # make the spectral weights comparable to the others
        if k.startswith("Wr") or k.startswith("Wi"):
            p[k] = p[k] * C * C * 0.3
    mats = fo.dft_mats(T, M)
    rng = np.random.default_rng(2)
    X0 = rng.standard_normal((B, T, cin))
    target = rng.standard_normal((B, T))

    def loss(pp):
        y, _ = fo.forward(pp, X0, mats)
        return fo.mse(y, target)[0]

    y, cache = fo.forward(p, X0, mats)
    _, dy = fo.mse(y, target)
    g = fo.backward(p, cache, dy, mats)
    assert set(g) == set(p)
    h = 1e-6
    for name, w in p.items():
        flat = w.reshape(-1)
        idx = rng.choice(flat.size, size=min(6, flat.size), replace=False)
        for i in idx:
            old = flat[i]
            flat[i] = old + h
            lp = loss(p)
            flat[i] = old - h
            lm = loss(p)
            flat[i] = old
            fd = (lp - lm) / (2 * h)
            an = g[name].reshape(-1)[i]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, i, fd, an)


def test_zero_spectral_weights_is_pointwise_mlp():
    B, T, cin, C, M, Hd, L = 3, 10, 2, 5, 4, 6, 3
    p = fo.init_params(cin, C, M, Hd, L, seed=3)
    for l in range(L):
        p[f"Wr{l}"][:] = 0.0
        p[f"Wi{l}"][:] = 0.0
    X0 = np.random.default_rng(4).standard_normal((B, T, cin))
    y, _ = fo.forward(p, X0, fo.dft_mats(T, M))
    H = X0 @ p["P"] + p["bP"]
    for l in range(L):
        Z = H @ p[f"W{l}"] + p[f"b{l}"]
        H = fo.gelu(Z) if l < L - 1 else Z
    ref = (fo.gelu(H @ p["Q1"] + p["bQ1"]) @ p["Q2"] + p["bQ2"])[..., 0]
    np.testing.assert_allclose(y, ref, atol=1e-12)


def test_f0_moments_closed_forms():
    """Two-beam f0 on a wide velocity box: (rho_0, rho_1 cos, rho_1 sin, M0,
    M2, M4) = (1, A/2, 0, 1, s^2 + u^2, u^4 + 6 u^2 s^2 + 3 s^4) to
    quadrature accuracy, and decoding gives (A, u, s) back; a shifted phase
    moves the mode into the sine part."""
    rng = np.random.default_rng(0)
    nx, nv = 32, 256
    for _ in range(4):
        A, u, s = rng.uniform(0.01, 0.5), rng.uniform(0.0, 3.0), rng.uniform(0.8, 1.2)
        V = 10 * s + u
        v = -V + np.arange(nv) * (2 * V / (nv - 1))
        M = lambda w: np.exp(-0.5 * (w / s) ** 2) / (s * math.sqrt(2 * math.pi))
        fv = 0.5 * (M(v - u) + M(v + u))
        x = 2 * math.pi * np.arange(nx) / nx
        f0 = ((1 + A * np.cos(x))[:, None] * fv[None, :])[None]
        m = fo.f0_moments(f0, [V])[0]
        want = (1.0, A / 2, 0.0, 1.0, s * s + u * u, u ** 4 + 6 * u * u * s * s + 3 * s ** 4)
        np.testing.assert_allclose(m, want, rtol=1e-10, atol=1e-12)
        A2, u2, s2 = fo.decode_f0_moments(m[None])
        np.testing.assert_allclose([A2[0], s2[0]], [A, s], rtol=1e-8)
        assert abs(u2[0] - u) < 1e-3 * max(u, 1.0) + 2e-3  # u^4 near 0 amplifies rounding for small u
        f1 = ((1 + A * np.sin(x))[:, None] * fv[None, :])[None]
        m1 = fo.f0_moments(f1, [V])[0]
        np.testing.assert_allclose(m1[1:3], [0.0, -A / 2], atol=1e-12)
