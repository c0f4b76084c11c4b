"""Float64 NumPy restatement of the Fourier-operator surrogate (forward and
backward), the CPU oracle for the GPU training and inference kernels.

TEST INFRASTRUCTURE ONLY: imported by tests/ (gradient checks of the GPU
network against this file, finite-difference checks of this file).  The
surrogate itself runs only on the GPU kernels of the library.

No reference counterpart: the reference has no learned model.  The network
is a 1D Fourier neural operator over the time axis:

    H0 = X0 P + bP                                  (lift, X0: [B, T, cin])
    for l < L:  Z = G(R_l(F H)) + H W_l + b_l,      H = gelu(Z) (l < L-1), Z (l = L-1)
    y = gelu(H Q1 + bQ1) Q2 + bQ2                   (projection, y: [B, T])

F keeps the first M real-DFT modes along T (Fc = cos, Fs = -sin), R_l
multiplies mode m by the complex C x C matrix Wr_l[m] + i Wi_l[m], and G is
the matching inverse real DFT (irfft of the truncated spectrum: modes
1..M-1 count twice, the imaginary part of mode 0 is dropped, scale 1/T).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf


def dft_mats(T: int, M: int):
    """Fc, Fs [M, T]; Gc, Gs [T, M] (see the module docstring)."""
    t = np.arange(T)[None, :]
    m = np.arange(M)[:, None]
    ang = 2.0 * math.pi * m * t / T
    Fc = np.cos(ang)
    Fs = -np.sin(ang)
    a = np.where(np.arange(M) == 0, 1.0, 2.0)[None, :]
    Gc = a * np.cos(ang.T) / T
    Gs = -a * np.sin(ang.T) / T
    return Fc, Fs, Gc, Gs


def init_params(cin: int, width: int, modes: int, hidden: int, layers: int, seed: int = 0) -> dict:
    rng = np.random.default_rng(seed)
    C, M = width, modes

    def u(shape, fan_in):
        b = 1.0 / math.sqrt(fan_in)
        return rng.uniform(-b, b, size=shape)

    p = {"P": u((cin, C), cin), "bP": u((C,), cin), "Q1": u((C, hidden), C), "bQ1": u((hidden,), C),
         "Q2": u((hidden, 1), hidden), "bQ2": u((1,), hidden)}
    for l in range(layers):
        scale = 1.0 / (C * C)
        p[f"Wr{l}"] = scale * rng.uniform(0, 1, size=(M, C, C))
        p[f"Wi{l}"] = scale * rng.uniform(0, 1, size=(M, C, C))
        p[f"W{l}"] = u((C, C), C)
        p[f"b{l}"] = u((C,), C)
    return p


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def n_layers(p: dict) -> int:
    return sum(1 for k in p if k.startswith("Wr"))


def forward(p: dict, X0: np.ndarray, mats):
    """y [B, T] and the cache for ``backward``."""
    Fc, Fs, Gc, Gs = mats
    L = n_layers(p)
    H = X0 @ p["P"] + p["bP"]
    cache = {"X0": X0, "H": [], "Z": [], "Xr": [], "Xi": []}
    for l in range(L):
        Xr = np.einsum("mt,btc->bmc", Fc, H)
        Xi = np.einsum("mt,btc->bmc", Fs, H)
        Wr, Wi = p[f"Wr{l}"], p[f"Wi{l}"]
        Yr = np.einsum("bmc,mcd->bmd", Xr, Wr) - np.einsum("bmc,mcd->bmd", Xi, Wi)
        Yi = np.einsum("bmc,mcd->bmd", Xr, Wi) + np.einsum("bmc,mcd->bmd", Xi, Wr)
        Z = np.einsum("tm,bmd->btd", Gc, Yr) + np.einsum("tm,bmd->btd", Gs, Yi) + H @ p[f"W{l}"] + p[f"b{l}"]
        cache["H"].append(H)
        cache["Z"].append(Z)
        cache["Xr"].append(Xr)
        cache["Xi"].append(Xi)
        H = gelu(Z) if l < L - 1 else Z
    U = H @ p["Q1"] + p["bQ1"]
    V = gelu(U)
    y = (V @ p["Q2"] + p["bQ2"])[..., 0]
    cache.update(HL=H, U=U, V=V)
    return y, cache


def backward(p: dict, cache: dict, dy: np.ndarray, mats) -> dict:
    """Gradients of sum(dy * y) with respect to every parameter."""
    Fc, Fs, Gc, Gs = mats
    L = n_layers(p)
    g = {}
    dy3 = dy[..., None]
    V, U, H = cache["V"], cache["U"], cache["HL"]
    g["Q2"] = np.einsum("bth,bto->ho", V, dy3)
    g["bQ2"] = dy3.sum(axis=(0, 1))
    dU = (dy3 @ p["Q2"].T) * gelu_grad(U)
    g["Q1"] = np.einsum("btc,bth->ch", H, dU)
    g["bQ1"] = dU.sum(axis=(0, 1))
    dH = dU @ p["Q1"].T
    for l in reversed(range(L)):
        Z, Hin, Xr, Xi = cache["Z"][l], cache["H"][l], cache["Xr"][l], cache["Xi"][l]
        dZ = dH if l == L - 1 else dH * gelu_grad(Z)
        g[f"b{l}"] = dZ.sum(axis=(0, 1))
        g[f"W{l}"] = np.einsum("btc,btd->cd", Hin, dZ)
        dHn = dZ @ p[f"W{l}"].T
        dYr = np.einsum("tm,btd->bmd", Gc, dZ)
        dYi = np.einsum("tm,btd->bmd", Gs, dZ)
        Wr, Wi = p[f"Wr{l}"], p[f"Wi{l}"]
        dXr = np.einsum("bmd,mcd->bmc", dYr, Wr) + np.einsum("bmd,mcd->bmc", dYi, Wi)
        dXi = -np.einsum("bmd,mcd->bmc", dYr, Wi) + np.einsum("bmd,mcd->bmc", dYi, Wr)
        g[f"Wr{l}"] = np.einsum("bmc,bmd->mcd", Xr, dYr) + np.einsum("bmc,bmd->mcd", Xi, dYi)
        g[f"Wi{l}"] = -np.einsum("bmc,bmd->mcd", Xi, dYr) + np.einsum("bmc,bmd->mcd", Xr, dYi)
        dHn = dHn + np.einsum("mt,bmc->btc", Fc, dXr) + np.einsum("mt,bmc->btc", Fs, dXi)
        dH = dHn
    g["P"] = np.einsum("bti,btc->ic", cache["X0"], dH)
    g["bP"] = dH.sum(axis=(0, 1))
    return g


def mse(y, target):
    """Mean squared error and its gradient with respect to y."""
    d = y - target
    return float(np.mean(d * d)), 2.0 * d / d.size


def f0_moments(f0, vmax):
    """Encoding of the surrogate's input f0 [n, nx, nv] on the grids
    v_j = -vmax[r] + j dv, dv = 2 vmax[r]/(nv - 1) (the quantities the
    vp_nn_f0_moments kernel returns): per run
    ((1/nx) sum_i rho_i, (1/nx) sum_i rho_i cos(2 pi i/nx),
    -(1/nx) sum_i rho_i sin(2 pi i/nx), M0, M2, M4), rho_i = sum_j f_ij dv,
    M_p = sum_j fbar_j v_j^p dv, fbar = (1/nx) sum_i f_ij.
    Pinned by closed forms (tests/test_oracle_fno.py): for
    f0 = (1 + A cos(2 pi i/nx)) 1/2 [M_s(v - u) + M_s(v + u)] on a wide box the
    results are (1, A/2, 0, 1, s^2 + u^2, u^4 + 6 u^2 s^2 + 3 s^4)."""
    f0 = np.asarray(f0, dtype=np.float64)
    n, nx, nv = f0.shape
    out = np.empty((n, 6))
    ang = 2.0 * math.pi * np.arange(nx) / nx
    for r in range(n):
        V = float(vmax[r])
        dv = 2.0 * V / (nv - 1)
        v = -V + np.arange(nv) * dv
        rho = f0[r].sum(axis=1) * dv
        fbar = f0[r].sum(axis=0) / nx
        out[r] = (rho.sum() / nx, (rho * np.cos(ang)).sum() / nx, -(rho * np.sin(ang)).sum() / nx,
                  (fbar * dv).sum(), (fbar * v * v * dv).sum(), (fbar * v ** 4 * dv).sum())
    return out


def decode_f0_moments(m):
    """(amplitude, drift u, width s) of a two-beam f0 from its encoding:
    A = 2 |rho_1| / rho_0, u^4 = (3 M2'^2 - M4') / 2 and s^2 = M2' - u^2 with
    M2' = M2 / M0, M4' = M4 / M0 (the moments of 1/2 [M_s(v - u) + M_s(v + u)]
    are M2' = s^2 + u^2, M4' = u^4 + 6 u^2 s^2 + 3 s^4)."""
    m = np.asarray(m, dtype=np.float64)
    A = 2.0 * np.hypot(m[:, 1], m[:, 2]) / m[:, 0]
    m2 = m[:, 4] / m[:, 3]
    m4 = m[:, 5] / m[:, 3]
    u = np.maximum(0.5 * (3.0 * m2 * m2 - m4), 0.0) ** 0.25
    s = np.sqrt(np.maximum(m2 - u * u, 0.0))
    return A, u, s
