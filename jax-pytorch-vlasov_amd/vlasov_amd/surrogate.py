"""Fourier-operator surrogate of the field-energy history, trained and run
on the library's own GPU kernels (BF16 matrix-core GEMM with FP32
accumulation, GELU, bias, MSE and Adam kernels; torch only holds memory).

No reference counterpart: the reference has no learned model.  The network
(a 1D Fourier neural operator along time) maps the initial-condition
parameters (k, amplitude, drift, sigma) of a 1D1V run plus an encoding of
its initial distribution f0 on the 128 x 256 grid (density Fourier mode and
velocity moments, ``vp_nn_f0_moments``) to log W_E(t) at every step; it is
trained only on the ensemble solver's output, and the solver stays the
fallback and the judge of its error.  Two independently initialised
networks are trained; their disagreement, calibrated on held-out solver
runs, decides per query whether the network's answer is served:

    H0 = X0 P + bP                                  X0: [B, T, 9] features
    for l < L:  Z = G(R_l(F H)) + H W_l + b_l,      H = gelu(Z) (l < L-1)
    y = gelu(H Q1 + bQ1) Q2 + bQ2                   y: [B, T]

F keeps the first M real-DFT modes along T, R_l multiplies mode m by the
complex matrix Wr_l[m] + i Wi_l[m], G is the truncated inverse real DFT.
Every product is one strided, batched call of ``vp_nn_gemm``.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from ._lib import check, lib, stream_handle

# features: t/T; k, log10(amplitude), drift, sigma; from f0: log10 of the
# density mode amplitude, <v^2>, <v^4>/<v^2>^2; 1
CIN = 9


def _dft_mats(T: int, M: int, device):
    t = torch.arange(T, dtype=torch.float64)[None, :]
    m = torch.arange(M, dtype=torch.float64)[:, None]
    ang = 2.0 * math.pi * m * t / T
    a = torch.where(torch.arange(M) == 0, 1.0, 2.0).to(torch.float64)[None, :]
    mats = (torch.cos(ang), -torch.sin(ang), a * torch.cos(ang.T) / T, -a * torch.sin(ang.T) / T)
    return tuple(x.to(torch.float32).contiguous().to(device) for x in mats)


class _T:
    """(tensor, element offset) -> device address for the C calls."""

    __slots__ = ("t", "off")

    def __init__(self, t: torch.Tensor, off: int = 0):
        self.t = t
        self.off = off

    @property
    def ptr(self) -> int:
        return self.t.data_ptr() + 4 * self.off


def _as(x):
    return x if isinstance(x, _T) else _T(x)


class FNO:
    """Parameters (float32 master copies), gradients, Adam state and the
    forward/backward passes on the GPU kernels."""

    def __init__(self, T: int, width: int = 64, modes: int = 32, hidden: int = 128, layers: int = 4,
                 seed: int = 0, device="cuda", params: dict | None = None):
        self.T, self.C, self.M, self.Hd, self.L = T, width, modes, hidden, layers
        self.device = torch.device(device)
        if params is None:
            from_np = _init_params(CIN, width, modes, hidden, layers, seed)
        else:
            from_np = params
        self.p = {k: torch.as_tensor(np.asarray(v), dtype=torch.float32).contiguous().to(self.device)
                  for k, v in from_np.items()}
        self.g = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.step = 0
        self.Fc, self.Fs, self.Gc, self.Gs = _dft_mats(T, modes, self.device)
        self._bufs = {}

    # -- plumbing ---------------------------------------------------------
    def _buf(self, name, shape):
        b = self._bufs.get(name)
        n = int(np.prod(shape))
        if b is None or b.numel() < n:
            b = torch.empty(n, dtype=torch.float32, device=self.device)
            self._bufs[name] = b
        return b[:n].view(*shape)

    def _gemm(self, batch, M, N, K, alpha, A, sA, B, sB, beta, C, sC):
        A, B, C = _as(A), _as(B), _as(C)
        check(lib().vp_nn_gemm(batch, M, N, K, float(alpha), A.ptr, *sA, B.ptr, *sB, float(beta), C.ptr, *sC,
                               stream_handle(self.device)), "vp_nn_gemm")

    def _s(self):
        return stream_handle(self.device)

    # -- forward ----------------------------------------------------------
    def forward(self, X0: torch.Tensor) -> torch.Tensor:
        """X0 [B, T, CIN] float32 (device) -> y [B, T]."""
        B, T, C, M, Hd = X0.shape[0], self.T, self.C, self.M, self.Hd
        BT = B * T
        p = self.p
        H = self._buf("H0", (B, T, C))
        self._gemm(1, BT, C, CIN, 1.0, X0, (0, CIN, 1), p["P"], (0, C, 1), 0.0, H, (0, C, 1))
        check(lib().vp_nn_bias_add(H.data_ptr(), p["bP"].data_ptr(), BT, C, self._s()), "bias")
        self.cache = {"X0": X0, "B": B, "H": [], "Z": [], "Xh": []}
        for l in range(self.L):
            Xh = self._buf(f"Xh{l}", (B, M, 2 * C))
            # spectrum: Xr = Fc H, Xi = Fs H (per sample), stored [B][M][re | im]
            self._gemm(B, M, C, T, 1.0, self.Fc, (0, T, 1), H, (T * C, C, 1), 0.0, Xh, (M * 2 * C, 2 * C, 1))
            self._gemm(B, M, C, T, 1.0, self.Fs, (0, T, 1), H, (T * C, C, 1), 0.0, _T(Xh, C), (M * 2 * C, 2 * C, 1))
            Yh = self._buf("Yh", (B, M, 2 * C))
            Wr, Wi = p[f"Wr{l}"], p[f"Wi{l}"]
            sX = (2 * C, M * 2 * C, 1)  # batch over modes, rows over samples
            sW = (C * C, C, 1)
            self._gemm(M, B, C, C, 1.0, Xh, sX, Wr, sW, 0.0, Yh, sX)
            self._gemm(M, B, C, C, -1.0, _T(Xh, C), sX, Wi, sW, 1.0, Yh, sX)
            self._gemm(M, B, C, C, 1.0, Xh, sX, Wi, sW, 0.0, _T(Yh, C), sX)
            self._gemm(M, B, C, C, 1.0, _T(Xh, C), sX, Wr, sW, 1.0, _T(Yh, C), sX)
            Z = self._buf(f"Z{l}", (B, T, C))
            self._gemm(B, T, C, M, 1.0, self.Gc, (0, M, 1), Yh, (M * 2 * C, 2 * C, 1), 0.0, Z, (T * C, C, 1))
            self._gemm(B, T, C, M, 1.0, self.Gs, (0, M, 1), _T(Yh, C), (M * 2 * C, 2 * C, 1), 1.0, Z, (T * C, C, 1))
            self._gemm(1, BT, C, C, 1.0, H, (0, C, 1), p[f"W{l}"], (0, C, 1), 1.0, Z, (0, C, 1))
            check(lib().vp_nn_bias_add(Z.data_ptr(), p[f"b{l}"].data_ptr(), BT, C, self._s()), "bias")
            self.cache["H"].append(H)
            self.cache["Z"].append(Z)
            self.cache["Xh"].append(Xh)
            if l < self.L - 1:
                Hn = self._buf(f"H{l + 1}", (B, T, C))
                check(lib().vp_nn_gelu(Z.data_ptr(), Hn.data_ptr(), BT * C, self._s()), "gelu")
                H = Hn
            else:
                H = Z
        U = self._buf("U", (BT, Hd))
        self._gemm(1, BT, Hd, C, 1.0, H, (0, C, 1), p["Q1"], (0, Hd, 1), 0.0, U, (0, Hd, 1))
        check(lib().vp_nn_bias_add(U.data_ptr(), p["bQ1"].data_ptr(), BT, Hd, self._s()), "bias")
        V = self._buf("V", (BT, Hd))
        check(lib().vp_nn_gelu(U.data_ptr(), V.data_ptr(), BT * Hd, self._s()), "gelu")
        y = self._buf("y", (B, T))
        self._gemm(1, BT, 1, Hd, 1.0, V, (0, Hd, 1), p["Q2"], (0, 1, 1), 0.0, y, (0, 1, 1))
        check(lib().vp_nn_bias_add(y.data_ptr(), p["bQ2"].data_ptr(), BT, 1, self._s()), "bias")
        self.cache.update(HL=H, U=U, V=V)
        return y

    # -- backward ---------------------------------------------------------
    def backward(self, dy: torch.Tensor) -> None:
        """Gradients of sum(dy * y) into self.g (overwritten)."""
        c = self.cache
        B, T, C, M, Hd = c["B"], self.T, self.C, self.M, self.Hd
        BT = B * T
        p, g = self.p, self.g
        V, U, H = c["V"], c["U"], c["HL"]
        s = self._s()
        self._gemm(1, Hd, 1, BT, 1.0, V, (0, 1, Hd), dy, (0, 1, 1), 0.0, g["Q2"], (0, 1, 1))
        check(lib().vp_nn_bias_grad(dy.data_ptr(), g["bQ2"].data_ptr(), BT, 1, s), "bias_grad")
        dV = self._buf("dV", (BT, Hd))
        self._gemm(1, BT, Hd, 1, 1.0, dy, (0, 1, 1), p["Q2"], (0, 1, 1), 0.0, dV, (0, Hd, 1))
        dU = self._buf("dU", (BT, Hd))
        check(lib().vp_nn_gelu_bwd(U.data_ptr(), dV.data_ptr(), dU.data_ptr(), BT * Hd, s), "gelu_bwd")
        self._gemm(1, C, Hd, BT, 1.0, H, (0, 1, C), dU, (0, Hd, 1), 0.0, g["Q1"], (0, Hd, 1))
        check(lib().vp_nn_bias_grad(dU.data_ptr(), g["bQ1"].data_ptr(), BT, Hd, s), "bias_grad")
        dH = self._buf("dH_a", (B, T, C))
        self._gemm(1, BT, C, Hd, 1.0, dU, (0, Hd, 1), p["Q1"], (0, 1, Hd), 0.0, dH, (0, C, 1))
        sX = (2 * C, M * 2 * C, 1)
        for l in reversed(range(self.L)):
            Z, Hin, Xh = c["Z"][l], c["H"][l], c["Xh"][l]
            if l == self.L - 1:
                dZ = dH
            else:
                dZ = self._buf("dZ", (B, T, C))
                check(lib().vp_nn_gelu_bwd(Z.data_ptr(), dH.data_ptr(), dZ.data_ptr(), BT * C, s), "gelu_bwd")
            check(lib().vp_nn_bias_grad(dZ.data_ptr(), g[f"b{l}"].data_ptr(), BT, C, s), "bias_grad")
            self._gemm(1, C, C, BT, 1.0, Hin, (0, 1, C), dZ, (0, C, 1), 0.0, g[f"W{l}"], (0, C, 1))
            dHn = self._buf("dH_b" if dH.data_ptr() == self._bufs["dH_a"].data_ptr() else "dH_a", (B, T, C))
            self._gemm(1, BT, C, C, 1.0, dZ, (0, C, 1), p[f"W{l}"], (0, 1, C), 0.0, dHn, (0, C, 1))
            dYh = self._buf("dYh", (B, M, 2 * C))
            self._gemm(B, M, C, T, 1.0, self.Gc, (0, 1, M), dZ, (T * C, C, 1), 0.0, dYh, (M * 2 * C, 2 * C, 1))
            self._gemm(B, M, C, T, 1.0, self.Gs, (0, 1, M), dZ, (T * C, C, 1), 0.0, _T(dYh, C), (M * 2 * C, 2 * C, 1))
            Wr, Wi = p[f"Wr{l}"], p[f"Wi{l}"]
            sWt = (C * C, 1, C)  # W[m] transposed
            dXh = self._buf("dXh", (B, M, 2 * C))
            self._gemm(M, B, C, C, 1.0, dYh, sX, Wr, sWt, 0.0, dXh, sX)
            self._gemm(M, B, C, C, 1.0, _T(dYh, C), sX, Wi, sWt, 1.0, dXh, sX)
            self._gemm(M, B, C, C, -1.0, dYh, sX, Wi, sWt, 0.0, _T(dXh, C), sX)
            self._gemm(M, B, C, C, 1.0, _T(dYh, C), sX, Wr, sWt, 1.0, _T(dXh, C), sX)
            sXt = (2 * C, 1, M * 2 * C)  # per mode: rows over channels, k over samples
            sW = (C * C, C, 1)
            self._gemm(M, C, C, B, 1.0, Xh, sXt, dYh, sX, 0.0, g[f"Wr{l}"], sW)
            self._gemm(M, C, C, B, 1.0, _T(Xh, C), sXt, _T(dYh, C), sX, 1.0, g[f"Wr{l}"], sW)
            self._gemm(M, C, C, B, -1.0, _T(Xh, C), sXt, dYh, sX, 0.0, g[f"Wi{l}"], sW)
            self._gemm(M, C, C, B, 1.0, Xh, sXt, _T(dYh, C), sX, 1.0, g[f"Wi{l}"], sW)
            self._gemm(B, T, C, M, 1.0, self.Fc, (0, 1, T), dXh, (M * 2 * C, 2 * C, 1), 1.0, dHn, (T * C, C, 1))
            self._gemm(B, T, C, M, 1.0, self.Fs, (0, 1, T), _T(dXh, C), (M * 2 * C, 2 * C, 1), 1.0, dHn,
                       (T * C, C, 1))
            dH = dHn
        self._gemm(1, CIN, C, BT, 1.0, c["X0"], (0, 1, CIN), dH, (0, C, 1), 0.0, g["P"], (0, C, 1))
        check(lib().vp_nn_bias_grad(dH.data_ptr(), g["bP"].data_ptr(), BT, C, s), "bias_grad")

    def adam(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> None:
        self.step += 1
        for k, w in self.p.items():
            check(lib().vp_nn_adam(w.data_ptr(), self.g[k].data_ptr(), self.m[k].data_ptr(), self.v[k].data_ptr(),
                                   w.numel(), lr, b1, b2, eps, self.step, self._s()), "adam")

    def wmse(self, y: torch.Tensor, target: torch.Tensor, w: torch.Tensor, grad: torch.Tensor | None) -> torch.Tensor:
        """Weighted sum of squared errors (device scalar); grad <- 2 w (y - target)."""
        n = y.numel()
        part = self._buf("mse_part", ((n + 255) // 256,))
        check(lib().vp_nn_wmse(y.data_ptr(), target.data_ptr(), w.data_ptr(),
                               None if grad is None else grad.data_ptr(), part.data_ptr(), n, self._s()), "wmse")
        return part.sum()

    def mse(self, y: torch.Tensor, target: torch.Tensor, grad: torch.Tensor | None) -> torch.Tensor:
        """Sum of squared errors (device scalar); grad <- 2 (y - target) / n."""
        n = y.numel()
        part = self._buf("mse_part", ((n + 255) // 256,))
        check(lib().vp_nn_mse(y.data_ptr(), target.data_ptr(), None if grad is None else grad.data_ptr(),
                              part.data_ptr(), n, self._s()), "mse")
        return part.sum()


def _init_params(cin, width, modes, hidden, layers, seed):
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


# ---------------------------------------------------------------------------
# Data, training and serving
# ---------------------------------------------------------------------------
# Parameter box of the training distribution (the batched ensemble's
# generator plus a thermal width): k ~ U[0.2, 0.6], amplitude ~ log-U[1e-3,
# 0.5], beam drift ~ U[0, 3], sigma ~ U[0.8, 1.2].
BOX = np.array([[0.2, 0.6], [math.log10(1e-3), math.log10(0.5)], [0.0, 3.0], [0.8, 1.2]])


def sample_params(n: int, seed: int = 0) -> np.ndarray:
    """[n, 4] rows (k, amplitude, drift, sigma), seeded numpy PCG64."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(size=(n, 4))
    lo, hi = BOX[:, 0], BOX[:, 1]
    x = lo + u * (hi - lo)
    x[:, 1] = 10.0 ** x[:, 1]
    return x


def in_box(params: np.ndarray) -> np.ndarray:
    x = np.asarray(params, dtype=np.float64).copy()
    x[:, 1] = np.log10(np.maximum(x[:, 1], 1e-300))
    return np.all((x >= BOX[:, 0] - 1e-12) & (x <= BOX[:, 1] + 1e-12), axis=1)


def f0_batch(params: np.ndarray, nx: int = 128, nv: int = 256, device="cuda") -> torch.Tensor:
    """Initial distributions [n, nx, nv] (float64, device) of the runs:
    f0 = (1 + A cos(2 pi i / nx)) 1/2 [M_s(v - u) + M_s(v + u)] at x_i = i L / nx
    (L = 2 pi / k), v_j = -V + j 2V / (nv - 1), V = 6 s + u, M_s the
    Maxwellian of width s."""
    p = torch.as_tensor(np.asarray(params, dtype=np.float64), device=device)
    A, u, sg = p[:, 1:2], p[:, 2:3], p[:, 3:4]
    V = 6.0 * sg + u
    v = -V + torch.arange(nv, dtype=torch.float64, device=device)[None, :] * (2.0 * V / (nv - 1))
    fv = (torch.exp(-0.5 * ((v - u) / sg) ** 2) + torch.exp(-0.5 * ((v + u) / sg) ** 2)) / (
        2.0 * sg * math.sqrt(2.0 * math.pi))
    ang = torch.cos(2.0 * math.pi * torch.arange(nx, dtype=torch.float64, device=device) / nx)[None, :]
    return ((1.0 + A * ang)[:, :, None] * fv[:, None, :]).contiguous()


def vmax_of(params: np.ndarray) -> np.ndarray:
    """V_max = 6 s + u of each run."""
    p = np.asarray(params, dtype=np.float64)
    return 6.0 * p[:, 3] + p[:, 2]


def encode_f0(f0: torch.Tensor, vmax) -> torch.Tensor:
    """vp_nn_f0_moments on f0 [n, nx, nv]: [n, 6] (float64, device)."""
    n, nx, nv = f0.shape
    vm = torch.as_tensor(np.asarray(vmax, dtype=np.float64), device=f0.device).contiguous()
    f0 = f0.contiguous()
    out = torch.empty((n, 6), dtype=torch.float64, device=f0.device)
    check(lib().vp_nn_f0_moments(f0.data_ptr(), n, nx, nv, vm.data_ptr(), out.data_ptr(),
                                 stream_handle(f0.device)), "vp_nn_f0_moments")
    return out


def solve(params: np.ndarray, steps: int = 500, nx: int = 128, nv: int = 256, dt: float = 0.1,
          batch: int = 2048, device="cuda", return_moments: bool = False):
    """Field-energy histories [n, steps] (float64, device) of 1D1V runs on an
    nx x nv grid from f0_batch(params) (L = 2 pi / k, V_max = 6 s + u), run by
    the batched ensemble solver; with return_moments, also the encodings of
    their f0 (encode_f0)."""
    from .ensemble import Ensemble

    params = np.asarray(params, dtype=np.float64)
    out, mom = [], []
    for i0 in range(0, len(params), batch):
        chunk = params[i0:i0 + batch]
        f0 = f0_batch(chunk, nx, nv, device)
        if return_moments:
            mom.append(encode_f0(f0, vmax_of(chunk)))
        specs = [{"nx": nx, "nv": nv, "L": 2.0 * math.pi / k, "V_max": 6.0 * sig + u, "f0": f0[i]}
                 for i, (k, amp, u, sig) in enumerate(chunk)]
        ens = Ensemble(specs, dt, physics_mode=True, device=device)
        out.append(ens.run(steps).T.contiguous())
        del ens, f0
    W = torch.cat(out)
    return (W, torch.cat(mom)) if return_moments else W


def f0_features(moments: torch.Tensor) -> torch.Tensor:
    """[n, 3] from the f0 encoding: log10 of the density mode amplitude
    2 |rho_1| / rho_0, <v^2> = M2 / M0 and the kurtosis <v^4> / <v^2>^2
    (3 for one Maxwellian, lower for two separated beams)."""
    m = moments.double()
    A = 2.0 * torch.hypot(m[:, 1], m[:, 2]) / m[:, 0]
    m2 = m[:, 4] / m[:, 3]
    m4 = m[:, 5] / m[:, 3]
    return torch.stack([torch.log10(torch.clamp(A, min=1e-300)), m2, m4 / (m2 * m2)], dim=1)


# ranges of the f0 features over the parameter box (log10 A; <v^2> = s^2 + u^2; kurtosis)
F0_BOX = np.array([[math.log10(1e-3), math.log10(0.5)], [0.64, 10.44], [1.0, 3.0]])


def features(params: np.ndarray, T: int, moments: torch.Tensor | None = None, device="cuda") -> torch.Tensor:
    """Network input [n, T, 9]: t / T; the four parameters scaled to [-1, 1]
    over the box (amplitude in log10); the three f0 features (f0_features of
    `moments`, the f0 encodings; computed from f0_batch(params) when None)
    scaled over their ranges; a constant 1."""
    x = np.asarray(params, dtype=np.float64).copy()
    x[:, 1] = np.log10(x[:, 1])
    z = 2.0 * (x - BOX[:, 0]) / (BOX[:, 1] - BOX[:, 0]) - 1.0
    n = len(x)
    if moments is None:
        moments = torch.cat([encode_f0(f0_batch(params[i:i + 2048], device=device), vmax_of(params[i:i + 2048]))
                             for i in range(0, n, 2048)])
    ff = f0_features(moments).cpu().numpy()
    zf = 2.0 * (ff - F0_BOX[:, 0]) / (F0_BOX[:, 1] - F0_BOX[:, 0]) - 1.0
    X = np.empty((n, T, CIN), dtype=np.float32)
    X[:, :, 0] = (np.arange(T) / T)[None, :]
    X[:, :, 1:5] = z[:, None, :]
    X[:, :, 5:8] = zf[:, None, :]
    X[:, :, 8] = 1.0
    return torch.from_numpy(X).to(device)


def rel_l2(W_hat: torch.Tensor, W: torch.Tensor) -> torch.Tensor:
    """Per-run relative L2 error of a history: ||W_hat - W|| / ||W||."""
    return (W_hat - W).norm(dim=1) / W.norm(dim=1)


class Surrogate:
    """Two-member FNO surrogate of log W_E(t), trained on the solver's
    histories, with per-query routing by the members' disagreement."""

    def __init__(self, steps: int = 500, width: int = 64, modes: int = 32, hidden: int = 128, layers: int = 4,
                 members: int = 2, seed: int = 0, device="cuda"):
        self.steps = steps
        self.device = torch.device(device)
        self.nets = [FNO(steps, width, modes, hidden, layers, seed=seed + 7919 * i, device=self.device)
                     for i in range(members)]
        self.net = self.nets[0]
        self.mu = 0.0
        self.sd = 1.0
        self.cal_spread = None  # calibration runs: disagreement, sorted
        self.cal_err = None     # and the matching judged errors

    FLOOR = 1e-6  # log target floor, relative to each run's maximum

    def _log_target(self, W: torch.Tensor) -> torch.Tensor:
        floor = W.max(dim=1, keepdim=True).values * self.FLOOR
        return torch.log(torch.maximum(W, floor))

    def fit(self, params: np.ndarray, W: torch.Tensor, moments: torch.Tensor | None = None, iters: int = 16000,
            batch: int = 128, lr: float = 3e-3, seed: int = 0, log=None, uniform: float = 0.1) -> list:
        """Adam on the weighted mean squared error of the normalized log
        history, cosine learning-rate decay, `iters` iterations per member
        (each member its own initialisation and minibatch order).  Per run
        the weight of time t is uniform / T + (1 - uniform) W_t^2 / sum W^2:
        the second part makes the loss the squared relative L2 error of the
        history (to first order), the metric the solver judges the
        surrogate by.  `moments`: the f0 encodings of the training runs
        (computed when None).  Returns (member, iteration, loss) every 50
        iterations; `log` (a callable) receives them as they come."""
        logW = self._log_target(W)
        self.mu = float(logW.mean())
        self.sd = float(logW.std())
        Y = ((logW - self.mu) / self.sd).float().contiguous()
        W2 = W.double() ** 2
        Wt = (uniform / self.steps + (1.0 - uniform) * W2 / W2.sum(dim=1, keepdim=True)) / batch
        Wt = Wt.float().contiguous()
        X = features(params, self.steps, moments, self.device)
        n = len(params)
        dy = torch.empty((batch, self.steps), dtype=torch.float32, device=self.device)
        losses = []
        for mi, net in enumerate(self.nets):
            gen = torch.Generator(device="cpu").manual_seed(seed + 104729 * mi)
            for it in range(iters):
                idx = torch.randint(0, n, (batch,), generator=gen).to(self.device)
                xb = X.index_select(0, idx).contiguous()
                yb = Y.index_select(0, idx).contiguous()
                wb = Wt.index_select(0, idx).contiguous()
                out = net.forward(xb)
                sse = net.wmse(out, yb, wb, dy)
                net.backward(dy)
                lr_t = 1e-5 + 0.5 * (lr - 1e-5) * (1.0 + math.cos(math.pi * it / max(iters - 1, 1)))
                net.adam(lr_t)
                if it % 50 == 0 or it == iters - 1:
                    losses.append((mi, it, float(sse)))
                    if log is not None:
                        log(losses[-1])
        return losses

    def predict_members(self, params: np.ndarray, moments: torch.Tensor | None = None,
                        batch: int = 512) -> torch.Tensor:
        """log W_E histories of every member [members, n, steps] (float64)."""
        X = features(params, self.steps, moments, self.device)
        outs = []
        for net in self.nets:
            rows = []
            for i0 in range(0, len(params), batch):
                y = net.forward(X[i0:i0 + batch].contiguous())
                rows.append(y.double() * self.sd + self.mu)
            outs.append(torch.cat(rows))
        return torch.stack(outs)

    def predict(self, params: np.ndarray, moments: torch.Tensor | None = None, batch: int = 512,
                with_spread: bool = False):
        """W_E histories [n, steps] (float64, device): the exponential of the
        members' mean log history.  with_spread: also each run's
        disagreement, the largest relative L2 distance of a member's
        history from the served one."""
        logs = self.predict_members(params, moments, batch)
        W = torch.exp(logs.mean(dim=0))
        if not with_spread:
            return W
        spread = torch.stack([rel_l2(torch.exp(l), W) for l in logs]).max(dim=0).values
        return W, spread

    def calibrate(self, params: np.ndarray, W: torch.Tensor, moments: torch.Tensor | None = None) -> None:
        """Record disagreement and true error on held-out solver runs (not
        used for training): the routing threshold of `query` comes from them."""
        Wp, spread = self.predict(params, moments, with_spread=True)
        err = rel_l2(Wp, W)
        order = torch.argsort(spread)
        self.cal_spread = spread[order].cpu().numpy()
        self.cal_err = err[order].cpu().numpy()

    def threshold(self, tol: float, quantile: float = 0.95, min_runs: int = 20) -> float:
        """Largest disagreement s such that, over the calibration runs with
        disagreement <= s (at least min_runs of them), the fraction whose
        error exceeds tol is at most 1 - quantile (-inf without calibration
        or when no such s exists)."""
        if self.cal_spread is None:
            return -math.inf
        miss = np.cumsum(self.cal_err > tol) / np.arange(1, len(self.cal_err) + 1)
        ok = np.nonzero((miss <= 1.0 - quantile) & (np.arange(1, len(miss) + 1) >= min_runs))[0]
        return float(self.cal_spread[ok[-1]]) if len(ok) else -math.inf

    def query(self, params: np.ndarray, tol: float = 0.02, quantile: float = 0.95, judge_fraction: float = 0.05,
              seed: int = 0, **solver_kw) -> tuple:
        """Serve W_E histories, each query routed on its own.

        A query is answered by the network when its parameters lie in the
        training box and the members' disagreement is at most
        threshold(tol, quantile): on the held-out calibration runs, a
        fraction `quantile` of the answers below that disagreement were
        within `tol` (relative L2 error of the history).  Every other query
        is answered by the solver.  A random `judge_fraction` of the served
        queries (at least one) is also solved; those keep the solver's
        answer.  If the judged misses exceed what the calibration allows
        (more than (1 - quantile) n_judged + 2 sqrt((1 - quantile) n_judged)
        + 1), the request lies outside the calibrated distribution and every
        query goes to the solver.  Returns (histories [n, steps], info)."""
        params = np.asarray(params, dtype=np.float64)
        n = len(params)
        inside = in_box(params)
        W = torch.empty((n, self.steps), dtype=torch.float64, device=self.device)
        thr = self.threshold(tol, quantile)
        info = {"n": n, "outside_box": int((~inside).sum()), "threshold": thr, "served": 0, "solved": 0,
                "judged": 0, "judged_misses": 0, "judge_max_err": None, "fallback": False}
        to_solve = ~inside
        judged = np.zeros(n, dtype=bool)
        idx_in = np.nonzero(inside)[0]
        if len(idx_in):
            pred, spread = self.predict(params[idx_in], with_spread=True)
            sp = spread.cpu().numpy()
            ok = sp <= thr
            serve = idx_in[ok]
            to_solve[idx_in[~ok]] = True
            W[torch.from_numpy(idx_in).to(self.device)] = pred
            if len(serve):
                rng = np.random.default_rng(seed)
                nj = min(len(serve), max(1, int(round(judge_fraction * len(serve)))))
                jidx = serve[np.sort(rng.choice(len(serve), size=nj, replace=False))]
                Wj = solve(params[jidx], self.steps, **solver_kw)
                err = rel_l2(W[torch.from_numpy(jidx).to(self.device)], Wj).cpu().numpy()
                W[torch.from_numpy(jidx).to(self.device)] = Wj  # judged queries keep the solver's answer
                judged[jidx] = True
                misses = int((err > tol).sum())
                info.update(judged=nj, judged_misses=misses, judge_max_err=float(err.max()))
                expect = (1.0 - quantile) * nj
                if misses > expect + 2.0 * math.sqrt(expect) + 1.0:
                    info["fallback"] = True
                    to_solve[serve] = True
                    to_solve[jidx] = False
        sidx = np.nonzero(to_solve)[0]
        if len(sidx):
            W[torch.from_numpy(sidx).to(self.device)] = solve(params[sidx], self.steps, **solver_kw)
        served = ~(to_solve | judged)
        info.update(served=int(served.sum()), solved=int(n - served.sum()), served_mask=served)
        return W, info
