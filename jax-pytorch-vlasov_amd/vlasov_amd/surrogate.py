"""Fourier-operator surrogate of the field-energy history, trained and run
on the library's own GPU kernels (BF16 matrix-core GEMM with FP32
accumulation, GELU, bias, MSE and Adam kernels; torch only holds memory).

No reference counterpart: the reference has no learned model.  The network
(a 1D Fourier neural operator along time) maps the initial-condition
parameters (k, amplitude, drift, sigma) of a 1D1V run on a fixed grid to
log W_E(t) at every step; it is trained only on the ensemble solver's
output, and the solver stays the fallback and the judge of its error:

    H0 = X0 P + bP                                  X0: [B, T, 6] features
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

CIN = 6  # features: t/T, k, log10(amplitude), drift, sigma, 1


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


def solve(params: np.ndarray, steps: int = 500, nx: int = 128, nv: int = 256, dt: float = 0.1,
          batch: int = 2048, device="cuda") -> torch.Tensor:
    """Field-energy histories [n, steps] (float64, device) of 1D1V runs on an
    nx x nv grid, f0 = (1 + A cos(k x)) 1/2 [M_s(v - u) + M_s(v + u)] with the
    Maxwellian M_s of width s, L = 2 pi / k, V_max = 6 s + u, run by the
    batched ensemble solver."""
    from .ensemble import Ensemble

    out = []
    for i0 in range(0, len(params), batch):
        chunk = params[i0:i0 + batch]
        specs = [{"nx": nx, "nv": nv, "L": 2.0 * math.pi / k, "V_max": 6.0 * sig + u,
                  "f0": torch.zeros((nx, nv), dtype=torch.float64, device=device)} for k, amp, u, sig in chunk]
        ens = Ensemble(specs, dt, physics_mode=True, device=device)
        for i, (k, amp, u, sig) in enumerate(chunk):
            g = ens.grids[i]
            fv = (torch.exp(-0.5 * ((g.v - u) / sig) ** 2) + torch.exp(-0.5 * ((g.v + u) / sig) ** 2)) / (
                2.0 * sig * math.sqrt(2.0 * math.pi))
            ens.state(i).copy_((1.0 + amp * torch.cos(k * g.x))[:, None] * fv[None, :])
        out.append(ens.run(steps).T.contiguous())
        del ens
    return torch.cat(out)


def features(params: np.ndarray, T: int, device="cuda") -> torch.Tensor:
    """Network input [n, T, 6]: t / T and the four parameters scaled to [-1, 1]
    over the box (amplitude in log10), and a constant 1."""
    x = np.asarray(params, dtype=np.float64).copy()
    x[:, 1] = np.log10(x[:, 1])
    z = 2.0 * (x - BOX[:, 0]) / (BOX[:, 1] - BOX[:, 0]) - 1.0
    n = len(x)
    X = np.empty((n, T, CIN), dtype=np.float32)
    X[:, :, 0] = (np.arange(T) / T)[None, :]
    X[:, :, 1:5] = z[:, None, :]
    X[:, :, 5] = 1.0
    return torch.from_numpy(X).to(device)


def rel_l2(W_hat: torch.Tensor, W: torch.Tensor) -> torch.Tensor:
    """Per-run relative L2 error of a history: ||W_hat - W|| / ||W||."""
    return (W_hat - W).norm(dim=1) / W.norm(dim=1)


class Surrogate:
    """FNO surrogate of log W_E(t), trained on the solver's histories."""

    def __init__(self, steps: int = 500, width: int = 64, modes: int = 32, hidden: int = 128, layers: int = 4,
                 seed: int = 0, device="cuda"):
        self.steps = steps
        self.device = torch.device(device)
        self.net = FNO(steps, width, modes, hidden, layers, seed=seed, device=self.device)
        self.mu = 0.0
        self.sd = 1.0

    FLOOR = 1e-6  # log target floor, relative to each run's maximum

    def _log_target(self, W: torch.Tensor) -> torch.Tensor:
        floor = W.max(dim=1, keepdim=True).values * self.FLOOR
        return torch.log(torch.maximum(W, floor))

    def fit(self, params: np.ndarray, W: torch.Tensor, iters: int = 16000, batch: int = 128, lr: float = 3e-3,
            seed: int = 0, log=None, uniform: float = 0.1) -> list:
        """Adam on the weighted mean squared error of the normalized log
        history; cosine learning-rate decay.  Per run the weight of time t
        is uniform / T + (1 - uniform) W_t^2 / sum W^2: the second part
        makes the loss the squared relative L2 error of the history (to
        first order), the metric the solver judges the surrogate by.
        Returns (iteration, loss) every 50 iterations; `log` (a callable)
        receives them as they come."""
        logW = self._log_target(W)
        self.mu = float(logW.mean())
        self.sd = float(logW.std())
        Y = ((logW - self.mu) / self.sd).float().contiguous()
        W2 = W.double() ** 2
        Wt = (uniform / self.steps + (1.0 - uniform) * W2 / W2.sum(dim=1, keepdim=True)) / batch
        Wt = Wt.float().contiguous()
        X = features(params, self.steps, self.device)
        n = len(params)
        gen = torch.Generator(device="cpu").manual_seed(seed)
        dy = torch.empty((batch, self.steps), dtype=torch.float32, device=self.device)
        losses = []
        for it in range(iters):
            idx = torch.randint(0, n, (batch,), generator=gen).to(self.device)
            xb = X.index_select(0, idx).contiguous()
            yb = Y.index_select(0, idx).contiguous()
            wb = Wt.index_select(0, idx).contiguous()
            out = self.net.forward(xb)
            sse = self.net.wmse(out, yb, wb, dy)
            self.net.backward(dy)
            lr_t = 1e-5 + 0.5 * (lr - 1e-5) * (1.0 + math.cos(math.pi * it / max(iters - 1, 1)))
            self.net.adam(lr_t)
            if it % 50 == 0 or it == iters - 1:
                losses.append((it, float(sse)))
                if log is not None:
                    log(losses[-1])
        return losses

    def predict(self, params: np.ndarray, batch: int = 512) -> torch.Tensor:
        """W_E histories [n, steps] (float64, device)."""
        X = features(params, self.steps, self.device)
        outs = []
        for i0 in range(0, len(params), batch):
            y = self.net.forward(X[i0:i0 + batch].contiguous())
            outs.append(torch.exp(y.double() * self.sd + self.mu))
        return torch.cat(outs)

    def query(self, params: np.ndarray, tol: float = 0.02, judge_fraction: float = 0.05, seed: int = 0,
              **solver_kw) -> tuple:
        """Serve W_E histories with the solver as fallback and judge.

        Parameters outside the training box go to the solver.  A random
        `judge_fraction` of the rest (at least one) is also solved and
        compared; if any judged run misses `tol` (relative L2 error of the
        history) the whole request falls back to the solver.  Returns
        (histories [n, steps], info)."""
        params = np.asarray(params, dtype=np.float64)
        n = len(params)
        inside = in_box(params)
        W = torch.empty((n, self.steps), dtype=torch.float64, device=self.device)
        info = {"n": n, "outside_box": int((~inside).sum()), "judged": 0, "judge_max_err": None,
                "fallback": False}
        idx_in = np.nonzero(inside)[0]
        idx_out = np.nonzero(~inside)[0]
        if len(idx_out):
            W[torch.from_numpy(idx_out).to(self.device)] = solve(params[idx_out], self.steps, **solver_kw)
        if len(idx_in):
            pred = self.predict(params[idx_in])
            rng = np.random.default_rng(seed)
            nj = max(1, int(round(judge_fraction * len(idx_in))))
            jsel = rng.choice(len(idx_in), size=min(nj, len(idx_in)), replace=False)
            Wj = solve(params[idx_in[jsel]], self.steps, **solver_kw)
            err = rel_l2(pred[torch.from_numpy(jsel).to(self.device)], Wj)
            info["judged"] = len(jsel)
            info["judge_max_err"] = float(err.max())
            if float(err.max()) > tol:
                info["fallback"] = True
                pred = solve(params[idx_in], self.steps, **solver_kw)
            else:
                pred[torch.from_numpy(jsel).to(self.device)] = Wj  # judged runs keep the solver's answer
            W[torch.from_numpy(idx_in).to(self.device)] = pred
        return W, info
