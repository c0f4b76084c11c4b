"""Shift ("reflected stencil") forms of the sweeps: the reference's solver 1
(``advect_1d_matrix_diag``) and solver 2 (``precompute_x_drift``,
``advect_x_static``, ``build_v_tap_weights``, ``advect_v_static``,
``check_v_window``), shared by the 1D1V and 2D2V mirror modules.

The sweeps run in the library's kernels (``vp_advect_reflected_f64``,
``vp_advect_static_f64``): d = (v dt)/dx, s = floor(d), alpha = d - s,
out[i] = sum_k w_k(alpha) f[i - s - o_k].  The small host-side pieces
(static blocks, tap-weight tables) follow the reference's own NumPy/torch
arithmetic (vlasov1D1V.py:357-512, vlasov2D2V.py:670-941).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, require_gpu, stream_handle

OFFSETS = (-2, -1, 0, 1, 2, 3)


def static_stencil_offsets(stencil_offsets):
    """Stencil offsets as a tuple of Python ints (vlasov1D1V.py:339-355).
    A device tensor is read once (reading it waits for the device): the
    tuple is kept on the tensor object with its version counter, so the
    per-step checks of the drivers do not synchronise."""
    if isinstance(stencil_offsets, torch.Tensor):
        cached = getattr(stencil_offsets, "_vp_static_offsets", None)
        if cached is not None and cached[0] == stencil_offsets._version:
            return cached[1]
        t = tuple(int(o) for o in stencil_offsets.tolist())
        try:
            stencil_offsets._vp_static_offsets = (stencil_offsets._version, t)
        except (AttributeError, RuntimeError):
            pass
        return t
    return tuple(int(o) for o in stencil_offsets)


def _check_offsets(stencil_offsets):
    if stencil_offsets is not None and static_stencil_offsets(stencil_offsets) != OFFSETS:
        raise ValueError(f"the kernels implement the stencil offsets {OFFSETS}, got {stencil_offsets}")


def lagrange5_weights(alpha):
    # same torch expressions as the reference (power form)
    from .vlasov1D1V import lagrange5_weights as lw
    return lw(alpha)


def sweep_reflected(f4: torch.Tensor, axis: int, vel: torch.Tensor, vel_strides, dt: float, dx: float,
                    periodic: bool, window: int = -1) -> torch.Tensor:
    """Reflected sweep of a 4D view along `axis`; velocity of element i at
    vel[sum_d i_d * vel_strides[d]]."""
    dev = require_gpu(f4, vel)
    if f4.dtype != torch.float64:
        raise TypeError("the shift forms run float64")
    fc = f4.contiguous()
    out = torch.empty_like(fc)
    v = vel.to(torch.float64)
    check(lib().vp_advect_reflected_f64(fc.data_ptr(), out.data_ptr(), _lib.i64x4(fc.shape), int(axis), v.data_ptr(),
                                        _lib.i64x4(vel_strides), float(dt), float(dx), int(bool(periodic)),
                                        int(window), stream_handle(dev)), "vp_advect_reflected_f64")
    return out


def precompute_x_drift(v, dt, dx, dtype):
    """(blocks, W) of vlasov1D1V.py:357-387 / vlasov2D2V.py:670-714."""
    np_dtype = np.dtype("float64" if dtype == torch.float64 else "float32")
    v_np = v.detach().cpu().numpy().astype(np_dtype)
    d = (v_np * np_dtype.type(dt)) / np_dtype.type(dx)
    s = np.floor(d)
    alpha = d - s
    s = s.astype(np.int64)
    edges = np.flatnonzero(np.diff(s)) + 1
    starts = np.concatenate(([0], edges))
    stops = np.concatenate((edges, [len(s)]))
    blocks = [(int(a), int(b), int(s[a])) for a, b in zip(starts, stops)]
    W = lagrange5_weights(torch.as_tensor(alpha, dtype=dtype, device=v.device))
    return blocks, W


_SHIFT_CACHE: dict = {}


def block_shifts(blocks, device) -> torch.Tensor:
    """Per-line shifts of `blocks` as a device tensor.  The drivers pass the
    same blocks every step, so the uploaded tensor is kept (a few recent
    block lists per device): no host-to-device copy, which would stall the
    stream, on the step path."""
    key = (tuple(tuple(b) for b in blocks), str(torch.device(device)))
    t = _SHIFT_CACHE.get(key)
    if t is None:
        if len(_SHIFT_CACHE) >= 16:
            _SHIFT_CACHE.pop(next(iter(_SHIFT_CACHE)))
        t = torch.tensor([s for a, b, s in blocks for _ in range(b - a)], dtype=torch.int64, device=device)
        _SHIFT_CACHE[key] = t
    return t


def static_sweep(f4: torch.Tensor, axis: int, shifts: torch.Tensor, W: torch.Tensor, param_strides) -> torch.Tensor:
    """Periodic sweep with given per-line shift and weights (W: [6, P])."""
    dev = require_gpu(f4, shifts, W)
    fc = f4.contiguous()
    out = torch.empty_like(fc)
    Wc = W.to(torch.float64).contiguous()
    sh = shifts.to(torch.int64).contiguous()
    check(lib().vp_advect_static_f64(fc.data_ptr(), out.data_ptr(), _lib.i64x4(fc.shape), int(axis), sh.data_ptr(),
                                     Wc.data_ptr(), Wc.shape[1], _lib.i64x4(param_strides), stream_handle(dev)),
          "vp_advect_static_f64")
    return out


def roll_weights(n: int, device) -> torch.Tensor:
    """Weights of a pure integer shift (only the centre tap): [6, n]."""
    W = torch.zeros((6, n), dtype=torch.float64, device=device)
    W[2] = 1.0
    return W


def build_v_tap_weights(E, dt, dv, S, stencil_offsets):
    """[2S+5, *E.shape] tap weights on m = -S-2 .. S+2 (vlasov1D1V.py:450-477)."""
    d = (E * dt) / dv
    s = torch.floor(d)
    alpha = d - s
    s = s.to(torch.long)
    W6 = lagrange5_weights(alpha)
    rows = []
    offs = static_stencil_offsets(stencil_offsets)
    for m in range(-(S + 2), S + 3):
        acc = torch.zeros_like(alpha)
        for k, o in enumerate(offs):
            acc = acc + (s == m - o).to(W6.dtype) * W6[k]
        rows.append(acc)
    return torch.stack(rows, dim=0)


def check_v_window(E, dt, dv, S):
    """max |E dt / dv| (vlasov1D1V.py:503-512); the static window is exact below S."""
    return float((E * dt / dv).abs().max())


def build_diag_weights(stencil_offsets, shifts, alpha, n_points):
    """[6, B, N] banded weights of the solver-1 Dirichlet sweep
    (build_v_diag_weights, vlasov1D1V.py:279-296)."""
    W = lagrange5_weights(alpha)
    v = torch.arange(n_points, device=alpha.device)
    offs = stencil_offsets if isinstance(stencil_offsets, torch.Tensor) else torch.tensor(
        stencil_offsets, device=alpha.device)
    total = offs.to(shifts.dtype).unsqueeze(1) + shifts.unsqueeze(0)
    src = v.view(1, 1, -1) - total.unsqueeze(-1)
    valid = (src >= 0) & (src < n_points)
    return valid.to(W.dtype) * W.unsqueeze(-1)
