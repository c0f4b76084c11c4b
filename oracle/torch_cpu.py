"""PyTorch-CPU restatement of the reference solver step (the timed CPU baseline).

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module: the
``cpu_baseline`` leg of ``bench.py`` times it and ``tests/`` check it.

It follows the reference PyTorch solver's default path (solver 0) in its op
order, eager, in float64, on the CPU:

* ``advect_1d`` (vlasov2D2V.py:356-413): permute the swept axis last,
  reshape to [B, N], ``advect_1d_vectorized`` (vlasov2D2V.py:287-353:
  departure index, ``floor``, fraction, [B, N, 6] int64 stencil indices,
  ``remainder`` or clamp + mask, a flat gather, the power-form weight sum
  of ``lagrange5_interp_sum`` vlasov2D2V.py:238-260), permute back;
* ``compute_rho`` / ``compute_poisson_real`` (vlasov2D2V.py:416-517) with
  ``torch.fft.rfft2`` / ``irfft2``;
* ``vlasov_step_impl`` (vlasov2D2V.py:943-998) and the 1D1V ``vlasov_step``
  (vlasov1D1V.py:574-622, with ``solve_poisson_real`` :538-572).

One difference, for memory only: ``advect_1d_vectorized`` runs on blocks of
rows (at most ``ROW_BLOCK_ELEMS`` stencil entries each) instead of the whole
[B, N, 6] at once, which at 128^4 would hold several 13 GB index tensors.
Every element is computed by the same operations.  The reference runs these
functions under ``torch.compile``; eager mode is timed here.

Pinned by tests/test_oracle_torch_cpu.py: every sweep is array-equal to
``oracle.vlasov_oracle.advect_axis`` and the steps agree with
``vlasov_step_1d`` / ``vlasov_step_2d`` to 1e-13 (the sums and FFTs round
differently).
"""
from __future__ import annotations

import os

import torch

ROW_BLOCK_ELEMS = 1 << 25  # stencil entries per block of rows


def use_all_cores() -> int:
    """torch intra-op threads = the CPUs this process may run on, capped by
    OMP_NUM_THREADS when that is set (the CPU share of a shared machine)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    torch.set_num_threads(n)
    return n


def _interp_sum(alpha, vals):
    a = alpha
    a2 = a * a
    a3 = a2 * a
    a4 = a3 * a
    a5 = a4 * a
    w0 = (a / 20.0) - (a2 / 24.0) - (a3 / 24.0) + (a4 / 24.0) - (a5 / 120.0)
    w1 = -(a / 2.0) + (2.0 * a2 / 3.0) - (a3 / 24.0) - (a4 / 6.0) + (a5 / 24.0)
    w2 = 1.0 - (a / 3.0) - (5.0 * a2 / 4.0) + (5.0 * a3 / 12.0) + (a4 / 4.0) - (a5 / 12.0)
    w3 = a + (2.0 * a2 / 3.0) - (7.0 * a3 / 12.0) - (a4 / 6.0) + (a5 / 12.0)
    w4 = -(a / 4.0) - (a2 / 24.0) + (7.0 * a3 / 24.0) + (a4 / 24.0) - (a5 / 24.0)
    w5 = (a / 30.0) - (a3 / 24.0) + (a5 / 120.0)
    return (w0 * vals[..., 0] + w1 * vals[..., 1] + w2 * vals[..., 2] + w3 * vals[..., 3]
            + w4 * vals[..., 4] + w5 * vals[..., 5])


def advect_rows(f_batch, velocities, dt, dx, periodic=True):
    """advect_1d_vectorized on f_batch [B, N] with one velocity per row (or
    per point, [B, N])."""
    B, N = f_batch.shape
    out = torch.empty_like(f_batch)
    offsets = torch.arange(-2, 4, dtype=torch.int64)
    points = torch.arange(N, dtype=f_batch.dtype)
    rows = max(1, ROW_BLOCK_ELEMS // (6 * N))
    for b0 in range(0, B, rows):
        b1 = min(B, b0 + rows)
        fb = f_batch[b0:b1]
        nb = b1 - b0
        v = velocities[b0:b1]
        v = v.unsqueeze(1) if v.ndim == 1 else v
        depart = points.unsqueeze(0) - (v * dt) / dx
        idx_left = torch.floor(depart)
        alpha = depart - idx_left
        stencil = idx_left.to(torch.long).unsqueeze(-1) + offsets  # [nb, N, 6]
        base = torch.arange(nb, dtype=torch.int64).view(nb, 1, 1) * N
        flat = fb.reshape(-1)
        if periodic:
            vals = flat[(base + torch.remainder(stencil, N)).reshape(-1)].reshape(nb, N, 6)
        else:
            valid = (stencil >= 0) & (stencil < N)
            vals = flat[(base + torch.clamp(stencil, 0, N - 1)).reshape(-1)].reshape(nb, N, 6)
            vals = vals * valid.to(fb.dtype)
        out[b0:b1] = _interp_sum(alpha, vals)
    return out


def advect_axis(f, velocities, axis, dt, dx, periodic=True):
    """advect_1d: sweep along `axis`; velocities broadcast to the batch shape."""
    perm = list(range(f.ndim))
    perm.append(perm.pop(axis))
    fp = f.permute(*perm)
    batch_shape = fp.shape[:-1]
    flat = fp.reshape(-1, fp.shape[-1])
    vel = velocities.expand(batch_shape).reshape(-1)
    out = advect_rows(flat, vel, dt, dx, periodic).reshape(fp.shape)
    inv = [0] * f.ndim
    for i, p in enumerate(perm):
        inv[p] = i
    return out.permute(*inv)


def compute_poisson_real(rho, kx, ky, inv_ksq):
    """compute_poisson_real: (Ex, Ey, delta_rho, phi)."""
    Nx, Ny = rho.shape
    delta_rho = rho - 1
    rho_k = torch.view_as_real(torch.fft.rfft2(delta_rho))
    phi_k = rho_k * inv_ksq.unsqueeze(-1)
    Ex_k = torch.stack([phi_k[..., 1] * kx, -phi_k[..., 0] * kx], dim=-1)
    Ey_k = torch.stack([phi_k[..., 1] * ky, -phi_k[..., 0] * ky], dim=-1)
    phi = torch.fft.irfft2(torch.view_as_complex(phi_k), s=(Nx, Ny))
    Ex = torch.fft.irfft2(torch.view_as_complex(Ex_k.contiguous()), s=(Nx, Ny))
    Ey = torch.fft.irfft2(torch.view_as_complex(Ey_k.contiguous()), s=(Nx, Ny))
    return Ex, Ey, delta_rho, phi


def vlasov_step_2d(f, dt, vx, vy, kx, ky, inv_ksq, dx, dy, dvx, dvy, physics_mode=True):
    """vlasov_step_impl: returns (f, delta_rho, phi, Ex, Ey)."""
    f = advect_axis(f, vx[None, :, None], 0, dt / 2.0, dx, True)
    f = advect_axis(f, vy[None, None, :], 1, dt / 2.0, dy, True)
    Nx, Ny = f.shape[0], f.shape[1]
    if physics_mode:
        rho = torch.sum(f, dim=(2, 3)) * dvx * dvy
        Ex, Ey, drho, phi = compute_poisson_real(rho, kx, ky, inv_ksq)
    else:
        Ex = Ey = drho = phi = torch.zeros((Nx, Ny), dtype=f.dtype)
    f = advect_axis(f, Ex[:, :, None], 2, dt, dvx, False)
    f = advect_axis(f, Ey[:, :, None], 3, dt, dvy, False)
    f = advect_axis(f, vy[None, None, :], 1, dt / 2.0, dy, True)
    f = advect_axis(f, vx[None, :, None], 0, dt / 2.0, dx, True)
    return f, drho, phi, Ex, Ey


def vlasov_step_1d(f, dt, v, inv_kx, inv_kx_sq, dx, dv, physics_mode=True):
    """vlasov_step (1D1V): returns (f, rho, phi, E)."""
    f = advect_rows(f.T, v, dt / 2.0, dx, True).T
    if physics_mode:
        rho = torch.sum(f, dim=1) * dv - 1.0
        rk = torch.fft.rfft(rho)
        re, im = rk.real, rk.imag
        E = torch.fft.irfft(torch.complex(im * inv_kx, -re * inv_kx), n=rho.shape[0])
        phi = torch.fft.irfft(torch.complex(re * inv_kx_sq, im * inv_kx_sq), n=rho.shape[0])
    else:
        E = rho = phi = torch.zeros(f.shape[0], dtype=f.dtype)
    f = advect_rows(f, E, dt, dv, False)
    f = advect_rows(f.T, v, dt / 2.0, dx, True).T
    return f.contiguous(), rho, phi, E
