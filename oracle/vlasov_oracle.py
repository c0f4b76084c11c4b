"""CPU oracle for the semi-Lagrangian Vlasov-Poisson step (NumPy, float64).

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker (or as the timed CPU baseline).

It restates, operation for operation, the reference PyTorch solver
(``simulations/vlasov1d_1v/pytorch/src/vlasov1D1V.py`` = P1 and
``simulations/vlasov2d_2v/pytorch/src/vlasov2D2V.py`` = P2 of
yasahi-hpc/JAX-PyTorch-Vlasov).  Element-wise arithmetic follows the
reference's evaluation order exactly (true divisions by literal constants,
left-to-right sums, no fused multiply-add), so the advection sweeps are
bit-for-bit what PyTorch-CPU eager computes for the reference code.
Reductions (``numpy.sum``) and FFTs (``numpy.fft``, pocketfft) round
differently from ``torch.sum`` / ``torch.fft`` at the 1e-16 level.

How each function is pinned (tests/test_oracle_*.py):

* ``lagrange5_weights`` / ``lagrange5_interp_sum``: exactness on x^0..x^5 at
  alpha in {0.1, 0.25, 0.5, 0.75, 0.9}, alpha=0 returns the centre tap,
  partition of unity (P1 tests/test_interpolation.py:4-65), and equality with
  the generic Lagrange product form (P2:210-235).
* ``advect_rows`` / ``advect_axis_direct``: brute-force dense interpolation
  matrix on tiny grids, integer shifts are exact permutations, periodic drift
  conserves the sum, analytic Gaussian shifts (P1 tests/test_advection.py:50-195),
  fully out-of-domain Dirichlet shift gives zero.
* ``advect_axis`` (the line form the steps use: weights per (velocity, index),
  one gather per tap, large arrays in threaded blocks): array-equal to
  ``advect_axis_direct`` for every velocity layout of the steps, multi-cell
  and out-of-domain shifts, blocks forced on (tests/test_oracle_interp_advect.py).
* ``solve_poisson_real_1d`` / ``compute_poisson_real``: single-mode closed
  forms (P1 tests/test_poisson.py:6-64, P2 tests/test_poisson.py:6-75), a dense
  DFT matrix restatement of rfft/irfft including the dropped imaginary parts
  of the DC and Nyquist bins.
* ``vlasov_step_1d`` / ``vlasov_step_2d``: dt=0 is the identity, mass
  conservation, Strang composition with E=0, the Landau damping rate against
  the root of the plasma dispersion relation.
* ``field_energy_*`` / ``moments_*``: closed forms on separable inputs.
* ``advect_rows_reflected`` / ``advect_axis_reflected`` (solver 1): integer
  shifts are exact rolls or zero-filled shifts, agreement with the departure
  form to roundoff (tests/test_oracle_reflected.py).
* ``precompute_x_drift`` / ``advect_x_static`` / ``build_v_tap_weights``
  (solver 2): the static drift equals the reflected sweep, tap weights sum to
  one, the window is exact inside |E dt/dv| < S and drops outer taps beyond.
* ``vlasov_step_1d_reflected``: agreement with ``vlasov_step_1d`` to 1e-10
  (the reference's own tolerance, T2/test_advection.py:526).
* Committed fixtures (tests/golden/, written from this module) are re-derived
  by tests/test_golden.py, which also checks their Landau damping rate.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

STENCIL_OFFSETS = np.arange(-2, 4, dtype=np.int64)


# --------------------------------------------------------------------------
# Interpolation (P1:163-210, P2:238-285)
# --------------------------------------------------------------------------

def lagrange5_weights(a):
    """6-point Lagrange weights on nodes -2..3 in the reference's power form.

    Follows P1:191-210 / P2:263-285 term by term (same divisions by literal
    constants, same left-to-right association).  Returns a list of 6 arrays.
    """
    a = np.asarray(a, dtype=np.float64)
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
    return [w0, w1, w2, w3, w4, w5]


def lagrange5_interp_sum(alpha, vals):
    """sum_k w_k(alpha) * vals[..., k], left to right (P1:163-189)."""
    w = lagrange5_weights(alpha)
    return (w[0] * vals[..., 0] + w[1] * vals[..., 1] + w[2] * vals[..., 2]
            + w[3] * vals[..., 3] + w[4] * vals[..., 4] + w[5] * vals[..., 5])


def lagrange5_product_form(alpha):
    """Generic product-form Lagrange basis on nodes -2..3 (P2:210-235).

    Used only to pin ``lagrange5_weights`` (different rounding, same
    polynomials)."""
    alpha = np.asarray(alpha, dtype=np.float64)
    nodes = np.arange(-2, 4, dtype=np.float64)
    out = []
    for j in range(6):
        w = np.ones_like(alpha)
        for m in range(6):
            if m != j:
                w = w * (alpha - nodes[m]) / (nodes[j] - nodes[m])
        out.append(w)
    return out


# --------------------------------------------------------------------------
# Advection (P1:212-256, P2:287-413)
# --------------------------------------------------------------------------

def advect_rows(f_batch, velocities, dt, dx, periodic=True):
    """Semi-Lagrangian shift of each row of ``f_batch`` [B, N].

    Restates ``advect_1d_vectorized`` (P1:212-256, P2:287-353): departure
    index ``i - (v*dt)/dx``, ``floor``/fraction, six taps at offsets -2..3,
    ``remainder`` for periodic rows, clamp + zero mask for Dirichlet rows.
    ``velocities`` is [B] or [B, N].
    """
    f_batch = np.asarray(f_batch, dtype=np.float64)
    B, N = f_batch.shape
    vel = np.asarray(velocities, dtype=np.float64)
    if vel.ndim == 1:
        vel = vel[:, None]
    point_indices = np.arange(N, dtype=np.float64)
    displacement = (vel * dt) / dx
    depart = point_indices[None, :] - displacement
    idx_left = np.floor(depart)
    alpha = depart - idx_left
    idx_left = idx_left.astype(np.int64)
    stencil = idx_left[..., None] + STENCIL_OFFSETS          # [B, N, 6]
    rows = np.arange(B, dtype=np.int64)[:, None, None]
    if periodic:
        vals = f_batch[rows, np.remainder(stencil, N)]
    else:
        valid = (stencil >= 0) & (stencil < N)
        vals = f_batch[rows, np.clip(stencil, 0, N - 1)] * valid.astype(np.float64)
    alpha = np.broadcast_to(alpha, (B, N))
    return lagrange5_interp_sum(alpha, vals)


def advect_axis(f, velocities, axis, dt, dx, periodic=True):
    """``advect_1d`` (P2:356-413): sweep of an array along ``axis``.

    ``velocities`` broadcasts to the batch shape (all dims but ``axis``, in
    order), as in the reference, so the velocity is constant along every
    swept line.  Same arithmetic as ``advect_axis_direct`` element for
    element (P2:330-353): departure ``i - (v*dt)/dx``, ``floor``, fraction,
    power-form weights, taps summed left to right, remainder or clamp+mask;
    the weights and tap indices are evaluated once on the compact
    (velocity, index) shape instead of per element, and the taps are
    gathered one at a time, which keeps the memory at a few copies of f
    (the direct form builds [B, N, 6] int64 index arrays).  Results are
    bit-identical to ``advect_axis_direct`` (tests/test_oracle_interp_advect.py).
    """
    f = np.asarray(f, dtype=np.float64)
    n = f.shape[axis]
    nb = f.ndim - 1
    vel = np.asarray(velocities, dtype=np.float64)
    if vel.ndim > nb:
        raise ValueError("velocities have more dimensions than the batch")
    vel = vel.reshape((1,) * (nb - vel.ndim) + vel.shape)
    np.broadcast_shapes(vel.shape, f.shape[:axis] + f.shape[axis + 1:])  # raises like the reference's broadcast
    # compact shape with the swept axis in place: the gathers below then read
    # whole contiguous rows of the trailing axes
    displacement = np.expand_dims((vel * dt) / dx, axis)
    ishape = [1] * f.ndim
    ishape[axis] = n
    index = np.arange(n, dtype=np.float64).reshape(ishape)
    chunks = _line_chunks(f, axis)
    if len(chunks) == 1:
        return np.ascontiguousarray(_advect_block(f, displacement, index, axis, n, periodic))
    # large arrays: independent blocks along a non-swept axis on a thread pool
    # (numpy releases the GIL in gathers and ufuncs); every element is computed
    # by the same operations as in one block
    out = np.empty_like(f)

    def run(sl):
        d = displacement[sl] if all(displacement.shape[i] > 1 for i in range(f.ndim) if sl[i] != slice(None)) \
            else displacement
        out[sl] = _advect_block(f[sl], d, index, axis, n, periodic)

    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=_oracle_threads()) as ex:
        list(ex.map(run, chunks))
    return out


def _oracle_threads():
    import os
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


# arrays from this many elements up are swept in blocks on a thread pool
CHUNK_MIN_ELEMS = 1 << 21


def _line_chunks(f, axis):
    """Slices of f along its largest non-swept axis (one slice for small f)."""
    if f.size < CHUNK_MIN_ELEMS or f.ndim < 2:
        return [tuple(slice(None) for _ in range(f.ndim))]
    cand = [i for i in range(f.ndim) if i != axis]
    c = max(cand, key=lambda i: f.shape[i])
    parts = min(f.shape[c], 4 * _oracle_threads())
    bounds = [f.shape[c] * p // parts for p in range(parts + 1)]
    out = []
    for a, b in zip(bounds[:-1], bounds[1:]):
        if b > a:
            sl = [slice(None)] * f.ndim
            sl[c] = slice(a, b)
            out.append(tuple(sl))
    return out


def _advect_block(f, displacement, index, axis, n, periodic):
    depart = index - displacement
    idx_left = np.floor(depart)
    w = lagrange5_weights(depart - idx_left)
    idx_left = idx_left.astype(np.int64)
    out = None
    for k in range(6):
        j = idx_left + STENCIL_OFFSETS[k]
        if periodic:
            vals = np.take_along_axis(f, np.broadcast_to(np.remainder(j, n), f.shape), axis=axis)
        else:
            valid = ((j >= 0) & (j < n)).astype(np.float64)
            vals = np.take_along_axis(f, np.broadcast_to(np.clip(j, 0, n - 1), f.shape), axis=axis) * valid
        term = w[k] * vals
        out = term if out is None else out + term
    return out


def advect_axis_direct(f, velocities, axis, dt, dx, periodic=True):
    """``advect_1d`` (P2:356-413) as the reference writes it: permute the
    swept axis last, reshape to [B, N], ``advect_rows`` with one velocity per
    row, permute back.  ``velocities`` broadcasts to the batch shape."""
    f = np.asarray(f, dtype=np.float64)
    perm = list(range(f.ndim))
    perm.append(perm.pop(axis))
    fp = np.transpose(f, perm)
    batch_shape = fp.shape[:-1]
    n = fp.shape[-1]
    flat = np.ascontiguousarray(fp).reshape(-1, n)
    vel = np.broadcast_to(np.asarray(velocities, dtype=np.float64), batch_shape).reshape(-1)
    out = advect_rows(flat, vel, dt, dx, periodic=periodic).reshape(fp.shape)
    inv = [0] * f.ndim
    for i, p in enumerate(perm):
        inv[p] = i
    return np.ascontiguousarray(np.transpose(out, inv))


# --------------------------------------------------------------------------
# Grids (P1:94-122, P2:146-173)
# --------------------------------------------------------------------------

@dataclass
class Grid1D:
    Nx: int
    Nv: int
    L: float
    V_max: float
    dx: float
    dv: float
    x: np.ndarray
    v: np.ndarray
    inv_kx: np.ndarray
    inv_kx_sq: np.ndarray


def grid_1d(Nx, Nv, L, V_max):
    """Grid of P1:94-122 (float64): periodic x without endpoint, v with endpoint."""
    dx = L / Nx
    dv = 2 * V_max / (Nv - 1)
    x = np.linspace(0.0, L, Nx + 1)[:-1]
    v = np.linspace(-V_max, V_max, Nv)
    kx = 2 * math.pi * np.fft.rfftfreq(Nx, d=dx)
    with np.errstate(divide="ignore"):
        inv_kx = np.where(kx == 0, 0.0, 1.0 / np.where(kx == 0, 1.0, kx))
    return Grid1D(Nx, Nv, L, V_max, dx, dv, x, v, inv_kx, inv_kx * inv_kx)


@dataclass
class Grid2D:
    Nx: int
    Ny: int
    Nvx: int
    Nvy: int
    Lx: float
    Ly: float
    Vx_max: float
    Vy_max: float
    dx: float
    dy: float
    dvx: float
    dvy: float
    x: np.ndarray
    y: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    kx: np.ndarray
    ky: np.ndarray
    inv_ksq: np.ndarray


def grid_2d(Nx, Ny, Nvx, Nvy, Lx, Ly, Vx_max, Vy_max):
    """Grid of P2:146-173 (float64); kx is [Nx,1] fftfreq, ky is [1,Ny//2+1] rfftfreq."""
    dx, dy = Lx / Nx, Ly / Ny
    dvx, dvy = 2 * Vx_max / (Nvx - 1), 2 * Vy_max / (Nvy - 1)
    x = np.linspace(0.0, Lx, Nx + 1)[:-1]
    y = np.linspace(0.0, Ly, Ny + 1)[:-1]
    vx = np.linspace(-Vx_max, Vx_max, Nvx)
    vy = np.linspace(-Vy_max, Vy_max, Nvy)
    kx = (2 * math.pi * np.fft.fftfreq(Nx, d=dx))[:, None]
    ky = (2 * math.pi * np.fft.rfftfreq(Ny, d=dy))[None, :]
    ksq = kx ** 2 + ky ** 2
    inv_ksq = np.where(ksq == 0, 0.0, 1.0 / np.where(ksq == 0, 1.0, ksq))
    return Grid2D(Nx, Ny, Nvx, Nvy, Lx, Ly, Vx_max, Vy_max, dx, dy, dvx, dvy,
                  x, y, vx, vy, kx, ky, inv_ksq)


# --------------------------------------------------------------------------
# Field solves (P1:514-572, P2:416-523)
# --------------------------------------------------------------------------

def solve_poisson_real_1d(f, inv_kx, inv_kx_sq, dv):
    """``solve_poisson_real`` of P1:538-572.  Returns (E, rho, phi), rho = n - 1."""
    f = np.asarray(f, dtype=np.float64)
    n = np.sum(f, axis=1) * dv
    rho = n - 1.0
    rho_k = np.fft.rfft(rho)
    re, im = rho_k.real, rho_k.imag
    E_k = (im * inv_kx) + 1j * (-re * inv_kx)
    phi_k = (re * inv_kx_sq) + 1j * (im * inv_kx_sq)
    E = np.fft.irfft(E_k, n=rho.shape[0])
    phi = np.fft.irfft(phi_k, n=rho.shape[0])
    return E, rho, phi


def compute_rho(f, dvx, dvy):
    """``compute_rho`` of P2:416-431: sum over (vx, vy) times dvx, then dvy."""
    return np.sum(np.asarray(f, dtype=np.float64), axis=(2, 3)) * dvx * dvy


def compute_poisson_real(rho, kx, ky, inv_ksq):
    """``compute_poisson_real`` of P2:488-517.  Returns (Ex, Ey, delta_rho, phi)."""
    Nx, Ny = rho.shape
    delta_rho = rho - 1
    rho_k = np.fft.rfft2(delta_rho)
    phi_re = rho_k.real * inv_ksq
    phi_im = rho_k.imag * inv_ksq
    Ex_k = (phi_im * kx) + 1j * (-phi_re * kx)
    Ey_k = (phi_im * ky) + 1j * (-phi_re * ky)
    phi = np.fft.irfft2(phi_re + 1j * phi_im, s=(Nx, Ny))
    Ex = np.fft.irfft2(Ex_k, s=(Nx, Ny))
    Ey = np.fft.irfft2(Ey_k, s=(Nx, Ny))
    return Ex, Ey, delta_rho, phi


def solve_poisson_real_2d(f, kx, ky, inv_ksq, dvx, dvy):
    """``solve_poisson_real`` of P2:520-523."""
    return compute_poisson_real(compute_rho(f, dvx, dvy), kx, ky, inv_ksq)


# --------------------------------------------------------------------------
# Strang steps (P1:574-622, P2:943-998)
# --------------------------------------------------------------------------

def vlasov_step_1d(f, dt, v, inv_kx, inv_kx_sq, dx, dv, physics_mode=True):
    """``vlasov_step`` of P1:574-622 on f [Nx, Nv].  Returns (f, rho, phi, E)."""
    f = np.asarray(f, dtype=np.float64)
    f = advect_rows(f.T, v, dt / 2.0, dx, periodic=True).T
    if physics_mode:
        E, rho, phi = solve_poisson_real_1d(f, inv_kx, inv_kx_sq, dv)
    else:
        E = np.zeros(f.shape[0])
        rho = np.zeros(f.shape[0])
        phi = np.zeros(f.shape[0])
    f = advect_rows(f, E, dt, dv, periodic=False)
    f = advect_rows(f.T, v, dt / 2.0, dx, periodic=True).T
    return np.ascontiguousarray(f), rho, phi, E


def vlasov_step_2d(f, dt, g: Grid2D, physics_mode=True, grid_arrays=None):
    """``vlasov_step_impl`` of P2:943-998 on f [Nx, Ny, Nvx, Nvy].

    Returns (f, delta_rho, phi, Ex, Ey).  ``grid_arrays`` may override
    (vx, vy, kx, ky, inv_ksq) with the exact arrays a caller used.
    """
    vx, vy, kx, ky, inv_ksq = grid_arrays if grid_arrays is not None else (
        g.vx, g.vy, g.kx, g.ky, g.inv_ksq)
    f = advect_axis(f, vx[None, :, None], 0, dt / 2.0, g.dx, periodic=True)
    f = advect_axis(f, vy[None, None, :], 1, dt / 2.0, g.dy, periodic=True)
    if physics_mode:
        Ex, Ey, rho, phi = solve_poisson_real_2d(f, kx, ky, inv_ksq, g.dvx, g.dvy)
    else:
        Ex = np.zeros((g.Nx, g.Ny))
        Ey = np.zeros((g.Nx, g.Ny))
        rho = np.zeros((g.Nx, g.Ny))
        phi = np.zeros((g.Nx, g.Ny))
    f = advect_axis(f, Ex[:, :, None], 2, dt, g.dvx, periodic=False)
    f = advect_axis(f, Ey[:, :, None], 3, dt, g.dvy, periodic=False)
    f = advect_axis(f, vy[None, None, :], 1, dt / 2.0, g.dy, periodic=True)
    f = advect_axis(f, vx[None, :, None], 0, dt / 2.0, g.dx, periodic=True)
    return f, rho, phi, Ex, Ey


# --------------------------------------------------------------------------
# Solver-1/2 forms (reflected stencil)
# --------------------------------------------------------------------------

def advect_rows_reflected(f_batch, velocities, dt, dx, periodic=True, window=-1):
    """Shift form of the sweep on f_batch [B, N] (one velocity per row or per
    point): d = (v dt)/dx, s = floor(d), alpha = d - s, out[i] = sum_k
    w_k(alpha) f[i - s - o_k], o_k = k - 2, accumulated from 0 in k order.
    Follows advect_1d_matrix_diag (P1:258-337: _roll_columns/_roll_rows and
    build_v_diag_weights) and, with window >= 0, advect_v_static
    (P1:479-501), which keeps only taps m = s + o_k in [-window-2, window+2].
    """
    f = np.asarray(f_batch, dtype=np.float64)
    B, N = f.shape
    vel = np.broadcast_to(np.asarray(velocities, dtype=np.float64).reshape(B, -1), (B, N))
    d = (vel * dt) / dx
    s = np.floor(d)
    w = lagrange5_weights(d - s)
    s = s.astype(np.int64)
    i = np.arange(N)[None, :]
    out = np.zeros_like(f)
    rows = np.arange(B)[:, None]
    for k in range(6):
        o = k - 2
        j = i - s - o
        if periodic:
            vals = f[rows, j % N]
        else:
            valid = (j >= 0) & (j < N)
            if window >= 0:
                m = s + o
                valid &= (m >= -(window + 2)) & (m <= window + 2)
            vals = np.where(valid, f[rows, np.clip(j, 0, N - 1)], 0.0)
        out = out + vals * w[k]
    return out


def advect_axis_reflected(f, velocities, axis, dt, dx, periodic=True, window=-1):
    """``advect_rows_reflected`` along any axis (velocities broadcast against
    f with the swept axis removed, or of f's full shape)."""
    f = np.asarray(f, dtype=np.float64)
    v = np.asarray(velocities, dtype=np.float64)
    fm = np.moveaxis(f, axis, -1)
    if v.ndim == f.ndim and v.shape == f.shape:
        vm = np.moveaxis(v, axis, -1)
    else:
        vm = np.broadcast_to(np.moveaxis(np.broadcast_to(v, f.shape[:axis] + (1,) + f.shape[axis + 1:]), axis, -1),
                             fm.shape)
    out = advect_rows_reflected(fm.reshape(-1, fm.shape[-1]), vm.reshape(-1, fm.shape[-1]), dt, dx, periodic, window)
    return np.moveaxis(out.reshape(fm.shape), -1, axis)


def precompute_x_drift(v, dt, dx):
    """(blocks, W) of P1:357-387: s = floor((v dt)/dx) in contiguous
    (start, stop, shift) blocks and W = w_k(alpha) [6, Nv]."""
    d = (np.asarray(v, dtype=np.float64) * dt) / dx
    s = np.floor(d)
    alpha = d - s
    s = s.astype(np.int64)
    edges = np.flatnonzero(np.diff(s)) + 1
    starts = np.concatenate(([0], edges))
    stops = np.concatenate((edges, [len(s)]))
    blocks = [(int(a), int(b), int(s[a])) for a, b in zip(starts, stops)]
    return blocks, lagrange5_weights(alpha)


def advect_x_static(f, W, blocks, axis=0):
    """advect_x_static (P1:408-448; P2:756-819 with axis): periodic sweep
    along `axis` with the velocity index on axis + (1 in 1D, 2 in 2D)."""
    f = np.asarray(f, dtype=np.float64)
    vel_axis = 1 if f.ndim == 2 else axis + 2
    shifts = np.array([s for a, b, s in blocks for _ in range(b - a)], dtype=np.int64)
    n = f.shape[axis]
    shape = [1] * f.ndim
    shape[vel_axis] = len(shifts)
    sh = shifts.reshape(shape)
    ishape = [1] * f.ndim
    ishape[axis] = n
    rows = np.arange(n).reshape(ishape) - sh
    out = np.zeros_like(f)
    for k in range(6):
        idx = np.broadcast_to(np.remainder(rows - (k - 2), n), f.shape)
        out = out + np.take_along_axis(f, idx, axis=axis) * np.asarray(W[k]).reshape(shape)
    return out


def build_v_tap_weights(E, dt, dv, S):
    """[2S+5, ...] tap weights of P1:450-477 on m = -S-2 .. S+2."""
    d = (np.asarray(E, dtype=np.float64) * dt) / dv
    s = np.floor(d)
    W6 = lagrange5_weights(d - s)
    s = s.astype(np.int64)
    rows = []
    for m in range(-(S + 2), S + 3):
        acc = np.zeros_like(d)
        for k in range(6):
            acc = acc + (s == m - (k - 2)).astype(np.float64) * W6[k]
        rows.append(acc)
    return np.stack(rows)


def vlasov_step_1d_reflected(f, dt, v, inv_kx, inv_kx_sq, dx, dv, physics_mode=True, window=-1):
    """vlasov_step_matrix_diag (P1:625-665, window < 0) and the gather-free
    step (P1:668-755, window = S) on f [Nx, Nv]: (f, rho, phi, E)."""
    f = np.asarray(f, dtype=np.float64)
    f = advect_rows_reflected(f.T, v, dt / 2.0, dx, periodic=True).T
    if physics_mode:
        E, rho, phi = solve_poisson_real_1d(f, inv_kx, inv_kx_sq, dv)
    else:
        E = rho = phi = np.zeros(f.shape[0])
    f = advect_rows_reflected(f, E, dt, dv, periodic=False, window=window)
    f = advect_rows_reflected(f.T, v, dt / 2.0, dx, periodic=True).T
    return np.ascontiguousarray(f), rho, phi, E


# --------------------------------------------------------------------------
# Diagnostics (no reference function; defined by this library)
# --------------------------------------------------------------------------

def field_energy_1d(E, dx):
    """W_E = 1/2 sum E^2 dx."""
    return 0.5 * float(np.sum(np.asarray(E) ** 2)) * dx


def field_energy_2d(Ex, Ey, dx, dy):
    """W_E = 1/2 sum (Ex^2 + Ey^2) dx dy."""
    return 0.5 * float(np.sum(np.asarray(Ex) ** 2 + np.asarray(Ey) ** 2)) * dx * dy


def moments_1d(f, v, dx, dv):
    """(mass, kinetic energy, L2 norm squared) of f [Nx, Nv]."""
    f = np.asarray(f, dtype=np.float64)
    mass = float(np.sum(f)) * dx * dv
    kin = 0.5 * float(np.sum(f * (v[None, :] ** 2))) * dx * dv
    l2 = float(np.sum(f * f)) * dx * dv
    return mass, kin, l2


def moments_2d(f, vx, vy, dx, dy, dvx, dvy):
    """(mass, kinetic energy, L2 norm squared) of f [Nx, Ny, Nvx, Nvy]."""
    f = np.asarray(f, dtype=np.float64)
    cell = dx * dy * dvx * dvy
    v2 = vx[:, None] ** 2 + vy[None, :] ** 2
    mass = float(np.sum(f)) * cell
    kin = 0.5 * float(np.sum(f * v2[None, None])) * cell
    l2 = float(np.sum(f * f)) * cell
    return mass, kin, l2


# --------------------------------------------------------------------------
# Linear theory for the pinning tests
# --------------------------------------------------------------------------

def landau_root(k, guess=(1.4156, -0.1533)):
    """Root of 1 + (1 + zeta Z(zeta))/k^2 = 0, zeta = omega / (sqrt(2) k).

    Z is the plasma dispersion function, Z(z) = i sqrt(pi) w(z), with w the
    Faddeeva function (scipy.special.wofz).  Returns complex omega
    (omega_r - i gamma).  Newton iteration with a numerical derivative.
    """
    from scipy.special import wofz

    def D(om):
        z = om / (math.sqrt(2.0) * k)
        Z = 1j * math.sqrt(math.pi) * wofz(z)
        return 1.0 + (1.0 + z * Z) / (k * k)

    om = complex(guess[0], guess[1])
    for _ in range(100):
        h = 1e-7
        d = (D(om + h) - D(om - h)) / (2 * h)
        step = D(om) / d
        om -= step
        if abs(step) < 1e-14:
            break
    return om
