"""1D1V Vlasov-Poisson solver: drop-in for the reference PyTorch module.

Mirrors the public interface of
``simulations/vlasov1d_1v/pytorch/src/vlasov1D1V.py`` (cited as P1 below):
the same ``Grid`` / ``Field`` / ``Variables`` / ``SimulationState``
dataclasses, the same kernel-level functions (``advect_1d_vectorized``,
``solve_poisson``, ``solve_poisson_real``, ``vlasov_step``) with the same
positional signatures, and the same ``Vlasov`` driver.  Every computation
runs in the hand-written HIP kernels of ``libvlasov_amd.so``; tensors must be
on a ROCm GPU (``device='cuda'``), there is no CPU path.
"""
from __future__ import annotations

import math
import pathlib
import platform
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr, require_gpu, stream_handle


# ---------------------------------------------------------------------------
# Device helpers (P1:19-74)
# ---------------------------------------------------------------------------

def get_device_label(device: torch.device) -> str:
    if device.type == "cuda":
        idx = device.index if device.index is not None else torch.cuda.current_device()
        return f"{torch.cuda.get_device_name(idx)} (cuda:{idx})"
    return f"{platform.processor() or platform.machine()} (cpu)"


def synchronize(device: torch.device):
    if device.type == "cuda":
        torch.cuda.synchronize()


def reset_peak_memory(device: torch.device):
    if device.type == "cuda":
        torch.cuda.reset_peak_memory_stats(device)


def get_memory_stats(device: torch.device):
    synchronize(device)
    if device.type == "cuda":
        return {
            "allocated": torch.cuda.memory_allocated(device),
            "reserved": torch.cuda.memory_reserved(device),
            "peak_allocated": torch.cuda.max_memory_allocated(device),
            "peak_reserved": torch.cuda.max_memory_reserved(device),
        }
    return None


def format_memory(num_bytes: int) -> str:
    return f"{num_bytes / (1024 ** 2):.2f} MiB"


def default_device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("vlasov_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda")


# ---------------------------------------------------------------------------
# Data structures (P1:76-160)
# ---------------------------------------------------------------------------

@dataclass
class Grid:
    Nx: int
    Nv: int
    L: float
    V_max: float
    dx: float
    dv: float
    x: torch.Tensor
    v: torch.Tensor
    inv_kx: torch.Tensor
    inv_kx_sq: torch.Tensor
    x_indices: torch.Tensor
    v_indices: torch.Tensor
    x_batch_offsets: torch.Tensor
    v_batch_offsets: torch.Tensor
    stencil_offsets: torch.Tensor

    @classmethod
    def create(cls, Nx, Nv, L, V_max, device="cuda", dtype=torch.float64):
        """Grid of P1:94-122 (same torch expressions, so identical arrays)."""
        dx = L / Nx
        dv = 2 * V_max / (Nv - 1)
        x = torch.linspace(0, L, Nx + 1, device=device, dtype=dtype)[:-1]
        v = torch.linspace(-V_max, V_max, Nv, device=device, dtype=dtype)
        kx = 2 * math.pi * torch.fft.rfftfreq(Nx, d=dx, device=device, dtype=dtype)
        inv_kx = torch.where(kx == 0, torch.tensor(0.0, device=device, dtype=dtype), 1.0 / kx)
        inv_kx_sq = inv_kx * inv_kx
        index_dtype = torch.int64
        x_indices = torch.arange(Nx, device=device, dtype=dtype)
        v_indices = torch.arange(Nv, device=device, dtype=dtype)
        x_batch_offsets = (torch.arange(Nv, device=device, dtype=index_dtype) * Nx).view(Nv, 1, 1)
        v_batch_offsets = (torch.arange(Nx, device=device, dtype=index_dtype) * Nv).view(Nx, 1, 1)
        stencil_offsets = torch.arange(-2, 4, device=device, dtype=index_dtype)
        return Grid(Nx, Nv, L, V_max, dx, dv, x, v, inv_kx, inv_kx_sq,
                    x_indices, v_indices, x_batch_offsets, v_batch_offsets, stencil_offsets)

    def to(self, device) -> "Grid":
        for name in ("x", "v", "inv_kx", "inv_kx_sq", "x_indices", "v_indices",
                     "x_batch_offsets", "v_batch_offsets", "stencil_offsets"):
            setattr(self, name, getattr(self, name).to(device))
        return self


@dataclass
class Field:
    """Field data (rho, phi)."""
    rho: torch.Tensor
    phi: torch.Tensor


@dataclass
class Variables:
    """Distribution function f and the equilibrium f0."""
    fn: torch.Tensor
    f0: torch.Tensor


@dataclass
class SimulationState:
    field: Field
    variables: Variables
    t: float
    step: int


# ---------------------------------------------------------------------------
# Kernels
# ---------------------------------------------------------------------------

def _f64(t: torch.Tensor) -> torch.Tensor:
    if t.dtype != torch.float64:
        raise TypeError(f"the float64 path needs float64 tensors, got {t.dtype}")
    return t


def advect_rows(f_batch: torch.Tensor, velocities: torch.Tensor, dt: float, dx: float,
                periodic: bool = True) -> torch.Tensor:
    """Sweep every row of f_batch [B, N] (the advected axis is the last one).

    ``velocities`` is [B] (one velocity per row) or [B, N] (per point), as in
    advect_1d_vectorized (P1:212-256).  Returns a new [B, N] tensor.
    """
    dev = require_gpu(f_batch, velocities)
    B, N = f_batch.shape
    vel = velocities.to(torch.float64)
    if vel.ndim == 1:
        vel = vel.expand(B)
        vs = (0, 0, vel.stride(0), 0)
    else:
        vel = vel.expand(B, N)
        vs = (0, 0, vel.stride(0), vel.stride(1))
    return _advect(f_batch, (1, 1, B, N), 3, vel, vs, dt, dx, periodic, dev)


def _advect(f, shape, axis, vel, vs, dt, dx, periodic, dev, out=None):
    f_c = f.contiguous()
    if out is None:
        out = torch.empty_like(f_c)
    if f_c.dtype == torch.float64:
        fn = lib().vp_advect_f64
    elif f_c.dtype == torch.float32:
        fn = lib().vp_advect_f32
    else:
        raise TypeError(f"unsupported dtype {f_c.dtype}")
    rc = fn(f_c.data_ptr(), out.data_ptr(), _lib.i64x4(shape), int(axis), vel.data_ptr(),
            _lib.i64x4(vs), float(dt), float(dx), int(bool(periodic)), stream_handle(dev))
    check(rc, "vp_advect")
    return out


def advect_1d_vectorized(f_batch, velocities, dt, point_indices, batch_offsets, stencil_offsets, dx,
                         periodic=True):
    """Drop-in for advect_1d_vectorized (P1:212-256).

    ``point_indices``, ``batch_offsets`` and ``stencil_offsets`` are accepted
    for signature compatibility; the kernel derives the same indices itself
    (arange(N), row offsets, taps -2..3).
    """
    return advect_rows(f_batch, velocities, dt, dx, periodic=periodic)


def lagrange5_weights(alpha):
    """[6, ...] stacked power-form weights (P1:191-210); element-wise torch ops."""
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
    return torch.stack([w0, w1, w2, w3, w4, w5], dim=0)


def lagrange5_interp_sum(alpha, gather_vals):
    """sum_k w_k(alpha) gather_vals[..., k] (P1:163-189); element-wise torch ops."""
    w = lagrange5_weights(alpha)
    return (w[0] * gather_vals[..., 0] + w[1] * gather_vals[..., 1] + w[2] * gather_vals[..., 2]
            + w[3] * gather_vals[..., 3] + w[4] * gather_vals[..., 4] + w[5] * gather_vals[..., 5])


def compute_rho(f, dv):
    """rho = sum_v f * dv - 1 (P1:559-561), compensated float64 sum."""
    dev = require_gpu(f)
    f_c = f.contiguous()
    Nx, Nv = f_c.shape
    rho = torch.empty(Nx, dtype=torch.float64, device=dev)
    fn = lib().vp_row_sum_f64 if f_c.dtype == torch.float64 else lib().vp_row_sum_f32
    check(fn(f_c.data_ptr(), Nx, Nv, float(dv), 1.0, 1.0, rho.data_ptr(), stream_handle(dev)),
          "vp_row_sum")
    return rho


def poisson_from_rho(rho, inv_kx, inv_kx_sq):
    """E and phi from rho (P1:551-572)."""
    dev = require_gpu(rho, inv_kx, inv_kx_sq)
    Nx = rho.shape[0]
    E = torch.empty(Nx, dtype=torch.float64, device=dev)
    phi = torch.empty(Nx, dtype=torch.float64, device=dev)
    nbytes = lib().vp_poisson_1d_workspace(Nx)
    ws = _lib.Workspace.get(dev, nbytes, "poisson1d")
    rc = lib().vp_poisson_1d_f64(rho.contiguous().data_ptr(), Nx, _f64(inv_kx).contiguous().data_ptr(),
                                 _f64(inv_kx_sq).contiguous().data_ptr(), E.data_ptr(), phi.data_ptr(),
                                 ws.data_ptr(), ws.numel(), stream_handle(dev))
    check(rc, "vp_poisson_1d")
    return E, phi


def solve_poisson_real(f, inv_kx, inv_kx_sq, dv):
    """Drop-in for solve_poisson_real (P1:538-572).  Returns (E, rho, phi)."""
    rho = compute_rho(f, dv)
    E, phi = poisson_from_rho(rho, inv_kx, inv_kx_sq)
    return E, rho, phi


def solve_poisson(f, inv_kx, dv):
    """Drop-in for solve_poisson (P1:514-536): same results as the real form."""
    return solve_poisson_real(f, inv_kx, inv_kx * inv_kx, dv)


def vlasov_step(f, dt, grid_v, grid_inv_kx, grid_inv_kx_sq, x_indices, v_indices, x_batch_offsets,
                v_batch_offsets, stencil_offsets, grid_dx, grid_dv, grid_V_max, physics_mode):
    """Drop-in for vlasov_step (P1:574-622): X(dt/2), Poisson, V(dt), X(dt/2).

    Functional: returns new (f, rho, phi); f is not modified.
    """
    f, rho, phi, _ = step_with_field(f, dt, grid_v, grid_inv_kx, grid_inv_kx_sq, grid_dx, grid_dv,
                                     physics_mode)
    return f, rho, phi


def step_with_field(f, dt, grid_v, grid_inv_kx, grid_inv_kx_sq, grid_dx, grid_dv, physics_mode):
    """One vlasov_step that also returns the mid-step field E: (f, rho, phi, E).
    A float32 f runs the FP32-storage step (vp_step_1d1v_f32): f stays float32,
    rho, phi and E are float64."""
    dev = require_gpu(f, grid_v)
    if f.dtype not in (torch.float64, torch.float32):
        raise TypeError(f"f must be float64 or float32 (FP32 storage), got {f.dtype}")
    f_c = f.contiguous()
    Nx, Nv = f_c.shape
    out = torch.empty_like(f_c)
    E = torch.empty(Nx, dtype=torch.float64, device=dev)
    rho = torch.empty(Nx, dtype=torch.float64, device=dev)
    phi = torch.empty(Nx, dtype=torch.float64, device=dev)
    nbytes = lib().vp_step_1d1v_workspace(Nx, Nv)
    ws = _lib.Workspace.get(dev, nbytes, "step1d")
    pm = bool(physics_mode)
    fn = lib().vp_step_1d1v_f64 if f_c.dtype == torch.float64 else lib().vp_step_1d1v_f32
    rc = fn(
        f_c.data_ptr(), out.data_ptr(), Nx, Nv, float(dt), _f64(grid_v).contiguous().data_ptr(),
        ptr(grid_inv_kx.contiguous()) if pm else None, ptr(grid_inv_kx_sq.contiguous()) if pm else None,
        float(grid_dx), float(grid_dv), int(pm), E.data_ptr(), rho.data_ptr(), phi.data_ptr(),
        ws.data_ptr(), ws.numel(), stream_handle(dev))
    check(rc, "vp_step_1d1v")
    return out, rho, phi, E


# ---------------------------------------------------------------------------
# Solver 1 and solver 2 (shift forms, P1:258-512, P1:625-755)
# ---------------------------------------------------------------------------
from . import shift_forms as _sf  # noqa: E402

static_stencil_offsets = _sf.static_stencil_offsets
precompute_x_drift = _sf.precompute_x_drift
build_v_tap_weights = _sf.build_v_tap_weights
check_v_window = _sf.check_v_window


def build_v_diag_weights(stencil_offsets, shifts, alpha, n_points):
    """[6, Nx, Nv] banded weights of the solver-1 V sweep (P1:279-296)."""
    return _sf.build_diag_weights(stencil_offsets, shifts, alpha, n_points)


def advect_1d_matrix_diag(f_batch, velocities, dt, dx, stencil_offsets, periodic=True):
    """Drop-in for advect_1d_matrix_diag (P1:302-337): periodic -> sweep
    along dim 0 with one velocity per column ([Nv]); else along dim 1 with
    one velocity per row ([Nx]), zero outside."""
    _sf._check_offsets(stencil_offsets)
    Nx, Nv = f_batch.shape
    f4 = f_batch.reshape(1, 1, Nx, Nv)
    vel = velocities.reshape(-1)
    if periodic:
        out = _sf.sweep_reflected(f4, 2, vel, (0, 0, 0, 1), dt, dx, True)
    else:
        out = _sf.sweep_reflected(f4, 3, vel, (0, 0, 1, 0), dt, dx, False)
    return out.reshape(Nx, Nv)


def block_roll(f, blocks):
    """Roll each column block by its static shift along dim 0 (P1:389-406)."""
    Nx, Nv = f.shape
    if all(shift % Nx == 0 for _, _, shift in blocks):
        return f
    out = _sf.static_sweep(f.reshape(1, 1, Nx, Nv), 2, _sf.block_shifts(blocks, f.device),
                           _sf.roll_weights(Nv, f.device), (0, 0, 0, 1))
    return out.reshape(Nx, Nv)


def advect_x_static(f, W_x, blocks, stencil_offsets):
    """Drop-in for advect_x_static (P1:408-448): periodic X drift with the
    static shifts of `blocks` and the weights W_x [6, Nv]."""
    _sf._check_offsets(stencil_offsets)
    Nx, Nv = f.shape
    out = _sf.static_sweep(f.reshape(1, 1, Nx, Nv), 2, _sf.block_shifts(blocks, f.device), W_x, (0, 0, 0, 1))
    return out.reshape(Nx, Nv)


def advect_v_static(f, E, dt, dv, S, stencil_offsets):
    """Drop-in for advect_v_static (P1:479-501): Dirichlet V kick exact while
    |E dt/dv| < S (taps outside the static window are dropped, as there)."""
    _sf._check_offsets(stencil_offsets)
    Nx, Nv = f.shape
    out = _sf.sweep_reflected(f.reshape(1, 1, Nx, Nv), 3, E.reshape(-1), (0, 0, 1, 0), dt, dv, False, int(S))
    return out.reshape(Nx, Nv)


def _field(f, physics_mode, inv_kx, inv_kx_sq, dv):
    if physics_mode:
        return solve_poisson_real(f, inv_kx, inv_kx_sq, dv)
    z = f.new_zeros((f.shape[0],))
    return z, z.clone(), z.clone()


def _shift_step(f, dt, v, x_shift, W_x, inv_kx, inv_kx_sq, dx, dv, physics_mode, window):
    """One solver-1 (x_shift is None) or solver-2 step through
    vp_step_1d1v_shift_f64: (f, rho, phi, E)."""
    dev = require_gpu(f)
    if f.dtype != torch.float64:
        raise TypeError("the shift forms run float64")
    f_c = f.contiguous()
    Nx, Nv = f_c.shape
    out = torch.empty_like(f_c)
    E = torch.empty(Nx, dtype=torch.float64, device=dev)
    rho = torch.empty(Nx, dtype=torch.float64, device=dev)
    phi = torch.empty(Nx, dtype=torch.float64, device=dev)
    ws = _lib.Workspace.get(dev, lib().vp_step_1d1v_workspace(Nx, Nv), "step1d")
    pm = bool(physics_mode)
    if x_shift is None:
        v_c = _f64(v).contiguous()
        sh_p, w_p, w_stride = None, None, 0
    else:
        v_c = None
        W_c = W_x.to(torch.float64).contiguous()
        sh_p, w_p, w_stride = x_shift.data_ptr(), W_c.data_ptr(), W_c.shape[1]
    rc = lib().vp_step_1d1v_shift_f64(
        f_c.data_ptr(), out.data_ptr(), Nx, Nv, float(dt), ptr(v_c) if v_c is not None else None, sh_p, w_p,
        w_stride, ptr(inv_kx.contiguous()) if pm else None, ptr(inv_kx_sq.contiguous()) if pm else None,
        float(dx), float(dv), int(pm), int(window), E.data_ptr(), rho.data_ptr(), phi.data_ptr(), ws.data_ptr(),
        ws.numel(), stream_handle(dev))
    check(rc, "vp_step_1d1v_shift_f64")
    return out, rho, phi, E


def vlasov_step_matrix_diag(f, dt, grid_v, grid_inv_kx, grid_inv_kx_sq, x_indices, v_indices, x_batch_offsets,
                            v_batch_offsets, stencil_offsets, grid_dx, grid_dv, grid_V_max, physics_mode):
    """Drop-in for vlasov_step_matrix_diag (P1:625-665), the solver-1 step:
    advect_1d_matrix_diag X(dt/2), the field, V(dt, E), X(dt/2), in one
    library call."""
    _sf._check_offsets(stencil_offsets)
    f, rho, phi, _ = _shift_step(f, dt, grid_v, None, None, grid_inv_kx, grid_inv_kx_sq, grid_dx, grid_dv,
                                 physics_mode, -1)
    return f, rho, phi


def vlasov_step_gather_free_impl(f, dt, W_x, blocks, stencil_offsets, grid_inv_kx, grid_inv_kx_sq, grid_dv,
                                 physics_mode, S):
    """Drop-in for vlasov_step_gather_free_impl (P1:668-701), the solver-2
    core: advect_x_static, the field, advect_v_static, advect_x_static, in
    one library call."""
    _sf._check_offsets(stencil_offsets)
    x_shift = _sf.block_shifts(blocks, f.device)
    f, rho, phi, _ = _shift_step(f, dt, None, x_shift, W_x, grid_inv_kx, grid_inv_kx_sq, 1.0, grid_dv,
                                 physics_mode, int(S))
    return f, rho, phi


def vlasov_step_gather_free(f, dt, grid_v, grid_inv_kx, grid_inv_kx_sq, x_indices, v_indices, x_batch_offsets,
                            v_batch_offsets, stencil_offsets, grid_dx, grid_dv, grid_V_max, physics_mode, S=1):
    """Drop-in for vlasov_step_gather_free (P1:703-755): precompute, then the core."""
    blocks, W_x = precompute_x_drift(grid_v, dt / 2.0, grid_dx, f.dtype)
    return vlasov_step_gather_free_impl(f, dt, W_x, blocks, static_stencil_offsets(stencil_offsets), grid_inv_kx,
                                        grid_inv_kx_sq, grid_dv, physics_mode, S)


def field_energy(E, dx) -> torch.Tensor:
    """W_E = 1/2 sum E^2 dx (compensated), a [1] float64 device tensor."""
    dev = require_gpu(E)
    out = torch.empty(1, dtype=torch.float64, device=dev)
    check(lib().vp_field_energy_f64(_f64(E).contiguous().data_ptr(), None, E.numel(), float(dx), 1.0,
                                    out.data_ptr(), stream_handle(dev)), "vp_field_energy_f64")
    return out


def moments(f, grid) -> torch.Tensor:
    """(mass, kinetic energy, L2 norm squared) of f [Nx, Nv]:
    sum f dx dv, 1/2 sum f v^2 dx dv, sum f^2 dx dv (compensated), [3] float64."""
    dev = require_gpu(f)
    f_c = f.contiguous()
    Nx, Nv = f_c.shape
    out = torch.empty(3, dtype=torch.float64, device=dev)
    ws = _lib.Workspace.get(dev, lib().vp_moments_workspace(), "moments")
    fn = lib().vp_moments_f64 if f_c.dtype == torch.float64 else lib().vp_moments_f32
    check(fn(f_c.data_ptr(), Nx, Nv, grid.v.to(torch.float64).contiguous().data_ptr(), None, 0, float(grid.dx),
             float(grid.dv), out.data_ptr(), ws.data_ptr(), ws.numel(), stream_handle(dev)), "vp_moments")
    return out


# ---------------------------------------------------------------------------
# Driver (P1:757-994)
# ---------------------------------------------------------------------------

class Vlasov:
    dt: float = 0.05
    diag_steps: int = 10
    out_dir: str = "data"
    physics_mode: bool = False
    epsilon: float = 0.001
    dtype: torch.dtype = torch.float64

    def __init__(self, Nx: int = 128, Nv: int = 128, L: float = 2.0 * math.pi, V_max: float = 6.0,
                 nbiter: int = 10, diag_steps: int = 10, dt: float = 0.05, out_dir: str = "data",
                 physics_mode: bool = False, epsilon: float = 0.001, dtype: str = "float64",
                 solver: int = 0, device=None, snapshot_format: str = "npz") -> None:
        """Same parameters as the reference constructor (P1:770-876).

        ``solver`` selects the step as in the reference: 0 vlasov_step,
        1 vlasov_step_matrix_diag, 2 vlasov_step_gather_free (static shift
        data hoisted out of the loop).  dtype='float32' is the FP32-storage
        variant (solver 0): f is float32, the grid, rho, phi and E stay
        float64 (the reference's float32 mode computes everything in
        float32).
        """
        self.output_log = ""
        if dtype not in ("float64", "float32"):
            raise ValueError(f"dtype must be 'float64' or 'float32', got {dtype!r}")
        if dtype == "float32" and solver != 0:
            raise ValueError("solvers 1 and 2 run float64 (the FP32-storage variant uses solver 0)")
        self.f_dtype = torch.float64 if dtype == "float64" else torch.float32
        self.dtype = torch.float64
        self.device = torch.device(device) if device is not None else default_device()
        self.grid = Grid.create(Nx, Nv, L, V_max, device=self.device, dtype=self.dtype)
        self.nbiter = nbiter
        self.diag_steps = diag_steps
        self.dt = dt
        self.out_dir = out_dir
        if snapshot_format not in ("npz", "bin"):
            raise ValueError(f"snapshot_format must be 'npz' or 'bin', got {snapshot_format!r}")
        self.snapshot_format = snapshot_format
        self.physics_mode = physics_mode
        self.epsilon = epsilon
        solver_fns = {0: vlasov_step, 1: vlasov_step_matrix_diag, 2: vlasov_step_gather_free}
        solver_names = {0: "vlasov_step", 1: "vlasov_step_matrix_diag", 2: "vlasov_step_gather_free"}
        self.solver = solver
        self.solver_type = solver_names[solver]
        self._step_fn = solver_fns[solver]
        if solver == 2:
            # static drift data hoisted out of the loop, as in P1:843-865
            g = self.grid
            blocks, W_x = precompute_x_drift(g.v, self.dt / 2.0, g.dx, self.dtype)
            offsets = static_stencil_offsets(g.stencil_offsets)

            def _step_gather_free(f, dt, grid_v, grid_inv_kx, grid_inv_kx_sq, x_indices, v_indices,
                                  x_batch_offsets, v_batch_offsets, stencil_offsets, grid_dx, grid_dv, grid_V_max,
                                  physics_mode):
                return vlasov_step_gather_free_impl(f, dt, W_x, blocks, offsets, grid_inv_kx, grid_inv_kx_sq,
                                                    grid_dv, physics_mode, 1)

            self._step_fn = _step_gather_free
        self.energy_history = np.zeros(0)
        self.moment_history: list = []  # (step, t, mass, kinetic, L2) at every snapshot
        pathlib.Path(self.out_dir).mkdir(parents=True, exist_ok=True)
        num_steps = self.nbiter // self.diag_steps
        message = f"PyTorch version: {torch.__version__}\n"
        message += f"Starting simulation on {self.device} with {self.f_dtype} precision "
        message += "with Poisson solver" if self.physics_mode else "without Poisson solver"
        message += f" and solver type {self.solver_type}."
        self._append_log(message)
        self._append_log(f"Grid: {self.grid.Nx}x{self.grid.Nv}, Steps: {self.nbiter}, "
                         f"chunks: {num_steps}, diag_freq: {self.diag_steps}")

    def _append_log(self, message: str) -> None:
        self.output_log += f"{message}\n"

    def initialize(self) -> SimulationState:
        """Initial state of P1:881-911 (Landau damping or Gaussian pulse)."""
        x_2d = self.grid.x.unsqueeze(1)
        v_2d = self.grid.v.unsqueeze(0)
        if self.physics_mode:
            kx = 2.0 * math.pi / self.grid.L
            maxwellian = (1.0 / math.sqrt(2.0 * math.pi)) * torch.exp(-0.5 * v_2d ** 2)
            perturbation = 1.0 + self.epsilon * torch.cos(kx * x_2d)
            f = perturbation * maxwellian
            f0 = maxwellian.squeeze(0)
        else:
            x0, v0 = self.grid.L / 2.0, 0.0
            sigma_x, sigma_v = self.grid.L / 10.0, self.grid.V_max / 4.0
            f = torch.exp(-0.5 * ((x_2d - x0) ** 2 / sigma_x ** 2 + (v_2d - v0) ** 2 / sigma_v ** 2))
            f0 = torch.zeros_like(self.grid.v, dtype=self.dtype)
        f = f.to(self.f_dtype).contiguous()
        E, rho, phi = solve_poisson_real(f, self.grid.inv_kx, self.grid.inv_kx_sq, self.grid.dv)
        state = SimulationState(field=Field(rho, phi), variables=Variables(fn=f, f0=f0), t=0.0, step=0)
        self.save_diag(state)
        return state

    def save_diag(self, state: SimulationState) -> None:
        """Snapshot to ``field_%06d.npz`` (the reference writes netCDF via xarray),
        or with snapshot_format='bin' to the Kokkos solver's raw
        ``fn/f0/rho/phi_%06d.bin`` files; records the moments (mass, kinetic
        energy, L2) of f."""
        mom = moments(state.variables.fn, self.grid).cpu().numpy()
        self.moment_history.append((int(state.step), float(state.t), *[float(m) for m in mom]))
        if self.snapshot_format == "bin":
            from . import snapshots
            snapshots.write_snapshot(self.out_dir, state.step, state.variables.fn, state.variables.f0,
                                     snapshots.kokkos_rho(state.field.rho), state.field.phi)
            return
        filename = pathlib.Path(self.out_dir) / f"field_{int(state.step):06d}.npz"
        np.savez(filename, f=state.variables.fn.detach().cpu().numpy(),
                 f0=state.variables.f0.detach().cpu().numpy(),
                 rho=state.field.rho.detach().cpu().numpy(), phi=state.field.phi.detach().cpu().numpy(),
                 time=np.array([float(state.t)]), x=self.grid.x.cpu().numpy(), v=self.grid.v.cpu().numpy(),
                 step=int(state.step), dt=self.dt, physics_mode=int(self.physics_mode), moments=mom)

    def restart(self, step: int) -> SimulationState:
        """State of the raw binary snapshot of iteration ``step`` in out_dir
        (written with snapshot_format='bin', or by the reference's Kokkos
        solver); run() continues from it.  f takes this run's precision."""
        from . import snapshots
        g = self.grid
        a = snapshots.read_snapshot(self.out_dir, step, {"fn": (g.Nx, g.Nv), "f0": (g.Nv,), "rho": (g.Nx,),
                                                         "phi": (g.Nx,)})
        t = {k: torch.from_numpy(v.astype(np.float64)).to(self.device) for k, v in a.items()}
        t["fn"] = torch.from_numpy(a["fn"]).to(self.device).to(self.f_dtype)
        t["rho"] = snapshots.delta_rho_of(t["rho"])  # the files hold the full density
        return SimulationState(field=Field(t["rho"], t["phi"]), variables=Variables(fn=t["fn"], f0=t["f0"]),
                               t=float(step) * self.dt, step=int(step))

    def run(self, state: SimulationState) -> SimulationState:
        """Time loop of P1:943-994.  ``self.energy_history`` receives the field
        energy 1/2 sum E^2 dx of every step's mid-step field."""
        g = self.grid
        diag_counter = 0
        rho, phi = state.field.rho, state.field.phi
        f = state.variables.fn
        energy = torch.zeros(max(self.nbiter, 1), dtype=torch.float64, device=self.device)
        synchronize(self.device)
        for i in range(self.nbiter):
            if self.solver == 0:
                f, rho, phi, E = step_with_field(f, self.dt, g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv,
                                                 self.physics_mode)
            else:
                f, rho, phi = self._step_fn(f, self.dt, g.v, g.inv_kx, g.inv_kx_sq, g.x_indices, g.v_indices,
                                            g.x_batch_offsets, g.v_batch_offsets, g.stencil_offsets, g.dx, g.dv,
                                            g.V_max, self.physics_mode)
                # mid-step field from the returned rho (the step's own Poisson input)
                E, _ = poisson_from_rho(rho, g.inv_kx, g.inv_kx_sq) if self.physics_mode else (
                    torch.zeros(g.Nx, dtype=torch.float64, device=self.device), None)
            check(lib().vp_field_energy_f64(E.data_ptr(), None, g.Nx, g.dx, 1.0, energy.data_ptr() + 8 * i,
                                            stream_handle(self.device)), "vp_field_energy_f64")
            state.t += self.dt
            state.step += 1
            diag_counter += 1
            if diag_counter >= self.diag_steps:
                synchronize(self.device)
                self.save_diag(SimulationState(field=Field(rho, phi),
                                               variables=Variables(fn=f, f0=state.variables.f0),
                                               t=state.t, step=state.step))
                diag_counter = 0
        synchronize(self.device)
        self.energy_history = energy[:self.nbiter].cpu().numpy()
        return SimulationState(field=Field(rho, phi), variables=Variables(fn=f, f0=state.variables.f0),
                               t=state.t, step=state.step)


def main(argv=None) -> str:
    """The reference script's command line (P1 __main__): same flags, log
    lines and log file ``vlp1d_1v_<dtype>.txt``."""
    from . import cli
    a = cli.parse(argv, two_d=False)
    solver = Vlasov(Nx=a.nx, Nv=a.nv, L=a.lx, V_max=a.vmax, nbiter=a.nbiter, diag_steps=a.diag_steps, dt=a.dt,
                    out_dir=a.out_dir, physics_mode=a.physics_mode, epsilon=a.epsilon, dtype=a.dtype,
                    solver=a.solver)
    return cli.run(solver, f"vlp1d_1v_{a.dtype}.txt", (get_memory_stats, reset_peak_memory, format_memory))


if __name__ == "__main__":
    main()
