"""2D2V Vlasov-Poisson solver: drop-in for the reference PyTorch module.

Mirrors the public interface of
``simulations/vlasov2d_2v/pytorch/src/vlasov2D2V.py`` (cited as P2 below):
``Grid`` / ``Field`` / ``Variables`` / ``SimulationState``, the kernel-level
functions ``advect_1d_vectorized``, ``advect_1d``, ``compute_rho``,
``compute_poisson(_real)``, ``solve_poisson(_real)``, the step
``vlasov_step_impl`` (same 22 positional arguments) and the ``Vlasov``
driver.  Everything runs in the HIP kernels of ``libvlasov_amd.so`` on a
ROCm GPU; there is no CPU path.
"""
from __future__ import annotations

import math
import pathlib
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, require_gpu, stream_handle
from .vlasov1D1V import (default_device, format_memory, get_device_label, get_memory_stats,  # noqa: F401
                         lagrange5_interp_sum, lagrange5_weights, reset_peak_memory, synchronize)


@dataclass
class Grid:
    """Grid of the 2D2V solver (P2:88-189)."""
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
    x: torch.Tensor
    y: torch.Tensor
    vx: torch.Tensor
    vy: torch.Tensor
    kx: torch.Tensor
    ky: torch.Tensor
    inv_ksq: torch.Tensor
    x_indices: torch.Tensor
    y_indices: torch.Tensor
    vx_indices: torch.Tensor
    vy_indices: torch.Tensor
    stencil_offsets: torch.Tensor

    @classmethod
    def create(cls, Nx, Ny, Nvx, Nvy, Lx, Ly, Vx_max, Vy_max, device="cuda", dtype=torch.float64):
        """Same torch expressions as P2:146-173, so identical arrays."""
        dx, dy = Lx / Nx, Ly / Ny
        dvx, dvy = 2 * Vx_max / (Nvx - 1), 2 * Vy_max / (Nvy - 1)
        x = torch.linspace(0, Lx, Nx + 1, device=device, dtype=dtype)[:-1]
        y = torch.linspace(0, Ly, Ny + 1, device=device, dtype=dtype)[:-1]
        vx = torch.linspace(-Vx_max, Vx_max, Nvx, device=device, dtype=dtype)
        vy = torch.linspace(-Vy_max, Vy_max, Nvy, device=device, dtype=dtype)
        kx = 2 * math.pi * torch.fft.fftfreq(Nx, d=dx, device=device, dtype=dtype).unsqueeze(1)
        ky = 2 * math.pi * torch.fft.rfftfreq(Ny, d=dy, device=device, dtype=dtype).unsqueeze(0)
        k_sq = kx ** 2 + ky ** 2
        inv_ksq = torch.where(k_sq == 0, torch.tensor(0.0, device=device, dtype=dtype), 1.0 / k_sq)
        x_indices = torch.arange(Nx, device=device, dtype=dtype)
        y_indices = torch.arange(Ny, device=device, dtype=dtype)
        vx_indices = torch.arange(Nvx, device=device, dtype=dtype)
        vy_indices = torch.arange(Nvy, device=device, dtype=dtype)
        stencil_offsets = torch.arange(-2, 4, device=device, dtype=torch.int64)
        return cls(Nx, Ny, Nvx, Nvy, Lx, Ly, Vx_max, Vy_max, dx, dy, dvx, dvy, x, y, vx, vy, kx, ky,
                   inv_ksq, x_indices, y_indices, vx_indices, vy_indices, stencil_offsets)

    def to(self, device) -> "Grid":
        for name in ("x", "y", "vx", "vy", "kx", "ky", "inv_ksq", "x_indices", "y_indices",
                     "vx_indices", "vy_indices", "stencil_offsets"):
            setattr(self, name, getattr(self, name).to(device))
        return self


@dataclass
class Field:
    """Field data (rho, phi)."""
    rho: torch.Tensor
    phi: torch.Tensor


@dataclass
class Variables:
    """Distribution function f and f0."""
    fn: torch.Tensor
    f0: torch.Tensor


@dataclass
class SimulationState:
    field: Field
    variables: Variables
    t: float
    step: int


def get_lagrange5_weights(alpha):
    """[..., 6] weights in the generic product form (P2:210-235); torch element-wise."""
    weights = torch.empty((*alpha.shape, 6), device=alpha.device, dtype=alpha.dtype)
    nodes = torch.arange(-2, 4, device=alpha.device, dtype=alpha.dtype)
    for j in range(6):
        w = torch.ones_like(alpha)
        for m in range(6):
            if m != j:
                w = w * (alpha - nodes[m]) / (nodes[j] - nodes[m])
        weights[..., j] = w
    return weights


# ---------------------------------------------------------------------------
# Sweeps
# ---------------------------------------------------------------------------

def _advect4(f, vel_f_strides, vel, axis, dt, dx, periodic, out=None):
    dev = require_gpu(f, vel)
    f_c = f.contiguous()
    if out is None:
        out = torch.empty_like(f_c)
    fn = lib().vp_advect_f64 if f_c.dtype == torch.float64 else lib().vp_advect_f32
    if f_c.dtype not in (torch.float64, torch.float32):
        raise TypeError(f"unsupported dtype {f_c.dtype}")
    rc = fn(f_c.data_ptr(), out.data_ptr(), _lib.i64x4(f_c.shape), int(axis), vel.data_ptr(),
            _lib.i64x4(vel_f_strides), float(dt), float(dx), int(bool(periodic)), stream_handle(dev))
    check(rc, "vp_advect")
    return out


def advect_1d(f_data, velocities, axis, dt, dx, periodic=True, point_indices=None, stencil_offsets=None):
    """Drop-in for advect_1d (P2:356-413): sweep of f(x, y, vx, vy) along ``axis``.

    ``velocities`` broadcasts to the batch shape (all dims except ``axis``),
    exactly as in the reference; the sweep runs in place of the
    permute/reshape/gather, reading each line once through LDS.
    """
    if f_data.ndim != 4:
        raise ValueError("advect_1d expects a 4D f(x, y, vx, vy)")
    axis = int(axis) % 4
    batch_shape = tuple(s for i, s in enumerate(f_data.shape) if i != axis)
    vel = velocities.to(device=f_data.device, dtype=torch.float64).expand(batch_shape)
    strides = list(vel.stride())
    strides.insert(axis, 0)
    return _advect4(f_data, strides, vel, axis, dt, dx, periodic)


def advect_1d_vectorized(f_batch, velocities, dt, dx, point_indices=None, stencil_offsets=None, periodic=True):
    """Drop-in for advect_1d_vectorized (P2:287-353) on rows [B, N]."""
    B, N = f_batch.shape
    vel = velocities.to(device=f_batch.device, dtype=torch.float64)
    if vel.ndim == 1:
        vel = vel.expand(B)
        strides = (0, 0, vel.stride(0), 0)
    else:
        vel = vel.expand(B, N)
        strides = (0, 0, vel.stride(0), vel.stride(1))
    return _advect4(f_batch.reshape(1, 1, B, N), strides, vel, 3, dt, dx, periodic).reshape(B, N)


# ---------------------------------------------------------------------------
# Field solve
# ---------------------------------------------------------------------------

def compute_rho(f_data, dvx, dvy):
    """Drop-in for compute_rho (P2:416-431), compensated float64 sum."""
    dev = require_gpu(f_data)
    f_c = f_data.contiguous()
    Nx, Ny, Nvx, Nvy = f_c.shape
    rho = torch.empty((Nx, Ny), dtype=torch.float64, device=dev)
    fn = lib().vp_row_sum_f64 if f_c.dtype == torch.float64 else lib().vp_row_sum_f32
    check(fn(f_c.data_ptr(), Nx * Ny, Nvx * Nvy, float(dvx), float(dvy), 0.0, rho.data_ptr(),
             stream_handle(dev)), "vp_row_sum")
    return rho


def compute_poisson_real(rho, kx, ky, inv_ksq):
    """Drop-in for compute_poisson_real (P2:488-517).  Returns (Ex, Ey, delta_rho, phi)."""
    dev = require_gpu(rho, kx, ky, inv_ksq)
    Nx, Ny = rho.shape
    outs = [torch.empty((Nx, Ny), dtype=torch.float64, device=dev) for _ in range(4)]
    Ex, Ey, drho, phi = outs
    ws = _lib.Workspace.get(dev, lib().vp_poisson_2d_workspace(Nx, Ny), "poisson2d")
    f64 = torch.float64
    rc = lib().vp_poisson_2d_f64(rho.to(f64).contiguous().data_ptr(), Nx, Ny,
                                 kx.to(f64).contiguous().data_ptr(), ky.to(f64).contiguous().data_ptr(),
                                 inv_ksq.to(f64).contiguous().data_ptr(), Ex.data_ptr(), Ey.data_ptr(),
                                 drho.data_ptr(), phi.data_ptr(), ws.data_ptr(), ws.numel(), stream_handle(dev))
    check(rc, "vp_poisson_2d")
    return Ex, Ey, drho, phi


def compute_poisson(rho, kx, ky, inv_ksq):
    """Drop-in for compute_poisson (P2:434-465): same results as the real form."""
    return compute_poisson_real(rho, kx, ky, inv_ksq)


def solve_poisson_real(f_data, kx, ky, inv_ksq, dvx, dvy):
    """Drop-in for solve_poisson_real (P2:520-523)."""
    return compute_poisson_real(compute_rho(f_data, dvx, dvy), kx, ky, inv_ksq)


def solve_poisson(f_data, kx, ky, inv_ksq, dvx, dvy):
    """Drop-in for solve_poisson (P2:468-485)."""
    return solve_poisson_real(f_data, kx, ky, inv_ksq, dvx, dvy)


# ---------------------------------------------------------------------------
# Step
# ---------------------------------------------------------------------------

def step_with_field(f, dt, grid_vx, grid_vy, grid_kx, grid_ky, grid_inv_ksq, grid_dx, grid_dy, grid_dvx,
                    grid_dvy, physics_mode, out=None):
    """One Strang step; returns (f, delta_rho, phi, Ex, Ey).  float32 f runs
    the FP32-storage variant (fields stay float64)."""
    dev = require_gpu(f, grid_vx, grid_vy)
    if f.dtype == torch.float32:
        # FP32 storage: one step of the in-place runner on a copy of f
        from types import SimpleNamespace

        from .runner import Run2D
        g = SimpleNamespace(vx=grid_vx, vy=grid_vy, kx=grid_kx, ky=grid_ky, inv_ksq=grid_inv_ksq, dx=float(grid_dx),
                            dy=float(grid_dy), dvx=float(grid_dvx), dvy=float(grid_dvy))
        run = Run2D(g, f, dt, bool(physics_mode))
        run.advance(1)
        res = run.state if out is None else out.copy_(run.state)
        return res, run.delta_rho.clone(), run.phi.clone(), run.Ex.clone(), run.Ey.clone()
    if f.dtype != torch.float64:
        raise TypeError(f"f must be float64 or float32 (FP32 storage), got {f.dtype}")
    f_c = f.contiguous()
    Nx, Ny = f_c.shape[0], f_c.shape[1]
    if out is None:
        out = torch.empty_like(f_c)
    fields = torch.empty((4, Nx, Ny), dtype=torch.float64, device=dev)
    Ex, Ey, drho, phi = fields[0], fields[1], fields[2], fields[3]
    shape = _lib.i64x4(f_c.shape)
    ws = _lib.Workspace.get(dev, lib().vp_step_2d2v_workspace(shape), "step2d")
    pm = bool(physics_mode)
    f64 = torch.float64
    kx = grid_kx.to(f64).contiguous() if pm else None
    ky = grid_ky.to(f64).contiguous() if pm else None
    iks = grid_inv_ksq.to(f64).contiguous() if pm else None
    rc = lib().vp_step_2d2v_f64(
        f_c.data_ptr(), out.data_ptr(), shape, float(dt), grid_vx.to(f64).contiguous().data_ptr(),
        grid_vy.to(f64).contiguous().data_ptr(), _lib.ptr(kx), _lib.ptr(ky), _lib.ptr(iks), float(grid_dx),
        float(grid_dy), float(grid_dvx), float(grid_dvy), int(pm), Ex.data_ptr(), Ey.data_ptr(),
        drho.data_ptr(), phi.data_ptr(), ws.data_ptr(), ws.numel(), stream_handle(dev))
    check(rc, "vp_step_2d2v")
    return out, drho, phi, Ex, Ey


def vlasov_step_impl(f, f0, dt, grid_x, grid_y, grid_vx, grid_vy, grid_kx, grid_ky, grid_inv_ksq,
                     grid_x_indices, grid_y_indices, grid_vx_indices, grid_vy_indices, grid_stencil_offsets,
                     grid_dx, grid_dy, grid_dvx, grid_dvy, grid_Nx, grid_Ny, physics_mode):
    """Drop-in for vlasov_step_impl (P2:943-998).  Returns (f, rho, phi)."""
    f, rho, phi, _, _ = step_with_field(f, dt, grid_vx, grid_vy, grid_kx, grid_ky, grid_inv_ksq, grid_dx,
                                        grid_dy, grid_dvx, grid_dvy, physics_mode)
    return f, rho, phi


# ---------------------------------------------------------------------------
# Solver 1 and solver 2 (shift forms, P2:526-941, P2:1001-1132)
# ---------------------------------------------------------------------------
from . import shift_forms as _sf  # noqa: E402

static_stencil_offsets = _sf.static_stencil_offsets
precompute_x_drift = _sf.precompute_x_drift
build_v_tap_weights = _sf.build_v_tap_weights
check_v_window = _sf.check_v_window
build_diag_weights = _sf.build_diag_weights


def advect_1d_matrix_diag_core(f_batch, velocities, dt, dx, stencil_offsets, periodic=True):
    """Drop-in for advect_1d_matrix_diag_core (P2:557-590): rows [B, N], one
    velocity per row ([B]), sweep along the last axis."""
    _sf._check_offsets(stencil_offsets)
    B, N = f_batch.shape
    vel = velocities.reshape(-1).to(torch.float64).expand(B)
    out = _sf.sweep_reflected(f_batch.reshape(1, 1, B, N), 3, vel, (0, 0, vel.stride(0), 0), dt, dx, periodic)
    return out.reshape(B, N)


def advect_1d_matrix_diag(f_data, velocities, axis, dt, dx, periodic=True, point_indices=None, stencil_offsets=None):
    """Drop-in for advect_1d_matrix_diag (P2:592-641): the solver-1 sweep of
    f(x, y, vx, vy) along ``axis`` (velocities broadcast to the batch
    shape), without the permute/reshape copies."""
    _sf._check_offsets(stencil_offsets)
    axis = int(axis) % 4
    batch_shape = tuple(n for i, n in enumerate(f_data.shape) if i != axis)
    vel = velocities.to(device=f_data.device, dtype=torch.float64).expand(batch_shape)
    strides = list(vel.stride())
    strides.insert(axis, 0)
    return _sf.sweep_reflected(f_data, axis, vel, strides, dt, dx, periodic)


def block_roll(f_data, blocks, axis, vel_axis):
    """Roll each velocity block by its static shift along ``axis`` (P2:717-754)."""
    if all(shift % f_data.shape[axis] == 0 for _, _, shift in blocks):
        return f_data
    strides = [0, 0, 0, 0]
    strides[vel_axis] = 1
    return _sf.static_sweep(f_data, axis, _sf.block_shifts(blocks, f_data.device),
                            _sf.roll_weights(f_data.shape[vel_axis], f_data.device), strides)


def advect_x_static(f_data, W, blocks, stencil_offsets, axis):
    """Drop-in for advect_x_static (P2:756-819): periodic drift along axis 0
    (velocity vx) or 1 (vy) with static shifts and weights W [6, Nv]."""
    _sf._check_offsets(stencil_offsets)
    strides = [0, 0, 0, 0]
    strides[axis + 2] = 1
    return _sf.static_sweep(f_data, axis, _sf.block_shifts(blocks, f_data.device), W, strides)


def advect_v_static(f_data, E, dt, dv, S, stencil_offsets, axis):
    """Drop-in for advect_v_static (P2:868-913): Dirichlet kick along axis 2
    (Ex) or 3 (Ey), exact while |E dt/dv| < S (taps outside the static window
    are dropped, as there)."""
    _sf._check_offsets(stencil_offsets)
    Ny = f_data.shape[1]
    return _sf.sweep_reflected(f_data, axis, E.to(torch.float64).contiguous(), (Ny, 1, 0, 0), dt, dv, False,
                               int(S))


def _fields(f, physics_mode, kx, ky, inv_ksq, dvx, dvy, Nx, Ny):
    if physics_mode:
        return solve_poisson_real(f, kx, ky, inv_ksq, dvx, dvy)
    z = torch.zeros((Nx, Ny), dtype=torch.float64, device=f.device)
    return z, z.clone(), z.clone(), z.clone()


def vlasov_step_matrix_diag(f, f0, dt, grid_x, grid_y, grid_vx, grid_vy, grid_kx, grid_ky, grid_inv_ksq,
                            grid_x_indices, grid_y_indices, grid_vx_indices, grid_vy_indices, grid_stencil_offsets,
                            grid_dx, grid_dy, grid_dvx, grid_dvy, grid_Nx, grid_Ny, physics_mode):
    """Drop-in for vlasov_step_matrix_diag (P2:1001-1040), the solver-1 step."""
    so = grid_stencil_offsets
    f = advect_1d_matrix_diag(f, grid_vx[None, :, None], 0, dt / 2.0, grid_dx, stencil_offsets=so, periodic=True)
    f = advect_1d_matrix_diag(f, grid_vy[None, None, :], 1, dt / 2.0, grid_dy, stencil_offsets=so, periodic=True)
    Ex, Ey, rho, phi = _fields(f, physics_mode, grid_kx, grid_ky, grid_inv_ksq, grid_dvx, grid_dvy, grid_Nx,
                               grid_Ny)
    f = advect_1d_matrix_diag(f, Ex[:, :, None], 2, dt, grid_dvx, stencil_offsets=so, periodic=False)
    f = advect_1d_matrix_diag(f, Ey[:, :, None], 3, dt, grid_dvy, stencil_offsets=so, periodic=False)
    f = advect_1d_matrix_diag(f, grid_vy[None, None, :], 1, dt / 2.0, grid_dy, stencil_offsets=so, periodic=True)
    f = advect_1d_matrix_diag(f, grid_vx[None, :, None], 0, dt / 2.0, grid_dx, stencil_offsets=so, periodic=True)
    return f, rho, phi


def vlasov_step_gather_free_impl(f, dt, W_x, W_y, blocks_x, blocks_y, stencil_offsets, grid_kx, grid_ky,
                                 grid_inv_ksq, grid_dvx, grid_dvy, grid_Nx, grid_Ny, physics_mode, S):
    """Drop-in for vlasov_step_gather_free_impl (P2:1043-1083), the solver-2 core."""
    f = advect_x_static(f, W_x, blocks_x, stencil_offsets, axis=0)
    f = advect_x_static(f, W_y, blocks_y, stencil_offsets, axis=1)
    Ex, Ey, rho, phi = _fields(f, physics_mode, grid_kx, grid_ky, grid_inv_ksq, grid_dvx, grid_dvy, grid_Nx,
                               grid_Ny)
    f = advect_v_static(f, Ex, dt, grid_dvx, S, stencil_offsets, axis=2)
    f = advect_v_static(f, Ey, dt, grid_dvy, S, stencil_offsets, axis=3)
    f = advect_x_static(f, W_y, blocks_y, stencil_offsets, axis=1)
    f = advect_x_static(f, W_x, blocks_x, stencil_offsets, axis=0)
    return f, rho, phi


def vlasov_step_gather_free(f, f0, dt, grid_x, grid_y, grid_vx, grid_vy, grid_kx, grid_ky, grid_inv_ksq,
                            grid_x_indices, grid_y_indices, grid_vx_indices, grid_vy_indices, grid_stencil_offsets,
                            grid_dx, grid_dy, grid_dvx, grid_dvy, grid_Nx, grid_Ny, physics_mode, S=1):
    """Drop-in for vlasov_step_gather_free (P2:1086-1132)."""
    blocks_x, W_x = precompute_x_drift(grid_vx, dt / 2.0, grid_dx, f.dtype)
    blocks_y, W_y = precompute_x_drift(grid_vy, dt / 2.0, grid_dy, f.dtype)
    return vlasov_step_gather_free_impl(f, dt, W_x, W_y, blocks_x, blocks_y,
                                        static_stencil_offsets(grid_stencil_offsets), grid_kx, grid_ky,
                                        grid_inv_ksq, grid_dvx, grid_dvy, grid_Nx, grid_Ny, physics_mode, S)


def moments(f, grid) -> torch.Tensor:
    """(mass, kinetic energy, L2 norm squared) of f [Nx, Ny, Nvx, Nvy]:
    sum f, 1/2 sum f (vx^2 + vy^2), sum f^2, each times dx dy dvx dvy
    (compensated float64 sums), a [3] float64 device tensor."""
    dev = require_gpu(f)
    f_c = f.contiguous()
    Nx, Ny, Nvx, Nvy = f_c.shape
    out = torch.empty(3, dtype=torch.float64, device=dev)
    ws = _lib.Workspace.get(dev, lib().vp_moments_workspace(), "moments")
    fn = lib().vp_moments_f64 if f_c.dtype == torch.float64 else lib().vp_moments_f32
    cell = grid.dx * grid.dy * grid.dvx * grid.dvy
    check(fn(f_c.data_ptr(), Nx * Ny, Nvx * Nvy, grid.vx.to(torch.float64).contiguous().data_ptr(),
             grid.vy.to(torch.float64).contiguous().data_ptr(), Nvy, float(cell), 1.0, out.data_ptr(), ws.data_ptr(),
             ws.numel(), stream_handle(dev)), "vp_moments")
    return out


# ---------------------------------------------------------------------------
# Driver (P2:1223-1435)
# ---------------------------------------------------------------------------

class Vlasov:
    dt: float = 0.05
    diag_steps: int = 10
    out_dir: str = "data"
    physics_mode: bool = False
    epsilon: float = 0.001
    dtype: torch.dtype = torch.float64

    def __init__(self, Nx=32, Ny=32, Nvx=32, Nvy=32, Lx=2.0 * math.pi, Ly=2.0 * math.pi, Vx_max=6.0,
                 Vy_max=6.0, nbiter=10, diag_steps=10, dt=0.05, out_dir="data", physics_mode=False,
                 epsilon=0.001, dtype: str = "float64", solver: int = 0, device=None,
                 fp32_storage: str = "deviation", snapshot_format: str = "npz"):
        """Same parameters as the reference constructor (P2:1236-1314).

        ``solver`` selects the step as the reference's ``solver_fns`` does:
        0 vlasov_step_impl, 1 vlasov_step_matrix_diag, 2
        vlasov_step_gather_free (all on the library's kernels).
        dtype='float32' is the FP32-storage variant: f is float32, the grid,
        charge density, Poisson solve and fields stay float64 (the
        reference's float32 mode computes everything in float32).
        ``fp32_storage`` (float32 only): "deviation" (the default) stores
        g = f - feq in float32, feq = fx(vx) fy(vy) the separable part of the
        (x, y) mean of the initial f with float64 profiles, so that rounding
        scales with the perturbation g instead of f (used when the shape has
        those kernels, nvy = 128; otherwise f itself is stored); "plain"
        stores f itself.
        """
        self.output_log = ""
        if dtype not in ("float64", "float32"):
            raise ValueError(f"dtype must be 'float64' or 'float32', got {dtype!r}")
        self.f_dtype = torch.float64 if dtype == "float64" else torch.float32
        if fp32_storage not in ("deviation", "plain"):
            raise ValueError(f"fp32_storage must be 'deviation' or 'plain', got {fp32_storage!r}")
        self.fp32_storage = fp32_storage
        self._f_init64 = None  # float64 initial f, split exactly by deviation storage
        self.dtype = torch.float64
        self.device = torch.device(device) if device is not None else default_device()
        self.grid = Grid.create(Nx, Ny, Nvx, Nvy, Lx, Ly, Vx_max, Vy_max, device=self.device, dtype=self.dtype)
        self.nbiter = nbiter
        self.diag_steps = diag_steps
        self.dt = dt
        self.out_dir = out_dir
        if snapshot_format not in ("npz", "bin"):
            raise ValueError(f"snapshot_format must be 'npz' or 'bin', got {snapshot_format!r}")
        self.snapshot_format = snapshot_format
        self.physics_mode = physics_mode
        self.epsilon = epsilon
        solver_fns = {0: vlasov_step_impl, 1: vlasov_step_matrix_diag, 2: vlasov_step_gather_free}
        solver_names = {0: "vlasov_step_impl", 1: "vlasov_step_matrix_diag", 2: "vlasov_step_gather_free"}
        self.solver = solver
        self.solver_type = solver_names[solver]
        self._step_fn = solver_fns[solver]
        if solver != 0 and self.f_dtype != torch.float64:
            raise ValueError("solvers 1 and 2 run float64 (the FP32-storage variant uses solver 0)")
        if solver == 2:
            # static drift data hoisted out of the loop, as in P2:1277-1303
            g = self.grid
            bx, Wx = precompute_x_drift(g.vx, self.dt / 2.0, g.dx, self.dtype)
            by, Wy = precompute_x_drift(g.vy, self.dt / 2.0, g.dy, self.dtype)
            offsets = static_stencil_offsets(g.stencil_offsets)

            def _step_gather_free(f, f0, dt, grid_x, grid_y, grid_vx, grid_vy, grid_kx, grid_ky, grid_inv_ksq,
                                  grid_x_indices, grid_y_indices, grid_vx_indices, grid_vy_indices,
                                  grid_stencil_offsets, grid_dx, grid_dy, grid_dvx, grid_dvy, grid_Nx, grid_Ny,
                                  physics_mode):
                return vlasov_step_gather_free_impl(f, dt, Wx, Wy, bx, by, offsets, grid_kx, grid_ky, grid_inv_ksq,
                                                    grid_dvx, grid_dvy, grid_Nx, grid_Ny, physics_mode, 1)

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
        self._append_log(f"Grid: {Nx}x{Ny}x{Nvx}x{Nvy}, Steps: {self.nbiter}, chunks: {num_steps}, "
                         f"diag_freq: {self.diag_steps}")

    def _append_log(self, message: str) -> None:
        self.output_log += f"{message}\n"

    def initialize(self) -> SimulationState:
        """Initial state of P2:1319-1355."""
        g = self.grid
        x_mesh = g.x[:, None, None, None]
        y_mesh = g.y[None, :, None, None]
        vx_mesh = g.vx[:, None]
        vy_mesh = g.vy[None, :]
        if self.physics_mode:
            kx_mode = 1 * 2.0 * math.pi / g.Lx
            ky_mode = 1 * 2.0 * math.pi / g.Ly
            energy = 0.5 * (vx_mesh ** 2 + vy_mesh ** 2)
            f0 = (1.0 / math.sqrt(2.0 * math.pi)) * torch.exp(-energy)
            f_data = (1.0 + self.epsilon * torch.cos(kx_mode * x_mesh) * torch.cos(ky_mode * y_mesh)) \
                * f0[None, None, :, :]
        else:
            x0, y0 = g.Lx / 2.0, g.Ly / 2.0
            sigma_x = g.Lx / 10.0
            sigma_v = g.Vx_max / 4.0
            dx2 = (x_mesh - x0) ** 2 / sigma_x ** 2
            dy2 = (y_mesh - y0) ** 2 / sigma_x ** 2
            vx2 = (vx_mesh[None, None, :, :] - 0.0) ** 2 / sigma_v ** 2
            vy2 = (vy_mesh[None, None, :, :] - 0.0) ** 2 / sigma_v ** 2
            f_data = torch.exp(-0.5 * (dx2 + dy2 + vx2 + vy2))
            f0 = torch.zeros((g.Nvx, g.Nvy), device=self.device, dtype=self.dtype)
        if self.f_dtype == torch.float32:
            self._f_init64 = f_data.to(torch.float64).contiguous()
        f_data = f_data.to(self.f_dtype).contiguous()
        *_, rho, phi = solve_poisson_real(f_data, g.kx, g.ky, g.inv_ksq, g.dvx, g.dvy)
        state = SimulationState(field=Field(rho, phi), variables=Variables(fn=f_data, f0=f0), t=0.0, step=0)
        self.save_diag(state)
        return state

    def save_diag(self, state: SimulationState) -> None:
        """Snapshot to ``field_%06d.npz`` (the reference writes netCDF via xarray),
        or with snapshot_format='bin' to the Kokkos solver's raw
        ``fn/f0/rho/phi_%06d.bin`` files; records the moments (mass, kinetic
        energy, L2) of f."""
        g = self.grid
        mom = moments(state.variables.fn, g).cpu().numpy()
        self.moment_history.append((int(state.step), float(state.t), *[float(m) for m in mom]))
        if self.snapshot_format == "bin":
            from . import snapshots
            snapshots.write_snapshot(self.out_dir, state.step, state.variables.fn, state.variables.f0,
                                     snapshots.kokkos_rho(state.field.rho), state.field.phi)
            return
        filename = pathlib.Path(self.out_dir) / f"field_{int(state.step):06d}.npz"
        np.savez(filename, f=state.variables.fn.detach().cpu().numpy(), f0=state.variables.f0.detach().cpu().numpy(),
                 rho=state.field.rho.detach().cpu().numpy(), phi=state.field.phi.detach().cpu().numpy(),
                 time=np.array([float(state.t)]), x=g.x.cpu().numpy(), y=g.y.cpu().numpy(),
                 vx=g.vx.cpu().numpy(), vy=g.vy.cpu().numpy(), step=int(state.step), dt=self.dt,
                 Lx=g.Lx, Ly=g.Ly, Vx_max=g.Vx_max, Vy_max=g.Vy_max, physics_mode=int(self.physics_mode),
                 moments=mom)

    def restart(self, step: int) -> SimulationState:
        """State of the raw binary snapshot of iteration ``step`` in out_dir
        (written with snapshot_format='bin', or by the reference's Kokkos
        solver); run() continues from it.  f keeps the file's precision
        (float32 files give the FP32-storage variant's f)."""
        from . import snapshots
        g = self.grid
        a = snapshots.read_snapshot(self.out_dir, step, {"fn": (g.Nx, g.Ny, g.Nvx, g.Nvy), "f0": (g.Nvx, g.Nvy),
                                                         "rho": (g.Nx, g.Ny), "phi": (g.Nx, g.Ny)})
        fn = torch.from_numpy(a["fn"]).to(self.device).to(self.f_dtype)
        t = {k: torch.from_numpy(a[k].astype(np.float64)).to(self.device) for k in ("f0", "rho", "phi")}
        t["rho"] = snapshots.delta_rho_of(t["rho"])  # the files hold the full density
        return SimulationState(field=Field(t["rho"], t["phi"]), variables=Variables(fn=fn, f0=t["f0"]),
                               t=float(step) * self.dt, step=int(step))

    def run(self, state: SimulationState) -> SimulationState:
        """Time loop of P2:1392-1435.

        Same steps, diagnostics cadence and returned state as the reference.
        The steps between two snapshots run in place through
        ``runner.Run2D`` (fused kernels where the shape has them, else the
        generic sweeps; both equal repeated vlasov_step_impl calls).
        ``self.energy_history`` receives the field energy
        1/2 sum(Ex^2+Ey^2) dx dy of every step (mid-step field).
        """
        from .runner import Run2D

        g = self.grid
        rho, phi = state.field.rho, state.field.phi
        f = state.variables.fn
        energy = torch.zeros(max(self.nbiter, 1), dtype=torch.float64, device=self.device)
        synchronize(self.device)
        if self.solver != 0:
            return self._run_steps(state, energy)
        from .runner import deviation_supported
        if (self.f_dtype == torch.float32 and self.fp32_storage == "deviation"
                and deviation_supported(tuple(f.shape))):
            # split the exact float64 initial f when this run starts from it
            f_split = self._f_init64 if (state.step == 0 and self._f_init64 is not None) else f
            runner = Run2D(g, f_split, self.dt, self.physics_mode, storage="deviation")
        else:
            runner = Run2D(g, f, self.dt, self.physics_mode)
        done = 0
        while done < self.nbiter:
            chunk = min(self.diag_steps, self.nbiter - done)
            runner.advance(chunk, energy=energy[done:done + chunk])
            done += chunk
            for _ in range(chunk):
                state.t += self.dt
            state.step += chunk
            if chunk == self.diag_steps:
                synchronize(self.device)
                self.save_diag(SimulationState(field=Field(runner.delta_rho.clone(), runner.phi.clone()),
                                               variables=Variables(fn=runner.state, f0=state.variables.f0),
                                               t=state.t, step=state.step))
        f = runner.state
        if self.nbiter > 0:
            rho, phi = runner.delta_rho.clone(), runner.phi.clone()
        synchronize(self.device)
        self.energy_history = energy[:self.nbiter].cpu().numpy()
        return SimulationState(field=Field(rho, phi), variables=Variables(fn=f, f0=state.variables.f0),
                               t=state.t, step=state.step)


def _vlasov_run_steps(self, state, energy):
    """Step-by-step loop of P2:1392-1435 through ``self._step_fn`` (solvers 1, 2)."""
    g = self.grid
    rho, phi = state.field.rho, state.field.phi
    f = state.variables.fn
    diag_counter = 0
    for i in range(self.nbiter):
        f, rho, phi = self._step_fn(f, state.variables.f0, self.dt, g.x, g.y, g.vx, g.vy, g.kx, g.ky, g.inv_ksq,
                                    g.x_indices, g.y_indices, g.vx_indices, g.vy_indices, g.stencil_offsets, g.dx,
                                    g.dy, g.dvx, g.dvy, g.Nx, g.Ny, self.physics_mode)
        if self.physics_mode:
            # mid-step field from the step's own charge density (rho returned = delta_rho)
            Ex, Ey, _, _ = compute_poisson_real(rho + 1.0, g.kx, g.ky, g.inv_ksq)
            check(lib().vp_field_energy_f64(Ex.data_ptr(), Ey.data_ptr(), g.Nx * g.Ny, g.dx, g.dy,
                                            energy.data_ptr() + 8 * i, stream_handle(self.device)),
                  "vp_field_energy_f64")
        state.t += self.dt
        state.step += 1
        diag_counter += 1
        if diag_counter >= self.diag_steps:
            synchronize(self.device)
            self.save_diag(SimulationState(field=Field(rho, phi), variables=Variables(fn=f, f0=state.variables.f0),
                                           t=state.t, step=state.step))
            diag_counter = 0
    synchronize(self.device)
    self.energy_history = energy[:self.nbiter].cpu().numpy()
    return SimulationState(field=Field(rho, phi), variables=Variables(fn=f, f0=state.variables.f0), t=state.t,
                           step=state.step)


Vlasov._run_steps = _vlasov_run_steps


def main(argv=None) -> str:
    """The reference script's command line (P2 __main__): same flags, log
    lines and log file ``vlp2d_2v_<dtype>.txt``."""
    from . import cli
    a = cli.parse(argv, two_d=True)
    solver = Vlasov(Nx=a.nx, Ny=a.ny, Nvx=a.nvx, Nvy=a.nvy, Lx=a.lx, Ly=a.ly, Vx_max=a.Vx_max, Vy_max=a.Vy_max,
                    nbiter=a.nbiter, diag_steps=a.diag_steps, dt=a.dt, out_dir=a.out_dir,
                    physics_mode=a.physics_mode, epsilon=a.epsilon, dtype=a.dtype, solver=a.solver)
    return cli.run(solver, f"vlp2d_2v_{a.dtype}.txt", (get_memory_stats, reset_peak_memory, format_memory))


__all__ = [name for name in dir() if not name.startswith("_")]
_ = format_memory

if __name__ == "__main__":
    main()
