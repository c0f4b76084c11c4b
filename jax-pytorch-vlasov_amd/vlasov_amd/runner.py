"""Allocation-free multi-step drivers.

``Run2D`` / ``Run1D`` own two ping-pong copies of f, the field arrays and
the workspace, and advance them in place step after step: no tensor is
allocated inside ``step()``.  Each step issues the same kernels as the
functional ``vlasov_step_impl`` / ``vlasov_step`` (bit-identical results),
one C-ABI call per kernel so a caller can bracket any of them with events.
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import check, lib, stream_handle

_KERNELS_2D = ("x_drift", "y_drift", "rho", "poisson", "vx_kick", "vy_kick", "y_drift2", "x_drift2")


class Run2D:
    """In-place 2D2V Strang stepping on f[Nx, Ny, Nvx, Nvy] (float64)."""

    def __init__(self, grid, f: torch.Tensor, dt: float, physics_mode: bool = True):
        self.g = grid
        self.dt = float(dt)
        self.physics_mode = bool(physics_mode)
        dev = f.device
        self.device = dev
        self.shape = tuple(f.shape)
        Nx, Ny = self.shape[0], self.shape[1]
        self.f = [f.contiguous().clone(), torch.empty_like(f)]
        self.cur = 0
        self.fields = torch.zeros((4, Nx, Ny), dtype=torch.float64, device=dev)  # Ex, Ey, drho, phi
        self.rho = torch.empty((Nx, Ny), dtype=torch.float64, device=dev)
        self.pws = torch.empty(lib().vp_poisson_2d_workspace(Nx, Ny), dtype=torch.uint8, device=dev)
        f64 = torch.float64
        self.vx = grid.vx.to(f64).contiguous()
        self.vy = grid.vy.to(f64).contiguous()
        self.kx = grid.kx.to(f64).contiguous()
        self.ky = grid.ky.to(f64).contiguous()
        self.iks = grid.inv_ksq.to(f64).contiguous()
        self.c_shape = _lib.i64x4(self.shape)
        self.vs_x = _lib.i64x4((0, 0, 1, 0))
        self.vs_y = _lib.i64x4((0, 0, 0, 1))
        self.vs_e = _lib.i64x4((Ny, 1, 0, 0))

    @property
    def state(self) -> torch.Tensor:
        return self.f[self.cur]

    @property
    def Ex(self):
        return self.fields[0]

    @property
    def Ey(self):
        return self.fields[1]

    @property
    def delta_rho(self):
        return self.fields[2]

    @property
    def phi(self):
        return self.fields[3]

    def _adv(self, src, dst, axis, vel, vs, dt, dx, periodic, s):
        check(lib().vp_advect_f64(src.data_ptr(), dst.data_ptr(), self.c_shape, axis, vel.data_ptr(), vs,
                                  dt, dx, periodic, s), "vp_advect_f64")

    def step(self, events: dict | None = None) -> None:
        """Advance one step.  ``events`` maps a kernel name in
        ("x_drift", "y_drift", "rho", "poisson", "vx_kick", "vy_kick", "y_drift2", "x_drift2")
        to a list that receives (start, end) torch.cuda.Event pairs."""
        g = self.g
        s = stream_handle(self.device)
        a, b = self.f[self.cur], self.f[1 - self.cur]
        half = self.dt / 2.0
        Nx, Ny, Nvx, Nvy = self.shape
        ev = events or {}

        def run(name, fn):
            lst = ev.get(name)
            if lst is None:
                fn()
                return
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            lst.append((e0, e1))

        run("x_drift", lambda: self._adv(a, b, 0, self.vx, self.vs_x, half, g.dx, 1, s))
        run("y_drift", lambda: self._adv(b, a, 1, self.vy, self.vs_y, half, g.dy, 1, s))
        Ex, Ey, drho, phi = self.fields[0], self.fields[1], self.fields[2], self.fields[3]
        if self.physics_mode:
            run("rho", lambda: check(lib().vp_row_sum_f64(a.data_ptr(), Nx * Ny, Nvx * Nvy, g.dvx, g.dvy, 0.0,
                                                          self.rho.data_ptr(), s), "vp_row_sum_f64"))
            run("poisson", lambda: check(lib().vp_poisson_2d_f64(
                self.rho.data_ptr(), Nx, Ny, self.kx.data_ptr(), self.ky.data_ptr(), self.iks.data_ptr(),
                Ex.data_ptr(), Ey.data_ptr(), drho.data_ptr(), phi.data_ptr(), self.pws.data_ptr(),
                self.pws.numel(), s), "vp_poisson_2d_f64"))
        run("vx_kick", lambda: self._adv(a, b, 2, Ex, self.vs_e, self.dt, g.dvx, 0, s))
        run("vy_kick", lambda: self._adv(b, a, 3, Ey, self.vs_e, self.dt, g.dvy, 0, s))
        run("y_drift2", lambda: self._adv(a, b, 1, self.vy, self.vs_y, half, g.dy, 1, s))
        run("x_drift2", lambda: self._adv(b, a, 0, self.vx, self.vs_x, half, g.dx, 1, s))
        # result is back in `a`
