"""Allocation-free multi-step drivers.

``Run2D`` owns the state f, the field arrays and the workspace and advances
them step after step without allocating.  Two kernel paths:

* fused (float64: nx, ny, nvx, nvy in {32, 64, 128}; float32 storage: nx,
  ny, nvy in {64, 128}, nvx in {32, 64, 128}): one C-ABI call per kernel (so
  a caller can bracket any of them with events), drifts in place, the kick
  out of place into a second f-sized buffer, the closing X(dt/2) of a step
  and the opening X(dt/2) of the next run as one pass;
* generic (any shape): the six separate sweeps of the reference step,
  ping-ponging between two copies of f.

For float64 f both paths give the same f as repeated calls of
``vlasov_step_impl``.  float32 f is the FP32-storage variant: f stored and
interpolated in float32, charge density, Poisson solve, fields and field
energy in float64.  ``storage="deviation"`` (fused, nvy = 128) stores
g = f - feq in float32 instead, feq(vx, vy) = fx(vx) fy(vy) the separable
(rank-one) part of the (x, y) mean of the initial f with float64 profiles:
the drifts run on g, the kick on g in float32 plus the float64 correction
(Kvx fx)(Kvy fy) - fx fy, and rounding scales with the perturbation instead
of with f.
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import check, lib, stream_handle

FUSED_KERNELS = ("x_drift", "y_drift_rho", "poisson", "kick", "y_drift", "xx_drift", "field_energy")
GENERIC_KERNELS = ("x_drift", "y_drift", "rho", "poisson", "vx_kick", "vy_kick", "y_drift2", "x_drift2",
                   "field_energy")


def fused_supported(shape, dtype=torch.float64) -> bool:
    """True when the library has fused kernels for this f shape and dtype."""
    if len(shape) != 4:
        return False
    q = lib().vp_run_2d2v_f32_workspace if dtype == torch.float32 else lib().vp_run_2d2v_workspace
    return q(_lib.i64x4(shape)) > 0


def deviation_supported(shape) -> bool:
    """True when FP32 deviation storage runs for this f shape."""
    return len(shape) == 4 and int(shape[3]) == 128 and fused_supported(shape, torch.float32)


def _suffix(dtype) -> str:
    if dtype == torch.float64:
        return "f64"
    if dtype == torch.float32:
        return "f32"
    raise TypeError(f"f must be float64 or float32 (FP32 storage), got {dtype}")


class Run2D:
    """In-place 2D2V Strang stepping of f[Nx, Ny, Nvx, Nvy] (float64, or
    float32 for the FP32-storage variant)."""

    def __init__(self, grid, f: torch.Tensor, dt: float, physics_mode: bool = True, fused: bool | None = None,
                 storage: str = "plain"):
        self.g = grid
        self.dt = float(dt)
        self.physics_mode = bool(physics_mode)
        dev = f.device
        self.device = dev
        self.shape = tuple(int(n) for n in f.shape)
        if storage not in ("plain", "deviation"):
            raise ValueError(f"storage must be 'plain' or 'deviation', got {storage!r}")
        self.deviation = storage == "deviation"
        if self.deviation:
            if not deviation_supported(self.shape) or fused is False:
                raise ValueError(f"FP32 deviation storage needs fused float32 kernels and nvy = 128, got {self.shape}")
            self.sfx = "f32"
            self.fused = True
        else:
            self.sfx = _suffix(f.dtype)
            self.fused = fused_supported(self.shape, f.dtype) if fused is None else bool(fused)
            if self.fused and not fused_supported(self.shape, f.dtype):
                raise ValueError(f"no fused kernels for shape {self.shape} and {f.dtype}")
        Nx, Ny = self.shape[0], self.shape[1]
        if self.deviation:
            # split on the device: feq = fx fy from the (x, y) mean of f, g = float32(f - feq);
            # self.feq holds the float64 profiles fx (nvx values) then fy (nvy values)
            f64 = f.contiguous() if f.dtype == torch.float64 else f.contiguous().to(torch.float64)
            sh = _lib.i64x4(self.shape)
            self.feq = torch.empty(self.shape[2] + self.shape[3], dtype=torch.float64, device=dev)
            self.rho_shift = torch.empty(1, dtype=torch.float64, device=dev)
            g = torch.empty(self.shape, dtype=torch.float32, device=dev)
            sws = torch.empty(lib().vp_dev_split_workspace(sh), dtype=torch.uint8, device=dev)
            check(lib().vp_dev_split_f32(f64.data_ptr(), sh, float(grid.dvx), float(grid.dvy), self.feq.data_ptr(),
                                         self.rho_shift.data_ptr(), g.data_ptr(), sws.data_ptr(), sws.numel(),
                                         stream_handle(dev)), "vp_dev_split_f32")
            self.f = [g, torch.empty_like(g)]
        else:
            self.f = [f.contiguous().clone(), torch.empty_like(f)]
        self.fields = torch.zeros((4, Nx, Ny), dtype=torch.float64, device=dev)  # Ex, Ey, delta_rho, phi
        f64 = torch.float64
        self.vx = grid.vx.to(f64).contiguous()
        self.vy = grid.vy.to(f64).contiguous()
        self.kx = grid.kx.to(f64).contiguous()
        self.ky = grid.ky.to(f64).contiguous()
        self.iks = grid.inv_ksq.to(f64).contiguous()
        self.c_shape = _lib.i64x4(self.shape)
        if self.fused:
            # pieces workspace: the kick's second buffer is self.f[1]
            nbytes = lib().vp_fused_workspace(self.c_shape, 8 if self.sfx == "f64" else 4)
            self.ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            self._tables_ready = False
        else:
            self.rho = torch.empty((Nx, Ny), dtype=f64, device=dev)
            self.pws = torch.empty(lib().vp_poisson_2d_workspace(Nx, Ny), dtype=torch.uint8, device=dev)
            self.vs_x = _lib.i64x4((0, 0, 1, 0))
            self.vs_y = _lib.i64x4((0, 0, 0, 1))
            self.vs_e = _lib.i64x4((Ny, 1, 0, 0))
        self.steps_done = 0

    @property
    def state(self) -> torch.Tensor:
        """f: the stored tensor, or feq + g rounded to float32 (a new tensor) with
        deviation storage."""
        if self.deviation:
            out = torch.empty(self.shape, dtype=torch.float32, device=self.device)
            check(lib().vp_dev_join_f32(self.f[0].data_ptr(), self.feq.data_ptr(), _lib.i64x4(self.shape),
                                        out.data_ptr(), None, stream_handle(self.device)), "vp_dev_join_f32")
            return out
        return self.f[0]

    def state_f64(self) -> torch.Tensor:
        """f in float64 (feq + g exactly with deviation storage)."""
        if not self.deviation:
            return self.f[0].to(torch.float64)
        out = torch.empty(self.shape, dtype=torch.float64, device=self.device)
        check(lib().vp_dev_join_f32(self.f[0].data_ptr(), self.feq.data_ptr(), _lib.i64x4(self.shape), None,
                                    out.data_ptr(), stream_handle(self.device)), "vp_dev_join_f32")
        return out

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

    @staticmethod
    def _timed(events, name, fn):
        lst = None if events is None else events.get(name)
        if lst is None:
            fn()
            return
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        lst.append((e0, e1))

    def advance(self, nsteps: int, events: dict | None = None, energy: torch.Tensor | None = None) -> None:
        """Advance ``nsteps`` steps.  ``events`` maps kernel names (FUSED_KERNELS
        or GENERIC_KERNELS) to lists receiving (start, end) event pairs;
        ``energy`` (float64 [>= nsteps], device) receives the field energy of
        each step."""
        if nsteps <= 0:
            return
        if self.fused:
            self._advance_fused(nsteps, events, energy)
        else:
            for st in range(nsteps):
                self._step_generic(events, None if energy is None else energy[st:st + 1])
        self.steps_done += nsteps

    def _advance_fused(self, nsteps, events, energy):
        g = self.g
        s = stream_handle(self.device)
        L = lib()
        sfx = self.sfx
        tables = getattr(L, f"vp_fused_tables_{sfx}")
        xdrift = getattr(L, f"vp_fused_xdrift_{sfx}")
        ydrift = getattr(L, f"vp_fused_ydrift_{sfx}")
        if self.deviation:
            feq, rsh = self.feq.data_ptr(), self.rho_shift.data_ptr()

            def poisson(sh, dvx, dvy, kx, ky, iks, Ex, Ey, drho, phi, ws, wsn, s):
                return L.vp_fused_poisson_dev_f32(sh, dvx, dvy, rsh, kx, ky, iks, Ex, Ey, drho, phi, ws, wsn, s)

            def kick(fin, fout, sh, Ex, Ey, dt, dvx, dvy, s):
                return L.vp_fused_kick_dev_f32(fin, fout, sh, feq, Ex, Ey, dt, dvx, dvy, s)
        else:
            poisson = getattr(L, f"vp_fused_poisson_{sfx}")
            kick = getattr(L, f"vp_fused_kick_{sfx}")
        fp = self.f[0].data_ptr()
        fp2 = self.f[1].data_ptr()  # the kick writes here; the closing Y drift reads it back
        ws, wsn = self.ws.data_ptr(), self.ws.numel()
        Ex, Ey, drho, phi = (t.data_ptr() for t in self.fields)
        sh = self.c_shape
        if not self._tables_ready:
            check(tables(sh, self.vx.data_ptr(), self.vy.data_ptr(), self.dt / 2.0, g.dx, g.dy, ws, wsn, s),
                  "fused tables")
            self._tables_ready = True
        if not self.physics_mode:
            self.fields.zero_()
        t = self._timed
        t(events, "x_drift", lambda: check(xdrift(fp, fp, sh, 1, ws, wsn, s), "xdrift"))
        for st in range(nsteps):
            t(events, "y_drift_rho", lambda: check(ydrift(fp, fp, sh, int(self.physics_mode), ws, wsn, s),
                                                   "ydrift"))
            if self.physics_mode:
                t(events, "poisson", lambda: check(poisson(
                    sh, g.dvx, g.dvy, self.kx.data_ptr(), self.ky.data_ptr(), self.iks.data_ptr(), Ex, Ey, drho,
                    phi, ws, wsn, s), "poisson"))
            if energy is not None:
                ep = energy.data_ptr() + 8 * st
                t(events, "field_energy", lambda: check(L.vp_field_energy_f64(
                    Ex, Ey, self.shape[0] * self.shape[1], g.dx, g.dy, ep, s), "field_energy"))
            t(events, "kick", lambda: check(kick(fp, fp2, sh, Ex, Ey, self.dt, g.dvx, g.dvy, s), "kick"))
            t(events, "y_drift", lambda: check(ydrift(fp2, fp, sh, 0, ws, wsn, s), "ydrift"))
            last = st == nsteps - 1
            t(events, "x_drift" if last else "xx_drift",
              lambda: check(xdrift(fp, fp, sh, 1 if last else 2, ws, wsn, s), "xdrift"))

    def _adv(self, src, dst, axis, vel, vs, dt, dx, periodic, s):
        fn = getattr(lib(), f"vp_advect_{self.sfx}")
        check(fn(src.data_ptr(), dst.data_ptr(), self.c_shape, axis, vel.data_ptr(), vs, dt, dx, periodic, s),
              f"vp_advect_{self.sfx}")

    def _step_generic(self, events, energy):
        g = self.g
        s = stream_handle(self.device)
        a, b = self.f[0], self.f[1]
        half = self.dt / 2.0
        Nx, Ny, Nvx, Nvy = self.shape
        t = self._timed
        Ex, Ey, drho, phi = self.fields[0], self.fields[1], self.fields[2], self.fields[3]
        t(events, "x_drift", lambda: self._adv(a, b, 0, self.vx, self.vs_x, half, g.dx, 1, s))
        t(events, "y_drift", lambda: self._adv(b, a, 1, self.vy, self.vs_y, half, g.dy, 1, s))
        if self.physics_mode:
            row_sum = getattr(lib(), f"vp_row_sum_{self.sfx}")
            t(events, "rho", lambda: check(row_sum(a.data_ptr(), Nx * Ny, Nvx * Nvy, g.dvx, g.dvy, 0.0,
                                                   self.rho.data_ptr(), s), "vp_row_sum"))
            t(events, "poisson", lambda: check(lib().vp_poisson_2d_f64(
                self.rho.data_ptr(), Nx, Ny, self.kx.data_ptr(), self.ky.data_ptr(), self.iks.data_ptr(),
                Ex.data_ptr(), Ey.data_ptr(), drho.data_ptr(), phi.data_ptr(), self.pws.data_ptr(),
                self.pws.numel(), s), "vp_poisson_2d_f64"))
        else:
            self.fields.zero_()
        if energy is not None:
            t(events, "field_energy", lambda: check(lib().vp_field_energy_f64(
                Ex.data_ptr(), Ey.data_ptr(), Nx * Ny, g.dx, g.dy, energy.data_ptr(), s), "field_energy"))
        t(events, "vx_kick", lambda: self._adv(a, b, 2, Ex, self.vs_e, self.dt, g.dvx, 0, s))
        t(events, "vy_kick", lambda: self._adv(b, a, 3, Ey, self.vs_e, self.dt, g.dvy, 0, s))
        t(events, "y_drift2", lambda: self._adv(a, b, 1, self.vy, self.vs_y, half, g.dy, 1, s))
        t(events, "x_drift2", lambda: self._adv(b, a, 0, self.vx, self.vs_x, half, g.dx, 1, s))
