"""Batched 1D1V ensemble: many independent Vlasov-Poisson instances of
different sizes advanced together by the library's ensemble kernels
(``vp_ens_run_f64``), optionally replayed from a captured HIP graph.

No reference counterpart: the reference runs one simulation per process.
Each instance follows ``vlasov_step`` (vlasov1D1V.py:574-622) on its own
grid built by ``vlasov1D1V.Grid.create`` (vlasov1D1V.py:94-122), so the
instances equal independent single-instance runs (bit-identical sweeps;
the charge density is summed in another, compensated order).

``Ensemble.autotune()`` times the launch-shape candidates of the two
sweep kernels on a scratch copy of the pool and keeps the fastest.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, stream_handle
from .vlasov1D1V import Grid

_INST_DTYPE = np.dtype([("f_off", "<i8"), ("v_off", "<i8"), ("k_off", "<i8"), ("nx", "<i4"), ("nv", "<i4"),
                        ("dx", "<f8"), ("dv", "<f8")])
assert _INST_DTYPE.itemsize == 48

MAX_NX = 256
MAX_NV = 512


class Ensemble:
    """Instances given as dicts with ``nx``, ``nv``, ``L``, ``V_max`` and
    either ``f0`` (a [nx, nv] tensor) or ``k``, ``amplitude``, ``drift``
    (``configs.ensemble_f0``)."""

    def __init__(self, specs, dt: float, physics_mode: bool = True, device="cuda"):
        from . import configs

        self.device = torch.device(device)
        self.dt = float(dt)
        self.physics_mode = bool(physics_mode)
        self.specs = list(specs)
        n = len(self.specs)
        if n < 1:
            raise ValueError("empty ensemble")
        self.grids = []
        rec = np.zeros(n, dtype=_INST_DTYPE)
        f_off = 0
        t_off = 0
        tables = []
        for i, sp in enumerate(self.specs):
            nx, nv = int(sp["nx"]), int(sp["nv"])
            if not (6 <= nx <= MAX_NX and 6 <= nv <= MAX_NV):
                raise ValueError(f"instance {i}: nx={nx}, nv={nv} outside [6, {MAX_NX}] x [6, {MAX_NV}]")
            g = Grid.create(nx, nv, float(sp["L"]), float(sp["V_max"]), device=self.device, dtype=torch.float64)
            self.grids.append(g)
            nk = nx // 2 + 1
            rec[i] = (f_off, t_off, t_off + nv, nx, nv, g.dx, g.dv)
            tables += [g.v, g.inv_kx, g.inv_kx_sq]
            t_off += nv + 2 * nk
            f_off += nx * nv
        self.offsets = [int(r) for r in rec["f_off"]]
        self.shapes = [(int(r["nx"]), int(r["nv"])) for r in rec]
        self.tables = torch.cat(tables).contiguous()
        self.inst = torch.from_numpy(rec.view(np.uint8).copy()).to(self.device)
        self.pool = torch.empty(f_off, dtype=torch.float64, device=self.device)
        for i, sp in enumerate(self.specs):
            f0 = sp.get("f0")
            if f0 is None:
                f0 = configs.ensemble_f0(self.grids[i], sp["k"], sp["amplitude"], sp["drift"])
            self.state(i).copy_(f0)
        self.max_nx = max(s[0] for s in self.shapes)
        self.max_nv = max(s[1] for s in self.shapes)
        self.ws = torch.empty(lib().vp_ens_workspace(n, self.max_nx, self.max_nv), dtype=torch.uint8,
                              device=self.device)
        self._graphs = {}
        self.steps_done = 0

    @property
    def ninst(self) -> int:
        return len(self.specs)

    @property
    def cells(self) -> int:
        return int(self.pool.numel())

    def state(self, i: int) -> torch.Tensor:
        """View of instance i's f [nx, nv] inside the pool."""
        nx, nv = self.shapes[i]
        o = self.offsets[i]
        return self.pool[o:o + nx * nv].view(nx, nv)

    def _launch(self, nsteps: int, energy: torch.Tensor | None, pool: torch.Tensor | None = None) -> None:
        p = self.pool if pool is None else pool
        check(lib().vp_ens_run_f64(p.data_ptr(), self.inst.data_ptr(), self.ninst, self.max_nx, self.max_nv,
                                   self.tables.data_ptr(), int(nsteps), self.dt, int(self.physics_mode),
                                   None if energy is None else energy.data_ptr(), self.ws.data_ptr(),
                                   self.ws.numel(), stream_handle(self.device)), "vp_ens_run_f64")

    def run(self, nsteps: int, graph: bool = False) -> torch.Tensor:
        """Advance every instance by nsteps; returns the field-energy history
        [nsteps, ninst] (device).  graph=True captures the whole nsteps
        launch sequence once into a HIP graph (per nsteps) and replays it."""
        if nsteps <= 0:
            return torch.zeros((0, self.ninst), dtype=torch.float64, device=self.device)
        if not graph:
            energy = torch.zeros((nsteps, self.ninst), dtype=torch.float64, device=self.device)
            self._launch(nsteps, energy)
        else:
            cg, energy = self.prepare_graph(nsteps)
            cg.replay()
            energy = energy.clone()
        self.steps_done += nsteps
        return energy

    def prepare_graph(self, nsteps: int):
        """Capture (once per nsteps, without running it) the launch sequence
        of nsteps steps into a HIP graph; returns (graph, energy buffer)."""
        g = self._graphs.get(nsteps)
        if g is None:
            energy = torch.zeros((nsteps, self.ninst), dtype=torch.float64, device=self.device)
            # capture on a side stream (torch's rule); only the graph reads and writes the pool
            s = torch.cuda.Stream(device=self.device)
            s.wait_stream(torch.cuda.current_stream(self.device))
            cg = torch.cuda.CUDAGraph()
            with torch.cuda.graph(cg, stream=s):
                self._launch(nsteps, energy)
            torch.cuda.current_stream(self.device).wait_stream(s)
            g = (cg, energy)
            self._graphs[nsteps] = g
        return g

    def autotune(self, steps: int = 2, reps: int = 3) -> dict:
        """Pick the ensemble launch shapes (elements per X-drift and kick work
        item, persistent workgroups per CU) by timing each candidate on a copy
        of the pool; sets the library knobs and returns the choice with
        timings."""
        L = lib()
        K = _lib.KNOBS
        scratch = self.pool.clone()

        def timed():
            self._launch(steps, None, scratch)  # warm
            torch.cuda.synchronize(self.device)
            t0 = time.perf_counter()
            for _ in range(reps):
                self._launch(steps, None, scratch)
            torch.cuda.synchronize(self.device)
            return (time.perf_counter() - t0) / (reps * steps)

        results = {}
        best = None
        for xt in (2048, 4096):
            for kt in (2048, 4096):
                for w in (1, 2):
                    L.vp_tune_set(K["ens_xtile"], xt)
                    L.vp_tune_set(K["ens_ktile"], kt)
                    L.vp_tune_set(K["ens_wgs"], w)
                    t = timed()
                    results[f"xtile={xt},ktile={kt},wgs={w}"] = t
                    if best is None or t < best[0]:
                        best = (t, xt, kt, w)
        L.vp_tune_set(K["ens_xtile"], best[1])
        L.vp_tune_set(K["ens_ktile"], best[2])
        L.vp_tune_set(K["ens_wgs"], best[3])
        self._graphs.clear()  # captured graphs hold the old launch shapes
        # the workspace bound covers every tile size, so it stays valid
        return {"ens_xtile": best[1], "ens_ktile": best[2], "ens_wgs": best[3], "step_s": best[0],
                "candidates_s": results}
