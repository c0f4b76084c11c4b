"""Write the committed regression fixtures of tests/golden/.

The reference ships no golden vectors (no .npy/.npz/.nc files) and its
PyTorch modules cannot be imported here (they import xarray, which this
image lacks), so these fixtures are written by the float64 oracle
(oracle/vlasov_oracle.py) after it has been pinned by the checks in
tests/test_oracle_*.py (closed forms, brute force, the Landau root).  They
anchor both the oracle (tests/test_golden.py re-derives them) and the GPU
path (tests/test_gpu_golden.py) on committed numbers.

    landau1d_128.npz   1D1V Landau damping, Nx = Nv = 128, L = 4 pi,
                       v in [-6, 6], f0 = (1 + 0.01 cos 0.5x) e^{-v^2/2}/sqrt(2 pi),
                       dt = 0.1, 600 steps: W_E[601] (entry 0 from f0's field,
                       entry n from step n's mid-step field), f after 10 and
                       after 600 steps (every other v column), and checksums.
    landau2d_32.npz    2D2V Landau, 32^4, L = 4 pi, v in [-6, 6],
                       f0 = (1 + 0.01 (cos 0.5x + cos 0.5y)) e^{-(vx^2+vy^2)/2}/(2 pi),
                       dt = 0.05, 8 steps: W_E[9], checksums after every
                       step, a strided sample of f after steps 2 and 8.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import math
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import vlasov_oracle as vo  # noqa: E402

HERE = pathlib.Path(__file__).resolve().parent

LANDAU1D = dict(Nx=128, Nv=128, L=4 * math.pi, V_max=6.0, eps=0.01, k=0.5, dt=0.1, steps=600)
LANDAU2D = dict(N=32, L=4 * math.pi, V_max=6.0, eps=0.01, k=0.5, dt=0.05, steps=8)
SAMPLE2D = (slice(None, None, 4), slice(1, None, 4), slice(None, None, 3), slice(2, None, 3))


def landau1d_initial():
    c = LANDAU1D
    g = vo.grid_1d(c["Nx"], c["Nv"], c["L"], c["V_max"])
    f0 = (1.0 + c["eps"] * np.cos(c["k"] * g.x[:, None])) * np.exp(-0.5 * g.v[None, :] ** 2) / math.sqrt(2 * math.pi)
    return g, f0


def landau1d(steps=None):
    c = LANDAU1D
    g, f = landau1d_initial()
    E0, _, _ = vo.solve_poisson_real_1d(f, g.inv_kx, g.inv_kx_sq, g.dv)
    W = [vo.field_energy_1d(E0, g.dx)]
    out = {}
    n = c["steps"] if steps is None else steps
    for s in range(1, n + 1):
        f, _, _, E = vo.vlasov_step_1d(f, c["dt"], g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv)
        W.append(vo.field_energy_1d(E, g.dx))
        if s == 10:
            out["f10"] = f[:, ::2].copy()
    out["W_E"] = np.array(W)
    out["f_last"] = f[:, ::2].copy()
    out["sum_last"] = np.array([f.sum(), (f * f).sum()])
    return out


def landau2d_initial():
    c = LANDAU2D
    N = c["N"]
    g = vo.grid_2d(N, N, N, N, c["L"], c["L"], c["V_max"], c["V_max"])
    x = g.x[:, None, None, None]
    y = g.y[None, :, None, None]
    vx = g.vx[None, None, :, None]
    vy = g.vy[None, None, None, :]
    f0 = (1.0 + c["eps"] * (np.cos(c["k"] * x) + np.cos(c["k"] * y))) * np.exp(-0.5 * (vx ** 2 + vy ** 2)) / (
        2 * math.pi)
    return g, f0


def landau2d(steps=None):
    c = LANDAU2D
    g, f = landau2d_initial()
    Ex, Ey, _, _ = vo.solve_poisson_real_2d(f, g.kx, g.ky, g.inv_ksq, g.dvx, g.dvy)
    W = [vo.field_energy_2d(Ex, Ey, g.dx, g.dy)]
    sums = []
    out = {}
    n = c["steps"] if steps is None else steps
    for s in range(1, n + 1):
        f, _, _, Ex, Ey = vo.vlasov_step_2d(f, c["dt"], g)
        W.append(vo.field_energy_2d(Ex, Ey, g.dx, g.dy))
        sums.append([f.sum(), (f * f).sum()])
        if s == 2:
            out["f2_sample"] = f[SAMPLE2D].copy()
    out["W_E"] = np.array(W)
    out["sums"] = np.array(sums)
    out["f_last_sample"] = f[SAMPLE2D].copy()
    return out


def main():
    np.savez_compressed(HERE / "landau1d_128.npz", **landau1d())
    np.savez_compressed(HERE / "landau2d_32.npz", **landau2d())
    for p in sorted(HERE.glob("*.npz")):
        print(p.name, p.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
