"""The GPU path against the committed fixtures of tests/golden/ (float64
oracle runs, see tests/golden/make_golden.py): the 600-step 1D Landau W_E
history and f, and the 2D 32^4 run through the fused in-place runner and the
generic step, all to 1e-12 of the largest value."""
import math
import pathlib
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))

import make_golden as mg  # noqa: E402


def _load(name):
    with np.load(GOLDEN / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b))))


def test_landau1d_history_and_f(device):
    from vlasov_amd import vlasov1D1V as v1
    fx = _load("landau1d_128.npz")
    c = mg.LANDAU1D
    g = v1.Grid.create(c["Nx"], c["Nv"], c["L"], c["V_max"], device=device)
    _, f0 = mg.landau1d_initial()
    f = torch.from_numpy(f0).to(device)
    E0, _, _ = v1.solve_poisson_real(f, g.inv_kx, g.inv_kx_sq, g.dv)
    W = [v1.field_energy(E0, g.dx)]
    for s in range(1, c["steps"] + 1):
        f, _, _, E = v1.step_with_field(f, c["dt"], g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv, True)
        W.append(v1.field_energy(E, g.dx))
        if s == 10:
            assert _rel(f[:, ::2].cpu().numpy(), fx["f10"]) <= 1e-12
    W = torch.cat(W).cpu().numpy()
    assert _rel(W, fx["W_E"]) <= 1e-12
    assert _rel(f[:, ::2].cpu().numpy(), fx["f_last"]) <= 1e-12


@pytest.mark.parametrize("fused", [True, False])
def test_landau2d_runner(device, fused):
    from vlasov_amd import vlasov2D2V as v2
    from vlasov_amd.runner import Run2D
    fx = _load("landau2d_32.npz")
    c = mg.LANDAU2D
    N = c["N"]
    g = v2.Grid.create(N, N, N, N, c["L"], c["L"], c["V_max"], c["V_max"], device=device)
    _, f0 = mg.landau2d_initial()
    run = Run2D(g, torch.from_numpy(f0).to(device), c["dt"], True, fused=fused)
    energy = torch.zeros(c["steps"], dtype=torch.float64, device=device)
    run.advance(2, energy=energy[:2])
    assert _rel(run.state[mg.SAMPLE2D].cpu().numpy(), fx["f2_sample"]) <= 1e-12
    run.advance(c["steps"] - 2, energy=energy[2:])
    assert _rel(energy.cpu().numpy(), fx["W_E"][1:]) <= 1e-12
    assert _rel(run.state[mg.SAMPLE2D].cpu().numpy(), fx["f_last_sample"]) <= 1e-12
    f = run.state.cpu().numpy()
    np.testing.assert_allclose([f.sum(), (f * f).sum()], fx["sums"][-1], rtol=1e-12)
