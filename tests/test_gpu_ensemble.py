"""GPU parity of the batched 1D1V ensemble (``vp_ens_run_f64``).

Every instance must follow ``vlasov_step`` (vlasov1D1V.py:574-622) on its
own grid: compared with the float64 oracle stepping each instance alone
(<= 1e-12 relative in f and in the field-energy history; sweeps are
bit-identical, so with the field off the result equals the oracle exactly),
with the library's single-instance step, with itself replayed from a HIP
graph (several replays, eager work in between) and after the launch-shape
autotuner changed the kernels' shapes.
"""
import math

import numpy as np
import pytest
import torch

from oracle import vlasov_oracle as vo

pytestmark = pytest.mark.gpu

# ragged sizes: not multiples of the X items' column blocks or the kick's row
# blocks, the extremes of the supported range, nv below one column block
SIZES = [(32, 64), (37, 70), (100, 129), (64, 512), (256, 100), (48, 33), (6, 6), (129, 257)]


def _specs(sizes, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for nx, nv in sizes:
        k = float(rng.uniform(0.2, 0.6))
        u = float(rng.uniform(0.0, 3.0))
        out.append({"nx": nx, "nv": nv, "k": k, "L": 2 * math.pi / k, "amplitude": float(rng.uniform(0.01, 0.3)),
                    "drift": u, "V_max": max(6.0, u + 6.0)})
    return out


def _ens(physics_mode=True, sizes=SIZES, dt=0.1):
    from vlasov_amd.ensemble import Ensemble
    return Ensemble(_specs(sizes), dt, physics_mode=physics_mode)


def _oracle_run(ens, i, nsteps, physics_mode=True):
    g = ens.grids[i]
    f = ens.state(i).cpu().numpy().copy()
    v, ik, ik2 = (a.cpu().numpy() for a in (g.v, g.inv_kx, g.inv_kx_sq))
    W = []
    for _ in range(nsteps):
        f, _, _, E = vo.vlasov_step_1d(f, ens.dt, v, ik, ik2, g.dx, g.dv, physics_mode=physics_mode)
        W.append(vo.field_energy_1d(E, g.dx))
    return f, np.array(W)


def test_ensemble_matches_oracle(device):
    ens = _ens()
    nsteps = 12
    refs = [_oracle_run(ens, i, nsteps) for i in range(ens.ninst)]
    energy = ens.run(nsteps).cpu().numpy()
    for i, (f_ref, W_ref) in enumerate(refs):
        f = ens.state(i).cpu().numpy()
        err = np.max(np.abs(f - f_ref)) / np.max(np.abs(f_ref))
        assert err <= 1e-12, (i, ens.shapes[i], err)
        werr = np.max(np.abs(energy[:, i] - W_ref)) / np.max(W_ref)
        assert werr <= 1e-12, (i, ens.shapes[i], werr)


def test_ensemble_free_streaming_equals_oracle_exactly(device):
    ens = _ens(physics_mode=False)
    refs = [_oracle_run(ens, i, 5, physics_mode=False)[0] for i in range(ens.ninst)]
    energy = ens.run(5)
    assert torch.count_nonzero(energy) == 0
    for i, f_ref in enumerate(refs):
        np.testing.assert_array_equal(ens.state(i).cpu().numpy(), f_ref)


def test_ensemble_matches_single_instance_library_path(device):
    from vlasov_amd import vlasov1D1V as v1
    ens = _ens(sizes=SIZES[:4])
    singles = []
    for i in range(ens.ninst):
        g = ens.grids[i]
        f = ens.state(i).clone()
        for _ in range(8):
            f, _, _, _ = v1.step_with_field(f, ens.dt, g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv, True)
        singles.append(f)
    ens.run(8)
    for i, f in enumerate(singles):
        a, b = ens.state(i), f
        assert float((a - b).abs().max() / b.abs().max()) <= 1e-13


def test_ensemble_graph_replays_match_eager(device):
    a = _ens()
    b = _ens()
    ea = [a.run(4, graph=True) for _ in range(3)]
    tmp = torch.ones(1000, device=device).sum()  # eager work between replays
    ea.append(a.run(4, graph=True))
    eb = [b.run(4) for _ in range(4)]
    assert float(tmp) == 1000.0
    for x, y in zip(ea, eb):
        assert torch.equal(x, y)
    assert torch.equal(a.pool, b.pool)


def test_autotuned_shapes_keep_results(device):
    from vlasov_amd import _lib
    lib = _lib.lib()
    keys = ("ens_xtile", "ens_ktile", "ens_wgs")
    old = {k: lib.vp_tune_get(_lib.KNOBS[k]) for k in keys}
    try:
        a = _ens()
        choice = a.autotune(steps=1, reps=1)
        assert choice["ens_xtile"] in (2048, 4096) and choice["ens_ktile"] in (2048, 4096)
        assert choice["ens_wgs"] in (1, 2)
        a.run(6)
        for k in keys:
            lib.vp_tune_set(_lib.KNOBS[k], 0)
        b = _ens()
        b.run(6)
        # the X item width changes the charge-density summation order only
        assert float((a.pool - b.pool).abs().max() / b.pool.abs().max()) <= 1e-13
        for xt, kt, w in ((2048, 4096, 2), (4096, 2048, 1)):
            lib.vp_tune_set(_lib.KNOBS["ens_xtile"], xt)
            lib.vp_tune_set(_lib.KNOBS["ens_ktile"], kt)
            lib.vp_tune_set(_lib.KNOBS["ens_wgs"], w)
            c = _ens()
            c.run(6)
            assert float((c.pool - b.pool).abs().max() / b.pool.abs().max()) <= 1e-13
    finally:
        for k, v in old.items():
            lib.vp_tune_set(_lib.KNOBS[k], v)


def test_ensemble_rejects_bad_sizes(device):
    from vlasov_amd import _lib
    from vlasov_amd.ensemble import Ensemble
    with pytest.raises(ValueError):
        Ensemble(_specs([(300, 64)]), 0.1)
    with pytest.raises(ValueError):
        Ensemble(_specs([(64, 600)]), 0.1)
    assert _lib.lib().vp_ens_workspace(4, 257, 64) == 0
    assert _lib.lib().vp_ens_workspace(4, 64, 513) == 0
