"""GPU parity of the fused in-place 2D2V kernels (every extent in {32, 64, 128}).

The fused X/Y drifts use the same exactly-rounded weights as the generic
sweeps (precomputed in a table) and the same tap order, so with E = 0 they
must be bit-identical to the oracle; with the field on, rho is summed in a
different (compensated) order, so results agree to <= 1e-13 per step and
<= 1e-12 over trajectories.
"""
import math

import numpy as np
import pytest
import torch

from oracle import vlasov_oracle as vo

pytestmark = pytest.mark.gpu


def _mods():
    from vlasov_amd import _lib, runner
    from vlasov_amd import vlasov2D2V as v2
    return v2, runner, _lib


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


def _grid(v2, shape, device, L=4 * math.pi, V=6.0):
    nx, ny, nvx, nvy = shape
    G = v2.Grid.create(nx, ny, nvx, nvy, L, 0.75 * L, V, 0.8 * V, device=device, dtype=torch.float64)
    g = vo.grid_2d(nx, ny, nvx, nvy, L, 0.75 * L, V, 0.8 * V)
    arrays = tuple(a.cpu().numpy() for a in (G.vx, G.vy, G.kx, G.ky, G.inv_ksq))
    return G, g, arrays


def _landau_like(shape, seed):
    nx, ny, nvx, nvy = shape
    rng = np.random.default_rng(seed)
    x = np.arange(nx) / nx * 2 * math.pi
    y = np.arange(ny) / ny * 2 * math.pi
    vx = np.linspace(-6, 6, nvx)
    vy = np.linspace(-4.8, 4.8, nvy)
    pert = 1 + 0.05 * (np.cos(x)[:, None] + np.sin(2 * y)[None, :])
    fv = np.exp(-0.5 * (vx[:, None] ** 2 + vy[None, :] ** 2)) / (2 * math.pi)
    return pert[:, :, None, None] * fv[None, None] * (1 + 0.01 * rng.random(shape))


@pytest.mark.parametrize("shape", [(32, 32, 32, 32), (64, 32, 32, 64), (32, 64, 128, 32)])
@pytest.mark.parametrize("physics_mode", [False, True])
def test_fused_step_matches_oracle(device, shape, physics_mode):
    v2, runner, _ = _mods()
    assert runner.fused_supported(shape)
    G, g, arrays = _grid(v2, shape, device)
    f0 = _landau_like(shape, 1)
    ref = vo.vlasov_step_2d(f0, 0.1, g, physics_mode=physics_mode, grid_arrays=arrays)
    out, drho, phi, Ex, Ey = v2.step_with_field(torch.from_numpy(f0).to(device), 0.1, G.vx, G.vy, G.kx, G.ky,
                                                G.inv_ksq, G.dx, G.dy, G.dvx, G.dvy, physics_mode)
    if physics_mode:
        assert rel(out.cpu().numpy(), ref[0]) <= 1e-13
        for o, r in zip((drho, phi, Ex, Ey), ref[1:]):
            assert rel(o.cpu().numpy(), r) <= 1e-13
    else:
        np.testing.assert_array_equal(out.cpu().numpy(), ref[0])


@pytest.mark.parametrize("physics_mode", [False, True])
def test_fused_run_matches_generic_steps(device, physics_mode):
    """Run2D fused (X halves merged across steps) == generic step-by-step."""
    v2, runner, _ = _mods()
    shape = (64, 64, 64, 64)
    G, g, arrays = _grid(v2, shape, device)
    f0 = torch.from_numpy(_landau_like(shape, 2)).to(device)
    a = runner.Run2D(G, f0, 0.1, physics_mode, fused=True)
    b = runner.Run2D(G, f0, 0.1, physics_mode, fused=False)
    ea = torch.zeros(6, dtype=torch.float64, device=device)
    eb = torch.zeros(6, dtype=torch.float64, device=device)
    a.advance(2, energy=ea[:2])
    a.advance(4, energy=ea[2:])
    b.advance(6, energy=eb)
    fa, fb = a.state.cpu().numpy(), b.state.cpu().numpy()
    if physics_mode:
        assert np.linalg.norm(fa - fb) / np.linalg.norm(fb) <= 1e-13
        assert rel(ea.cpu().numpy(), eb.cpu().numpy()) <= 1e-12
    else:
        np.testing.assert_array_equal(fa, fb)


def test_fused_xdrift_two_passes_equal_two_calls(device):
    v2, runner, L = _mods()
    shape = (128, 32, 64, 64)
    G, g, arrays = _grid(v2, shape, device)
    f = torch.rand(shape, dtype=torch.float64, device=device)
    sh = L.i64x4(shape)
    lib = L.lib()
    ws = torch.empty(lib.vp_fused_workspace(sh, 8), dtype=torch.uint8, device=device)
    s = torch.cuda.current_stream().cuda_stream
    L.check(lib.vp_fused_tables_f64(sh, G.vx.data_ptr(), G.vy.data_ptr(), 0.05, G.dx, G.dy, ws.data_ptr(),
                                    ws.numel(), s), "tables")
    one = f.clone()
    L.check(lib.vp_fused_xdrift_f64(one.data_ptr(), one.data_ptr(), sh, 1, ws.data_ptr(), ws.numel(), s), "x1")
    L.check(lib.vp_fused_xdrift_f64(one.data_ptr(), one.data_ptr(), sh, 1, ws.data_ptr(), ws.numel(), s), "x1")
    two = torch.empty_like(f)
    L.check(lib.vp_fused_xdrift_f64(f.data_ptr(), two.data_ptr(), sh, 2, ws.data_ptr(), ws.numel(), s), "x2")
    assert torch.equal(one, two)
    ref = vo.advect_axis(f.cpu().numpy(), G.vx.cpu().numpy()[None, :, None], 0, 0.05, G.dx, True)
    ref = vo.advect_axis(ref, G.vx.cpu().numpy()[None, :, None], 0, 0.05, G.dx, True)
    np.testing.assert_array_equal(two.cpu().numpy(), ref)


@pytest.mark.parametrize("knobs", [{"x_split": 1, "y_split": 1, "kick_prefetch": 0},
                                   {"x_split": 16, "y_split": 4, "kick_prefetch": 1},
                                   {"drift_dma": 1}, {"drift_dma": 2, "nontemporal": 8},
                                   {"drift_dma": 3, "x_split": 3, "kick_prefetch": 1, "nontemporal": 7}])
def test_launch_knobs_do_not_change_results(device, knobs):
    v2, runner, L = _mods()
    lib = L.lib()
    shape = (64, 64, 128, 128)
    G, g, arrays = _grid(v2, shape, device)
    f0 = torch.from_numpy(_landau_like(shape, 3)).to(device)
    base = runner.Run2D(G, f0, 0.05, True)
    base.advance(3)
    old = {k: lib.vp_tune_get(L.KNOBS[k]) for k in knobs}
    try:
        for k, v in knobs.items():
            assert lib.vp_tune_set(L.KNOBS[k], v) == 0
        other = runner.Run2D(G, f0, 0.05, True)
        other.advance(3)
    finally:
        for k, v in old.items():
            lib.vp_tune_set(L.KNOBS[k], v)
    assert torch.equal(base.state, other.state)


def test_vlasov2d_driver_fused_landau(device, tmp_path):
    """Vlasov driver at 32^4 (fused path): 20 steps vs the oracle, f and W_E."""
    v2, runner, _ = _mods()
    s = v2.Vlasov(Nx=32, Ny=32, Nvx=32, Nvy=32, Lx=4 * math.pi, Ly=4 * math.pi, Vx_max=6.0, Vy_max=6.0, nbiter=20,
                  diag_steps=10, dt=0.1, physics_mode=True, epsilon=0.05, out_dir=str(tmp_path), device=device)
    st = s.initialize()
    f_ref = st.variables.fn.cpu().numpy()
    g = vo.grid_2d(32, 32, 32, 32, 4 * math.pi, 4 * math.pi, 6.0, 6.0)
    arrays = tuple(a.cpu().numpy() for a in (s.grid.vx, s.grid.vy, s.grid.kx, s.grid.ky, s.grid.inv_ksq))
    st = s.run(st)
    W = []
    for _ in range(20):
        f_ref, drho, phi, Ex, Ey = vo.vlasov_step_2d(f_ref, 0.1, g, True, grid_arrays=arrays)
        W.append(vo.field_energy_2d(Ex, Ey, g.dx, g.dy))
    assert np.linalg.norm(st.variables.fn.cpu().numpy() - f_ref) / np.linalg.norm(f_ref) <= 1e-12
    assert np.max(np.abs(s.energy_history - np.array(W))) / max(W) <= 1e-12
    assert rel(st.field.phi.cpu().numpy(), phi) <= 1e-11
    assert st.step == 20 and (tmp_path / "field_000010.npz").exists() and (tmp_path / "field_000020.npz").exists()


@pytest.mark.parametrize("shape", [(8, 4, 64, 128), (4, 4, 128, 128), (4, 8, 32, 128), (8, 8, 128, 64),
                                   (8, 4, 64, 64), (4, 4, 32, 64)])
def test_kick_stream_and_plane_kernels_agree_on_edge_fields(device, shape):
    """Out-of-place kicks (row-split kernel at 128 or 64 vy columns with 2, 4,
    8 and 16 waves per plane and 4 or 8 prefetched rows, two-wave
    row-streaming kernel), the in-place whole-plane
    LDS kick and the oracle agree bit for bit on fields that exercise every
    tap path: E = 0, tiny E (rounding makes floor(r - d) - r non-constant:
    displacements of 1e-18, 3e-17 and 8e-17 cells move the first one or two
    rows of a plane, which the windowed walks leave to the per-row path;
    1e-15 cells moves seventeen, and the whole plane takes the per-row path),
    multi-cell and fully out-of-domain shifts, both signs."""
    v2, runner, L = _mods()
    lib = L.lib()
    nx, ny, nvx, nvy = shape
    rng = np.random.default_rng(11)
    f = rng.random(shape)
    dt, dvx, dvy = 0.1, 12.0 / (nvx - 1), 12.0 / (nvy - 1)
    choices = np.array([0.0, 1e-18, -1e-18, 3e-17 * dvx / dt, 8e-17 * dvx / dt, 1e-15 * dvx / dt, 5.3 * dvx / dt,
                        -5.3 * dvx / dt, 0.4 * dvx / dt, -0.7 * dvx / dt, 80.0 * dvx / dt, -80.0 * dvx / dt,
                        2.0 * dvx / dt])
    Ex = rng.choice(choices, size=(nx, ny))
    Ey = rng.choice(choices, size=(nx, ny)) * (dvy / dvx)
    ref = vo.advect_axis(f, Ex[:, :, None], 2, dt, dvx, periodic=False)
    ref = vo.advect_axis(ref, Ey[:, :, None], 3, dt, dvy, periodic=False)
    tf = torch.from_numpy(f).to(device)
    tEx, tEy = torch.from_numpy(Ex).to(device), torch.from_numpy(Ey).to(device)
    out = torch.empty_like(tf)
    s = torch.cuda.current_stream().cuda_stream
    sh = L.i64x4(shape)
    L.check(lib.vp_fused_kick_f64(tf.data_ptr(), out.data_ptr(), sh, tEx.data_ptr(), tEy.data_ptr(), dt, dvx, dvy,
                                  s), "kick out of place")
    outs = []
    old = {k: lib.vp_tune_get(L.KNOBS[k]) for k in ("kick_kernel", "kick_prefetch")}
    try:
        for kk, pf in ((0, 1), (0, 3), (2, 2), (2, 1), (3, 3), (4, 1), (5, 3), (5, 1), (6, 3), (6, 1)):
            lib.vp_tune_set(L.KNOBS["kick_kernel"], kk)
            lib.vp_tune_set(L.KNOBS["kick_prefetch"], pf)
            o = torch.empty_like(tf)
            L.check(lib.vp_fused_kick_f64(tf.data_ptr(), o.data_ptr(), sh, tEx.data_ptr(), tEy.data_ptr(), dt, dvx,
                                          dvy, s), "kick out of place")
            outs.append(o)
    finally:
        for k, v in old.items():
            lib.vp_tune_set(L.KNOBS[k], v)
    for o in outs:
        assert torch.equal(o, out)
    inplace = tf.clone()
    L.check(lib.vp_fused_kick_f64(inplace.data_ptr(), inplace.data_ptr(), sh, tEx.data_ptr(), tEy.data_ptr(), dt,
                                  dvx, dvy, s), "kick in place")
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    np.testing.assert_array_equal(inplace.cpu().numpy(), ref)
