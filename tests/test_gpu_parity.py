"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Tolerances: the sweeps follow the reference's element-wise operation order
with no fused multiply-add, so they must agree to <= 1e-14 relative (and are
expected to be bit-identical, checked separately); reductions and FFTs
round differently (compensated sums, exact-twiddle DFTs), so field results
agree to <= 1e-13 of their max magnitude and multi-step trajectories to
<= 1e-12 relative (the parity bar for the float64 path).
"""
import math

import numpy as np
import pytest
import torch

from oracle import vlasov_oracle as vo

pytestmark = pytest.mark.gpu


def _v1():
    from vlasov_amd import vlasov1D1V
    return vlasov1D1V


def _v2():
    from vlasov_amd import vlasov2D2V
    return vlasov2D2V


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("axis", [0, 1, 2, 3])
@pytest.mark.parametrize("shift_factor", [0.3, 1.0, 2.7, 6.0])
def test_advect_1d_axis_matches_oracle(device, axis, shift_factor):
    v2 = _v2()
    rng = np.random.default_rng(10 + axis)
    shape = (16, 12, 20, 24)
    f = rng.random(shape)
    batch = tuple(s for i, s in enumerate(shape) if i != axis)
    n = shape[axis]
    dx = 0.37
    dt = 0.1
    vel = rng.uniform(-1, 1, size=batch) * shift_factor * dx / dt
    periodic = axis < 2
    ref = vo.advect_axis(f, vel, axis, dt, dx, periodic=periodic)
    out = v2.advect_1d(torch.from_numpy(f).to(device), torch.from_numpy(vel).to(device), axis, dt, dx,
                       periodic=periodic).cpu().numpy()
    assert rel(out, ref) <= 1e-14, (axis, shift_factor, rel(out, ref), n)


@pytest.mark.parametrize("axis", [0, 1, 2, 3])
def test_advect_broadcast_velocities_bit_identical(device, axis):
    """Reference-style broadcast velocities (vx[None,:,None], E[:,:,None], ...)."""
    v2 = _v2()
    rng = np.random.default_rng(20 + axis)
    shape = (8, 10, 12, 14)
    f = rng.random(shape)
    dt, dx = 0.05, 0.2
    if axis == 0:
        vel = np.linspace(-6, 6, shape[2])[None, :, None]
    elif axis == 1:
        vel = np.linspace(-6, 6, shape[3])[None, None, :]
    else:
        vel = rng.normal(size=(shape[0], shape[1]))[:, :, None] * 3
    ref = vo.advect_axis(f, vel, axis, dt, dx, periodic=axis < 2)
    out = v2.advect_1d(torch.from_numpy(f).to(device), torch.from_numpy(np.ascontiguousarray(vel)).to(device),
                       axis, dt, dx, periodic=axis < 2).cpu().numpy()
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("shape,axis", [((700, 3, 37, 2), 0), ((2, 3, 600, 40), 2), ((1, 1, 4096, 33), 2),
                                        ((1, 1000, 64, 1), 1), ((3, 2, 5, 300), 3), ((1, 1, 7, 4096), 3),
                                        ((1, 9, 2, 2500), 3)])
@pytest.mark.parametrize("spread", [0.3, 4.0, 60.0, 3000.0])
@pytest.mark.parametrize("periodic", [True, False])
def test_advect_long_lines(device, shape, axis, spread, periodic):
    """Lines longer than 256, strided (row segments staged with their taps'
    halo; a displacement spread wider than the halo reads taps from global
    memory) and contiguous (segments of one line)."""
    v2 = _v2()
    rng = np.random.default_rng(int(shape[axis]) + int(spread * 10))
    f = rng.random(shape)
    dt, dx = 0.1, 0.25
    batch = [1 if i == axis else s for i, s in enumerate(shape)]
    vel = rng.uniform(-1, 1, size=batch) * spread * dx / dt
    vel = vel.reshape([s for i, s in enumerate(batch) if i != axis])
    ref = vo.advect_axis(f, vel, axis, dt, dx, periodic=periodic)
    out = v2.advect_1d(torch.from_numpy(f).to(device), torch.from_numpy(vel).to(device), axis, dt, dx,
                       periodic=periodic).cpu().numpy()
    np.testing.assert_array_equal(out, ref)


def test_segment_lengths_do_not_change_results(device):
    """The segment-length knobs of the long-line sweeps (strided rows per
    segment, contiguous indices per segment) choose launch shapes only."""
    from vlasov_amd import _lib as L
    v2 = _v2()
    lib = L.lib()
    rng = np.random.default_rng(41)
    f = torch.from_numpy(rng.random((1, 1, 1500, 1200))).to(device)
    vx = torch.from_numpy(rng.uniform(-30, 30, size=(1, 1, 1200))).to(device)
    ve = torch.from_numpy(rng.uniform(-8, 8, size=(1, 1, 1500))).to(device)
    outs = {}
    old = {k: lib.vp_tune_get(L.KNOBS[k]) for k in ("seg_rows", "seg_contig")}
    try:
        for rows, contig in ((0, 0), (16, 256), (32, 512), (64, 2048), (128, 4096)):
            lib.vp_tune_set(L.KNOBS["seg_rows"], rows)
            lib.vp_tune_set(L.KNOBS["seg_contig"], contig)
            a = v2.advect_1d(f, vx, 2, 0.1, 0.25, periodic=True)
            b = v2.advect_1d(f, ve, 3, 0.1, 0.25, periodic=False)
            outs[(rows, contig)] = (a, b)
    finally:
        for k, v in old.items():
            lib.vp_tune_set(L.KNOBS[k], v)
    a0, b0 = outs[(0, 0)]
    for a, b in outs.values():
        assert torch.equal(a, a0) and torch.equal(b, b0)
    ref = vo.advect_axis(f.cpu().numpy(), vx.cpu().numpy(), 2, 0.1, 0.25, periodic=True)
    np.testing.assert_array_equal(a0.cpu().numpy(), ref)


@pytest.mark.parametrize("periodic", [True, False])
def test_advect_rows_per_point_velocity(device, periodic):
    v1 = _v1()
    rng = np.random.default_rng(3)
    f = rng.random((33, 47))
    vel = rng.normal(size=(33, 47)) * 20
    ref = vo.advect_rows(f, vel, 0.1, 0.3, periodic=periodic)
    out = v1.advect_rows(torch.from_numpy(f).to(device), torch.from_numpy(vel).to(device), 0.1, 0.3,
                         periodic=periodic).cpu().numpy()
    np.testing.assert_array_equal(out, ref)


def test_advect_dirichlet_shift_out_is_zero(device):
    v1 = _v1()
    f = torch.rand(16, 16, dtype=torch.float64, device=device)
    out = v1.advect_rows(f, torch.ones(16, dtype=torch.float64, device=device), 12.0, 8.0 / 15, periodic=False)
    assert torch.count_nonzero(out).item() == 0


@pytest.mark.parametrize("nx", [8, 16, 128, 256, 512, 1000, 1024, 2048, 4096, 4098])
def test_rho_and_poisson_1d(device, nx):
    """Power-of-two nx up to 4096: FFTs, one stage per LDS pass below 1024
    and three from 1024 (ending on a pass of one stage at 1024, two at 2048,
    none at 4096); other nx: the many-workgroup DFT (512..4096) or the
    one-workgroup DFT."""
    v1 = _v1()
    g = vo.grid_1d(nx, 96, 4 * math.pi, 6.0)
    rng = np.random.default_rng(4)
    f = (1 + 0.1 * np.cos(0.5 * g.x) + 0.05 * np.sin(1.5 * g.x))[:, None] * np.exp(-0.5 * g.v ** 2)[None, :] \
        / math.sqrt(2 * math.pi) + 1e-3 * rng.random((nx, 96))
    E_ref, rho_ref, phi_ref = vo.solve_poisson_real_1d(f, g.inv_kx, g.inv_kx_sq, g.dv)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    E, rho, phi = v1.solve_poisson_real(t(f), t(g.inv_kx), t(g.inv_kx_sq), g.dv)
    assert np.max(np.abs(rho.cpu().numpy() - rho_ref)) <= 1e-14
    assert rel(E.cpu().numpy(), E_ref) <= 1e-13
    assert rel(phi.cpu().numpy(), phi_ref) <= 1e-13


@pytest.mark.parametrize("n", [(16, 16), (64, 32), (128, 128), (24, 20), (8, 8), (8, 512), (512, 16), (256, 256),
                               (32, 128), (128, 64), (512, 512), (96, 128)])
def test_poisson_2d(device, n):
    """Power-of-two grids run radix-2 FFTs (rows of 8..512 points, several
    rows per workgroup, Ex + i Ey through one inverse), other grids the
    exact-twiddle DFTs; both against the oracle (numpy.fft), and the FFT
    path against the DFT kernels forced by the knob."""
    v2 = _v2()
    from vlasov_amd import _lib
    Nx, Ny = n
    g = vo.grid_2d(Nx, Ny, 4, 4, 4 * math.pi, 3 * math.pi, 6.0, 6.0)
    rng = np.random.default_rng(5)
    rho = 1 + 0.01 * rng.normal(size=(Nx, Ny))
    ref = vo.compute_poisson_real(rho, g.kx, g.ky, g.inv_ksq)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    out = v2.compute_poisson_real(t(rho), t(g.kx), t(g.ky), t(g.inv_ksq))
    names = ["Ex", "Ey", "delta_rho", "phi"]
    for name, o, r in zip(names, out, ref):
        assert rel(o.cpu().numpy(), r) <= 1e-13, (name, rel(o.cpu().numpy(), r))
    lib = _lib.lib()
    old = lib.vp_tune_get(_lib.KNOBS["poisson_dft"])
    try:
        lib.vp_tune_set(_lib.KNOBS["poisson_dft"], 1)
        dft = v2.compute_poisson_real(t(rho), t(g.kx), t(g.ky), t(g.inv_ksq))
    finally:
        lib.vp_tune_set(_lib.KNOBS["poisson_dft"], old)
    for name, o, r in zip(names, out, dft):
        assert rel(o.cpu().numpy(), r.cpu().numpy()) <= 1e-13, name
        if name == "delta_rho":
            assert torch.equal(o, r)


def test_compute_rho_2d(device):
    v2 = _v2()
    rng = np.random.default_rng(6)
    f = rng.random((6, 5, 20, 18))
    ref = vo.compute_rho(f, 0.3, 0.7)
    out = v2.compute_rho(torch.from_numpy(f).to(device), 0.3, 0.7).cpu().numpy()
    assert rel(out, ref) <= 1e-15


def _landau1d(Nx, Nv, eps=0.01):
    g = vo.grid_1d(Nx, Nv, 4 * math.pi, 6.0)
    fM = np.exp(-0.5 * g.v ** 2) / math.sqrt(2 * math.pi)
    return g, (1 + eps * np.cos(0.5 * g.x))[:, None] * fM[None, :]


def test_step_1d_landau_trajectory(device):
    """1D1V Landau at 128x128: 100 steps, f and the field-energy history."""
    v1 = _v1()
    G = v1.Grid.create(128, 128, 4 * math.pi, 6.0, device=device, dtype=torch.float64)
    v_np, ik, ik2 = G.v.cpu().numpy(), G.inv_kx.cpu().numpy(), G.inv_kx_sq.cpu().numpy()
    _, f0 = _landau1d(128, 128)
    f_ref = f0.copy()
    f = torch.from_numpy(f0).to(device)
    W_ref, W = [], []
    for _ in range(100):
        f_ref, rho_r, phi_r, E_r = vo.vlasov_step_1d(f_ref, 0.1, v_np, ik, ik2, G.dx, G.dv)
        f, rho, phi, E = v1.step_with_field(f, 0.1, G.v, G.inv_kx, G.inv_kx_sq, G.dx, G.dv, True)
        W_ref.append(vo.field_energy_1d(E_r, G.dx))
        W.append(vo.field_energy_1d(E.cpu().numpy(), G.dx))
    fg = f.cpu().numpy()
    assert np.linalg.norm(fg - f_ref) / np.linalg.norm(f_ref) <= 1e-12
    W, W_ref = np.array(W), np.array(W_ref)
    assert np.max(np.abs(W - W_ref)) / np.max(W_ref) <= 1e-12


def test_step_1d_free_streaming(device):
    v1 = _v1()
    G = v1.Grid.create(64, 48, 2 * math.pi, 5.0, device=device, dtype=torch.float64)
    rng = np.random.default_rng(7)
    f0 = rng.random((64, 48))
    ref, *_ = vo.vlasov_step_1d(f0, 0.3, G.v.cpu().numpy(), None, None, G.dx, G.dv, physics_mode=False)
    out, rho, phi = v1.vlasov_step(torch.from_numpy(f0).to(device), 0.3, G.v, G.inv_kx, G.inv_kx_sq, G.x_indices,
                                   G.v_indices, G.x_batch_offsets, G.v_batch_offsets, G.stencil_offsets, G.dx,
                                   G.dv, G.V_max, False)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    assert torch.count_nonzero(rho).item() == 0 and torch.count_nonzero(phi).item() == 0


def _grid2_arrays(G):
    return (G.vx.cpu().numpy(), G.vy.cpu().numpy(), G.kx.cpu().numpy(), G.ky.cpu().numpy(),
            G.inv_ksq.cpu().numpy())


@pytest.mark.parametrize("physics_mode", [True, False])
def test_step_2d_random_16(device, physics_mode):
    v2 = _v2()
    G = v2.Grid.create(16, 16, 16, 16, 2 * math.pi, 2 * math.pi, 4.0, 4.0, device=device, dtype=torch.float64)
    g = vo.grid_2d(16, 16, 16, 16, 2 * math.pi, 2 * math.pi, 4.0, 4.0)
    rng = np.random.default_rng(8)
    f0 = rng.random((16, 16, 16, 16))
    ref = vo.vlasov_step_2d(f0, 0.05, g, physics_mode=physics_mode, grid_arrays=_grid2_arrays(G))
    out, drho, phi, Ex, Ey = v2.step_with_field(torch.from_numpy(f0).to(device), 0.05, G.vx, G.vy, G.kx, G.ky,
                                                G.inv_ksq, G.dx, G.dy, G.dvx, G.dvy, physics_mode)
    assert rel(out.cpu().numpy(), ref[0]) <= 1e-13
    if physics_mode:
        for o, r in zip((drho, phi, Ex, Ey), ref[1:]):
            assert rel(o.cpu().numpy(), r) <= 1e-13
    else:
        np.testing.assert_array_equal(out.cpu().numpy(), ref[0])


def test_step_2d_landau_trajectory(device):
    """2D2V Landau (C1 physics) on 24^2 x 32^2 for 20 steps: f and W_E history."""
    v2 = _v2()
    Nx, Nv = 24, 32
    L = 4 * math.pi
    G = v2.Grid.create(Nx, Nx, Nv, Nv, L, L, 6.0, 6.0, device=device, dtype=torch.float64)
    g = vo.grid_2d(Nx, Nx, Nv, Nv, L, L, 6.0, 6.0)
    ga = _grid2_arrays(G)
    x, vx = G.x.cpu().numpy(), G.vx.cpu().numpy()
    pert = 1 + 0.01 * (np.cos(0.5 * x)[:, None] + np.cos(0.5 * x)[None, :])
    fM = np.exp(-0.5 * (vx[:, None] ** 2 + vx[None, :] ** 2)) / (2 * math.pi)
    f_ref = pert[:, :, None, None] * fM[None, None]
    f = torch.from_numpy(f_ref.copy()).to(device)
    W, W_ref = [], []
    for _ in range(20):
        f_ref, _, _, Exr, Eyr = vo.vlasov_step_2d(f_ref, 0.05, g, True, grid_arrays=ga)
        f, _, _, Ex, Ey = v2.step_with_field(f, 0.05, G.vx, G.vy, G.kx, G.ky, G.inv_ksq, G.dx, G.dy, G.dvx,
                                             G.dvy, True)
        W_ref.append(vo.field_energy_2d(Exr, Eyr, G.dx, G.dy))
        W.append(vo.field_energy_2d(Ex.cpu().numpy(), Ey.cpu().numpy(), G.dx, G.dy))
    fg = f.cpu().numpy()
    assert np.linalg.norm(fg - f_ref) / np.linalg.norm(f_ref) <= 1e-12
    assert np.max(np.abs(np.array(W) - np.array(W_ref))) / max(W_ref) <= 1e-12


def test_vlasov_driver_1d(device, tmp_path):
    v1 = _v1()
    s = v1.Vlasov(Nx=32, Nv=32, nbiter=10, diag_steps=5, physics_mode=True, dt=0.1, out_dir=str(tmp_path),
                  device=device)
    st = s.initialize()
    f0 = st.variables.fn.cpu().numpy()
    st = s.run(st)
    g = s.grid
    f = f0
    for _ in range(10):
        f, rho, phi, E = vo.vlasov_step_1d(f, 0.1, g.v.cpu().numpy(), g.inv_kx.cpu().numpy(),
                                           g.inv_kx_sq.cpu().numpy(), g.dx, g.dv)
    assert rel(st.variables.fn.cpu().numpy(), f) <= 1e-13
    assert st.step == 10 and abs(st.t - 1.0) < 1e-12
    assert (tmp_path / "field_000010.npz").exists()
