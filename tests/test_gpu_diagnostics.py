"""GPU tests of the diagnostics (moments, field energy), the FP32-storage
Vlasov driver and the reference-style command line.

* moments (mass, kinetic energy, L2) of f against the float64 oracle
  (compensated sums: <= 1e-13 relative), float64 and float32 storage;
* the field energy (1/2 sum (Ex^2 + Ey^2)) dx dy at sizes that are not
  multiples of the kernel's load batches, against exactly rounded sums;
* the 2D2V driver with dtype='float32' tracks the float64 driver's
  field-energy history to 1e-5 and conserves mass;
* the command line writes the log lines the reference's parsers read
  (parse_cross_platform_vlp2d.py:88-90, parse_cross_platform_vlp1d.py).
"""
import math
import re

import numpy as np
import pytest
import torch

from oracle import vlasov_oracle as vo

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 100, 1023, 16384, 100003])
def test_field_energy_exactly_rounded_sums(device, n):
    """Compensated sums of E^2 (1D, Ey null) and Ex^2 + Ey^2 (2D): within 2
    ulps of math.fsum of the same float64 squares and sums, then scaled as
    the kernel scales them."""
    from vlasov_amd import _lib
    lib = _lib.lib()
    rng = np.random.default_rng(n)
    ex = rng.standard_normal(n) * 10.0 ** rng.uniform(-6, 2, n)
    ey = rng.standard_normal(n)
    dx, dy = 0.37, 1.9
    tx, ty = torch.from_numpy(ex).to(device), torch.from_numpy(ey).to(device)
    out = torch.empty(2, dtype=torch.float64, device=device)
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.vp_field_energy_f64(tx.data_ptr(), None, n, dx, 1.0, out.data_ptr(), s), "energy")
    _lib.check(lib.vp_field_energy_f64(tx.data_ptr(), ty.data_ptr(), n, dx, dy, out.data_ptr() + 8, s), "energy")
    got = out.cpu().numpy()
    ref1 = ((0.5 * math.fsum(ex * ex)) * dx) * 1.0
    ref2 = ((0.5 * math.fsum(ex * ex + ey * ey)) * dx) * dy
    for g_, r_ in ((got[0], ref1), (got[1], ref2)):
        assert abs(g_ - r_) <= 2 * np.finfo(np.float64).eps * abs(r_), (n, g_, r_)


def test_moments_1d_match_oracle(device):
    from vlasov_amd import vlasov1D1V as v1
    g = v1.Grid.create(100, 130, 4 * math.pi, 6.0, device=device)
    rng = np.random.default_rng(0)
    f = rng.random((100, 130))
    ref = vo.moments_1d(f, g.v.cpu().numpy(), g.dx, g.dv)
    for dt_ in (torch.float64, torch.float32):
        tf = torch.from_numpy(f).to(dt_).to(device)
        got = v1.moments(tf, g).cpu().numpy()
        r = vo.moments_1d(tf.double().cpu().numpy(), g.v.cpu().numpy(), g.dx, g.dv)
        np.testing.assert_allclose(got, r, rtol=1e-13)
    np.testing.assert_allclose(v1.moments(torch.from_numpy(f).to(device), g).cpu().numpy(), ref, rtol=1e-13)


def test_moments_2d_match_oracle(device):
    from vlasov_amd import vlasov2D2V as v2
    g = v2.Grid.create(12, 10, 16, 20, 4 * math.pi, 3 * math.pi, 6.0, 5.0, device=device)
    rng = np.random.default_rng(1)
    f = rng.random((12, 10, 16, 20))
    ref = vo.moments_2d(f, g.vx.cpu().numpy(), g.vy.cpu().numpy(), g.dx, g.dy, g.dvx, g.dvy)
    got = v2.moments(torch.from_numpy(f).to(device), g).cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-13)
    f32 = torch.from_numpy(f).float().to(device)
    ref32 = vo.moments_2d(f32.double().cpu().numpy(), g.vx.cpu().numpy(), g.vy.cpu().numpy(), g.dx, g.dy, g.dvx,
                          g.dvy)
    np.testing.assert_allclose(v2.moments(f32, g).cpu().numpy(), ref32, rtol=1e-13)


def test_vlasov2d_float32_driver_tracks_float64(device, tmp_path):
    from vlasov_amd import vlasov2D2V as v2
    kw = dict(Nx=64, Ny=64, Nvx=32, Nvy=64, Lx=4 * math.pi, Ly=4 * math.pi, Vx_max=6.0, Vy_max=6.0, nbiter=20,
              diag_steps=10, dt=0.1, physics_mode=True, epsilon=0.05, device=device)
    a = v2.Vlasov(dtype="float64", out_dir=str(tmp_path / "a"), **kw)
    b = v2.Vlasov(dtype="float32", out_dir=str(tmp_path / "b"), **kw)
    sa = a.run(a.initialize())
    sb = b.run(b.initialize())
    assert sb.variables.fn.dtype == torch.float32 and sa.variables.fn.dtype == torch.float64
    err = np.max(np.abs(b.energy_history - a.energy_history)) / np.max(a.energy_history)
    assert err <= 1e-5, err
    masses = np.array([m[2] for m in b.moment_history])
    assert np.max(np.abs(masses - masses[0])) / masses[0] <= 1e-6
    assert len(a.moment_history) == 3  # initial snapshot + two diagnostics


def test_vlasov1d_driver_records_energy_and_moments(device, tmp_path):
    from vlasov_amd import vlasov1D1V as v1
    s = v1.Vlasov(Nx=128, Nv=128, L=4 * math.pi, V_max=6.0, nbiter=30, diag_steps=10, dt=0.1, physics_mode=True,
                  epsilon=0.01, out_dir=str(tmp_path), device=device)
    st = s.initialize()
    f, g = st.variables.fn.cpu().numpy(), s.grid
    args = (g.v.cpu().numpy(), g.inv_kx.cpu().numpy(), g.inv_kx_sq.cpu().numpy(), g.dx, g.dv)
    W = []
    for _ in range(30):
        f, _, _, E = vo.vlasov_step_1d(f, 0.1, *args)
        W.append(vo.field_energy_1d(E, g.dx))
    s.run(st)
    assert np.max(np.abs(s.energy_history - np.array(W))) / max(W) <= 1e-12
    mass = np.array([m[2] for m in s.moment_history])
    # the zero (Dirichlet) velocity boundary lets mass out at the exp(-V_max^2/2) level
    assert len(mass) == 4 and np.max(np.abs(mass - mass[0])) / mass[0] <= 1e-9


# The reference's log parsers read these three patterns
# (python_scripts/plot/src/parse_cross_platform_vlp2d.py:88-90).
_ELAPSED = re.compile(r"Elapsed time:\s*([\d.eE+\-]+)\s*\[s\]")
_GRID = re.compile(r"Grid:\s*(\d+(?:\s*x\s*\d+)+)")
_STEPS = re.compile(r"(?:Steps|Total iterations):\s*(\d+)")


@pytest.mark.parametrize("two_d", [True, False])
def test_command_line_log_matches_reference_parsers(device, tmp_path, monkeypatch, two_d):
    monkeypatch.chdir(tmp_path)
    if two_d:
        from vlasov_amd import vlasov2D2V as mod
        argv = ["-nx", "32", "-ny", "32", "-nvx", "32", "-nvy", "32", "-nbiter", "4", "-diag_steps", "2",
                "-physics_mode", "-out_dir", str(tmp_path / "d")]
        name, grid = "vlp2d_2v_float64.txt", "32x32x32x32"
    else:
        from vlasov_amd import vlasov1D1V as mod
        argv = ["-nx", "64", "-nv", "64", "-nbiter", "6", "-diag_steps", "3", "-physics_mode", "-out_dir",
                str(tmp_path / "d")]
        name, grid = "vlp1d_1v_float64.txt", "64x64"
    mod.main(argv)
    text = (tmp_path / name).read_text()
    assert float(_ELAPSED.search(text).group(1)) > 0
    assert _GRID.search(text).group(1).replace(" ", "") == grid
    assert int(_STEPS.search(text).group(1)) == (4 if two_d else 6)


def test_binary_snapshots_and_restart_2d(device, tmp_path):
    """A 2D run that writes the Kokkos solver's raw snapshots, restarted from
    the middle one, ends bit for bit where the uninterrupted run ends."""
    from vlasov_amd import snapshots
    from vlasov_amd import vlasov2D2V as v2
    kw = dict(Nx=32, Ny=32, Nvx=32, Nvy=32, Lx=4 * math.pi, Ly=4 * math.pi, Vx_max=6.0, Vy_max=6.0,
              diag_steps=3, dt=0.1, physics_mode=True, epsilon=0.05, device=device, snapshot_format="bin")
    a = v2.Vlasov(nbiter=6, out_dir=str(tmp_path), **kw)
    sa = a.run(a.initialize())
    assert snapshots.snapshot_steps(tmp_path) == [0, 3, 6]
    b = v2.Vlasov(nbiter=3, out_dir=str(tmp_path), **kw)
    st = b.restart(3)
    assert st.step == 3 and st.variables.fn.shape == (32, 32, 32, 32)
    sb = b.run(st)
    assert torch.equal(sb.variables.fn, sa.variables.fn)
    np.testing.assert_array_equal(b.energy_history, a.energy_history[3:])
    last = snapshots.read_snapshot(tmp_path, 6, {"fn": (32, 32, 32, 32), "f0": (32, 32), "rho": (32, 32),
                                                "phi": (32, 32)})
    np.testing.assert_array_equal(last["fn"], sa.variables.fn.cpu().numpy())


def test_binary_snapshots_and_restart_1d(device, tmp_path):
    from vlasov_amd import vlasov1D1V as v1
    kw = dict(Nx=64, Nv=96, L=4 * math.pi, V_max=6.0, diag_steps=5, dt=0.1, physics_mode=True, epsilon=0.01,
              device=device, snapshot_format="bin")
    a = v1.Vlasov(nbiter=10, out_dir=str(tmp_path), **kw)
    sa = a.run(a.initialize())
    b = v1.Vlasov(nbiter=5, out_dir=str(tmp_path), **kw)
    sb = b.run(b.restart(5))
    assert torch.equal(sb.variables.fn, sa.variables.fn)
    np.testing.assert_array_equal(b.energy_history, a.energy_history[5:])
