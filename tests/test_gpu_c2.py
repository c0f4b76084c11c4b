"""The 2D2V two-stream configuration at 128^4 (FP64) and its 128-row kernels.

The benchmark workload is f[128, 128, 128, 128] with
f0 = (1 + 0.001 (cos kx + cos ky)) 1/2 [M(vx - 2.4) + M(vx + 2.4)] M(vy),
k = 0.2 (L = 10 pi), v in [-8, 8], dt = 0.05.  Its dominant kernel is the
Y(dt/2) drift with the charge-density partial sums over 128 rows
(drift_dma_kernel<double, 4, 1, true, true, 2, NT, 1>: 4 rows per thread,
a row reduction of 8 threads per row and 4 columns each, the vx loop split
12 ways at nvx = 64 and 128).

* One step at ny = 128 against the oracle (vlasov_step_impl,
  simulations/vlasov2d_2v/pytorch/src/vlasov2D2V.py:943-998) on the
  two-stream initial condition at shapes the oracle steps in seconds:
  array-equal with the field off, <= 1e-13 with it on, for the default vx
  split, for one split and for persistent Y-drift workgroups.
* At the full 128^4 shape, where the oracle would take minutes per step,
  size-independent properties: the fused runner (X halves of neighbouring
  steps merged) equals the generic six-sweep path, mass is conserved, and
  the x and y modes follow linear theory: the x mode grows at the root of
  the two-beam dispersion relation
  1 + (1/k^2) [1 + 1/2 (zeta_+ Z(zeta_+) + zeta_- Z(zeta_-))] = 0,
  zeta_+- = (omega -+ k u) / (sqrt(2) k)
  (gamma = 0.22584 at k = 0.2, u = 2.4), the y mode oscillates at the
  Landau root at k = 0.2 (omega = 1.06398).
"""
import math

import numpy as np
import pytest
import torch

from oracle import vlasov_oracle as vo

pytestmark = pytest.mark.gpu

L_BOX = 10 * math.pi
V_MAX = 8.0
DT = 0.05
K = 0.2
U = 2.4


def _mods():
    from vlasov_amd import _lib, configs, runner
    from vlasov_amd import vlasov2D2V as v2
    return v2, runner, configs, _lib


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


def _grid(v2, shape, device):
    nx, ny, nvx, nvy = shape
    G = v2.Grid.create(nx, ny, nvx, nvy, L_BOX, L_BOX, V_MAX, V_MAX, device=device, dtype=torch.float64)
    g = vo.grid_2d(nx, ny, nvx, nvy, L_BOX, L_BOX, V_MAX, V_MAX)
    arrays = tuple(a.cpu().numpy() for a in (G.vx, G.vy, G.kx, G.ky, G.inv_ksq))
    return G, g, arrays


_ORACLE = {}


def _oracle(shape, physics_mode, steps, device):
    """Oracle trajectory of the two-stream initial condition (cached)."""
    key = (shape, physics_mode, steps)
    if key not in _ORACLE:
        v2, _, configs, _ = _mods()
        G, g, arrays = _grid(v2, shape, device)
        f0 = configs.two_stream_2d(G, eps=0.001, k=K, u=U)
        f = f0.cpu().numpy()
        out = []
        for _ in range(steps):
            f, drho, phi, Ex, Ey = vo.vlasov_step_2d(f, DT, g, physics_mode=physics_mode, grid_arrays=arrays)
            out.append((f, drho, phi, Ex, Ey, vo.field_energy_2d(Ex, Ey, g.dx, g.dy)))
        _ORACLE[key] = (f0, out)
    return _ORACLE[key]


def _exact_delta_rho(f0, g, arrays):
    """delta_rho of the mid-step f from exactly rounded velocity sums
    (math.fsum), scaled as the reference scales them (sum * dvx * dvy - 1)."""
    vx, vy = arrays[0], arrays[1]
    h = vo.advect_axis(f0, vx[None, :, None], 0, DT / 2.0, g.dx, periodic=True)
    h = vo.advect_axis(h, vy[None, None, :], 1, DT / 2.0, g.dy, periodic=True)
    rows = h.reshape(g.Nx * g.Ny, -1)
    S = np.array([math.fsum(r) for r in rows]).reshape(g.Nx, g.Ny)
    return S * g.dvx * g.dvy - 1.0


@pytest.mark.parametrize("shape", [(32, 128, 32, 128), (32, 128, 64, 128)])
@pytest.mark.parametrize("physics_mode", [False, True])
def test_ny128_step_matches_oracle(device, shape, physics_mode):
    """f: array-equal with the field off, <= 1e-13 with it on.  The charge
    density: delta_rho = rho - 1 cancels the leading digits of rho ~ 1, and
    the two-stream perturbation is 2e-3, so sums of rho in two different
    orders (compensated partial sums here, pairwise numpy.sum in the oracle)
    differ by a few ulps of 1 = a few 1e-13 of max |delta_rho|.  Hence
    delta_rho is compared with the oracle to 1e-13 of max rho and with the
    exactly rounded sums to 2 ulps of 1, and phi, Ex, Ey (linear in
    delta_rho) to 1e-12 of their maxima."""
    v2, runner, configs, L = _mods()
    assert runner.fused_supported(shape)
    G, g, arrays = _grid(v2, shape, device)
    f0, traj = _oracle(shape, physics_mode, 1, device)
    ref = traj[0]
    exact = _exact_delta_rho(f0.cpu().numpy(), g, arrays) if physics_mode else None
    lib = L.lib()
    old = {k: lib.vp_tune_get(L.KNOBS[k]) for k in ("y_split", "y_units")}
    try:
        # (units mode, split): the default (one workgroup per (x, vx range),
        # 3 ranges at nvx = 32, 12 at 64 and 128), one range, and persistent
        # workgroups walking several units (two each by default; 3 ranges:
        # ranges of 10-11 tiles and one or two units per workgroup), with
        # rotated starts, spread over the XCDs
        for units, split in ((0, 0), (1, 1), (2, 0), (3, 3)):
            assert lib.vp_tune_set(L.KNOBS["y_units"], units) == 0
            assert lib.vp_tune_set(L.KNOBS["y_split"], split) == 0
            out, drho, phi, Ex, Ey = v2.step_with_field(f0, DT, G.vx, G.vy, G.kx, G.ky, G.inv_ksq, G.dx, G.dy,
                                                        G.dvx, G.dvy, physics_mode)
            torch.cuda.synchronize()
            if physics_mode:
                assert rel(out.cpu().numpy(), ref[0]) <= 1e-13, (units, split)
                d = drho.cpu().numpy()
                assert np.max(np.abs(d - ref[1])) <= 1e-13 * np.max(np.abs(ref[1] + 1.0)), (units, split)
                assert np.max(np.abs(d - exact)) <= 2 * np.finfo(np.float64).eps, (units, split)
                for o, r in zip((phi, Ex, Ey), ref[2:5]):
                    assert rel(o.cpu().numpy(), r) <= 1e-12, (units, split)
            else:
                np.testing.assert_array_equal(out.cpu().numpy(), ref[0])
    finally:
        for k, v in old.items():
            lib.vp_tune_set(L.KNOBS[k], v)


def test_ny128_fused_trajectory_matches_oracle(device):
    """Run2D's fused schedule (closing and opening X halves in one pass) over
    3 steps at (32, 128, 64, 128): f and the field-energy history <= 1e-12."""
    v2, runner, configs, L = _mods()
    shape = (32, 128, 64, 128)
    G, g, arrays = _grid(v2, shape, device)
    f0, traj = _oracle(shape, True, 3, device)
    run = runner.Run2D(G, f0, DT, True)
    assert run.fused
    W = torch.zeros(3, dtype=torch.float64, device=device)
    run.advance(3, energy=W)
    torch.cuda.synchronize()
    f_ref = traj[-1][0]
    assert np.linalg.norm(run.state.cpu().numpy() - f_ref) / np.linalg.norm(f_ref) <= 1e-12
    W_ref = np.array([t[5] for t in traj])
    assert np.max(np.abs(W.cpu().numpy() - W_ref)) / np.max(W_ref) <= 1e-12
    assert rel(run.Ex.cpu().numpy(), traj[-1][3]) <= 1e-12


@pytest.fixture(scope="module")
def c2_state(device):
    v2, runner, configs, _ = _mods()
    shape = (128, 128, 128, 128)
    G = v2.Grid.create(*shape, L_BOX, L_BOX, V_MAX, V_MAX, device=device, dtype=torch.float64)
    f0 = configs.two_stream_2d(G, eps=0.001, k=K, u=U)
    return G, f0


@pytest.mark.parametrize("physics_mode", [False, True])
def test_c2_fused_equals_generic(device, c2_state, physics_mode):
    """128^4 two-stream, 2 steps: fused schedule == generic six sweeps
    (identical with the field off, <= 1e-13 in f and <= 1e-12 in the field
    energy with it on: rho is summed in another order)."""
    v2, runner, configs, _ = _mods()
    G, f0 = c2_state
    a = runner.Run2D(G, f0, DT, physics_mode, fused=True)
    ea = torch.zeros(2, dtype=torch.float64, device=device)
    a.advance(2, energy=ea)
    fa = a.state.clone()
    del a
    b = runner.Run2D(G, f0, DT, physics_mode, fused=False)
    eb = torch.zeros(2, dtype=torch.float64, device=device)
    b.advance(2, energy=eb)
    torch.cuda.synchronize()
    if physics_mode:
        d = torch.linalg.vector_norm(fa - b.state) / torch.linalg.vector_norm(b.state)
        assert float(d) <= 1e-13
        assert rel(ea.cpu().numpy(), eb.cpu().numpy()) <= 1e-12
        assert float(eb.min()) > 0
    else:
        assert torch.equal(fa, b.state)
        assert float(ea.abs().max()) == 0.0


def _two_beam_root(k, u):
    from scipy.special import wofz

    def D(om):
        s = 0.0
        for uu in (u, -u):
            z = (om - k * uu) / (math.sqrt(2.0) * k)
            s += 0.5 * (1.0 + z * 1j * math.sqrt(math.pi) * wofz(z))
        return 1.0 + s / (k * k)

    om = 0.3j
    for _ in range(100):
        d = (D(om + 1e-7) - D(om - 1e-7)) / 2e-7
        step = D(om) / d
        om -= step
        if abs(step) < 1e-15:
            break
    assert abs(D(om)) < 1e-12
    return om


def _fit(t, a, modes, start):
    """Least-squares residual of a(t) against a sum of damped/growing modes;
    ``modes`` maps parameters to a list of basis functions of t."""
    from scipy.optimize import minimize

    def resid(p):
        B = np.stack(modes(p, t), 1)
        c, *_ = np.linalg.lstsq(B, a, rcond=None)
        return np.linalg.norm(B @ c - a) / np.linalg.norm(a)

    best = None
    for s in start:
        r = minimize(resid, s, method="Nelder-Mead", options=dict(xatol=1e-10, fatol=1e-14, maxiter=20000,
                                                                  maxfev=40000))
        if best is None or r.fun < best.fun:
            best = r
    return best


def test_c2_two_stream_linear_theory_and_mass(device, c2_state):
    """300 steps of the benchmark workload: mass conserved to 1e-9, the x
    mode grows at the two-beam root (within 1 %), the y mode oscillates at
    the Landau frequency (within 1 %), the field-energy kernel agrees with
    1/2 sum (Ex^2 + Ey^2) dx dy of the fields it was given."""
    v2, runner, configs, _ = _mods()
    G, f0 = c2_state
    steps = 300
    run = runner.Run2D(G, f0, DT, True)
    assert run.fused
    m0 = float(torch.sum(f0, dtype=torch.float64))
    sx = torch.sin(K * G.x)[:, None]
    sy = torch.sin(K * G.y)[None, :]
    amp = torch.zeros((steps, 2), dtype=torch.float64, device=device)
    W = torch.zeros(steps, dtype=torch.float64, device=device)
    for n in range(steps):
        run.advance(1, energy=W[n:n + 1])
        amp[n, 0] = 2 * torch.mean(run.Ex * sx)
        amp[n, 1] = 2 * torch.mean(run.Ey * sy)
    W_last = 0.5 * float(torch.sum(run.Ex ** 2 + run.Ey ** 2)) * G.dx * G.dy
    m1 = float(torch.sum(run.state, dtype=torch.float64))
    torch.cuda.synchronize()
    assert abs(m1 / m0 - 1.0) <= 1e-9
    assert abs(float(W[-1]) - W_last) <= 1e-12 * W_last
    amp = amp.cpu().numpy()
    assert np.all(np.isfinite(amp))
    t = (np.arange(steps) + 0.5) * DT  # step n's field is the mid-step field

    gamma = _two_beam_root(K, U).imag
    assert abs(gamma - 0.22584) < 1e-4
    m = t >= 2.0

    def xmodes(p, tt):
        g, g2, om, gi, om3, g3 = p
        return [np.exp(g * tt), np.exp(-g2 * tt), np.exp(gi * tt) * np.cos(om * tt), np.exp(gi * tt) * np.sin(om * tt),
                np.exp(-g3 * tt) * np.cos(om3 * tt), np.exp(-g3 * tt) * np.sin(om3 * tt)]

    rx = _fit(t[m], amp[m, 0], xmodes, ([0.3, 0.3, 1.3, 0.0, 1.2, 0.4], [0.2, 0.2, 1.4, -0.01, 1.1, 0.3]))
    assert rx.fun < 1e-3, rx
    assert abs(rx.x[0] / gamma - 1.0) < 0.01, (rx.x[0], gamma)

    omega = vo.landau_root(K, guess=(1.06, -1e-4)).real
    m = t >= 4.0

    def ymodes(p, tt):
        om, gi = p
        return [np.exp(gi * tt) * np.cos(om * tt), np.exp(gi * tt) * np.sin(om * tt)]

    ry = _fit(t[m], amp[m, 1], ymodes, ([1.0, 0.0],))
    assert ry.fun < 0.02, ry
    assert abs(ry.x[0] / omega - 1.0) < 0.01, (ry.x[0], omega)
