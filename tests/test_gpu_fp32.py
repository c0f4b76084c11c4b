"""GPU tests of the FP32-storage variant (f float32 in HBM, interpolated in
float32; charge density, Poisson solve, fields and field energy in float64).

* With E = 0 the fused float32 sweeps equal the generic float32 sweeps
  (same arithmetic: float64 weights rounded once, float32 six-tap sum) bit
  for bit.
* With the field on, fused and generic float32 runs differ only through the
  order of the float64 charge-density sums: <= 1e-6 relative.
* Against the float64 oracle the float32 trajectory stays within 1e-5
  relative in f and in the field-energy history (10 steps, 64x64x32x128).
"""
import functools
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


def _setup(shape, seed, L=4 * math.pi, V=6.0):
    v2, runner, _ = _mods()
    nx, ny, nvx, nvy = shape
    G = v2.Grid.create(nx, ny, nvx, nvy, L, L, V, V, device="cuda", dtype=torch.float64)
    rng = np.random.default_rng(seed)
    x = np.arange(nx) / nx * 2 * math.pi
    y = np.arange(ny) / ny * 2 * math.pi
    vx = np.linspace(-V, V, nvx)
    vy = np.linspace(-V, V, nvy)
    pert = 1 + 0.05 * (np.cos(x)[:, None] + np.sin(y)[None, :])
    fv = np.exp(-0.5 * (vx[:, None] ** 2 + vy[None, :] ** 2)) / (2 * math.pi)
    f0 = pert[:, :, None, None] * fv[None, None] * (1 + 0.01 * rng.random(shape))
    return G, f0


@pytest.mark.parametrize("shape", [(64, 64, 32, 64), (128, 64, 64, 128), (64, 128, 128, 64)])
def test_fp32_fused_free_streaming_equals_generic(device, shape):
    v2, runner, _ = _mods()
    assert runner.fused_supported(shape, torch.float32)
    G, f0 = _setup(shape, 1)
    f = torch.from_numpy(f0).float().to(device)
    a = runner.Run2D(G, f, 0.1, physics_mode=False, fused=True)
    b = runner.Run2D(G, f, 0.1, physics_mode=False, fused=False)
    a.advance(1)
    a.advance(2)
    b.advance(3)
    assert a.state.dtype == torch.float32
    assert torch.equal(a.state, b.state)


def test_fp32_fused_matches_generic_with_field(device):
    v2, runner, _ = _mods()
    shape = (64, 64, 64, 64)
    G, f0 = _setup(shape, 2)
    f = torch.from_numpy(f0).float().to(device)
    a = runner.Run2D(G, f, 0.1, physics_mode=True, fused=True)
    b = runner.Run2D(G, f, 0.1, physics_mode=True, fused=False)
    ea = torch.zeros(4, dtype=torch.float64, device=device)
    eb = torch.zeros(4, dtype=torch.float64, device=device)
    a.advance(4, energy=ea)
    b.advance(4, energy=eb)
    fa, fb = a.state.double().cpu().numpy(), b.state.double().cpu().numpy()
    assert np.max(np.abs(fa - fb)) / np.max(np.abs(fb)) <= 1e-6
    assert np.max(np.abs(ea.cpu().numpy() - eb.cpu().numpy())) / np.max(eb.cpu().numpy()) <= 1e-6


# Landau run at 64 x 64 x 32 x 128 (deviation storage needs 128 vy
# columns), 10 steps of the float64 oracle, shared by the trajectory tests.
_TRAJ_SHAPE = (64, 64, 32, 128)
_TRAJ_STEPS = 10


@functools.lru_cache(maxsize=1)
def _oracle_trajectory():
    """(f after the steps, field-energy history, last delta_rho) of the oracle."""
    G, f0 = _setup(_TRAJ_SHAPE, 4)
    g = vo.grid_2d(*_TRAJ_SHAPE, 4 * math.pi, 4 * math.pi, 6.0, 6.0)
    arrays = tuple(a.cpu().numpy() for a in (G.vx, G.vy, G.kx, G.ky, G.inv_ksq))
    f_ref = f0
    W = []
    drho = None
    for _ in range(_TRAJ_STEPS):
        f_ref, drho, _, Ex, Ey = vo.vlasov_step_2d(f_ref, 0.1, g, True, grid_arrays=arrays)
        W.append(vo.field_energy_2d(Ex, Ey, g.dx, g.dy))
    return f_ref, np.array(W), drho


def test_fp32_trajectory_vs_fp64_oracle(device):
    v2, runner, _ = _mods()
    G, f0 = _setup(_TRAJ_SHAPE, 4)
    run = runner.Run2D(G, torch.from_numpy(f0).float().to(device), 0.1, physics_mode=True)
    assert run.fused
    energy = torch.zeros(_TRAJ_STEPS, dtype=torch.float64, device=device)
    run.advance(_TRAJ_STEPS, energy=energy)
    f_ref, W, _ = _oracle_trajectory()
    f32 = run.state.double().cpu().numpy()
    err_f = np.max(np.abs(f32 - f_ref)) / np.max(np.abs(f_ref))
    err_w = np.max(np.abs(energy.cpu().numpy() - W)) / max(W)
    # the float32 input rounding alone is ~6e-8 relative
    assert err_f <= 1e-5, err_f
    assert err_w <= 1e-5, err_w


def test_fp32_xdrift_two_passes_equal_two_calls(device):
    v2, runner, L = _mods()
    shape = (128, 64, 32, 64)
    G, f0 = _setup(shape, 4)
    f = torch.from_numpy(f0).float().to(device)
    sh = L.i64x4(shape)
    lib = L.lib()
    ws = torch.empty(lib.vp_fused_workspace(sh, 4), dtype=torch.uint8, device=device)
    s = torch.cuda.current_stream().cuda_stream
    L.check(lib.vp_fused_tables_f32(sh, G.vx.data_ptr(), G.vy.data_ptr(), 0.05, G.dx, G.dy, ws.data_ptr(),
                                    ws.numel(), s), "tables")
    one = f.clone()
    for _ in range(2):
        L.check(lib.vp_fused_xdrift_f32(one.data_ptr(), one.data_ptr(), sh, 1, ws.data_ptr(), ws.numel(), s), "x1")
    two = torch.empty_like(f)
    L.check(lib.vp_fused_xdrift_f32(f.data_ptr(), two.data_ptr(), sh, 2, ws.data_ptr(), ws.numel(), s), "x2")
    assert torch.equal(one, two)
    gen = f.clone()
    tmp = torch.empty_like(f)
    vs = L.i64x4((0, 0, 1, 0))
    for src, dst in ((gen, tmp), (tmp, gen)):
        L.check(lib.vp_advect_f32(src.data_ptr(), dst.data_ptr(), sh, 0, G.vx.data_ptr(), vs, 0.05, G.dx, 1, s),
                "advect_f32")
    assert torch.equal(one, gen)


@pytest.mark.parametrize("shape", [(64, 128, 32, 128), (64, 128, 64, 64)])
def test_fp32_y_drift_launch_modes_agree(device, shape):
    """Float32 Y drifts over 128-row lines: persistent workgroups walking
    several (x, vx range) units (the default; rotated starts, also spread
    over the XCDs, also with uneven ranges) and one workgroup per unit give
    the same f bit for bit with the field off; with it on the charge
    density's float64 sums run in another order: <= 1e-6 after 3 steps."""
    v2, runner, L = _mods()
    G, f0 = _setup(shape, 7)
    f = torch.from_numpy(f0).float().to(device)
    lib = L.lib()
    old = {k: lib.vp_tune_get(L.KNOBS[k]) for k in ("y_split", "y_units")}
    free, field = [], []
    try:
        for units, split in ((1, 0), (0, 0), (3, 3), (2, 5)):
            lib.vp_tune_set(L.KNOBS["y_units"], units)
            lib.vp_tune_set(L.KNOBS["y_split"], split)
            a = runner.Run2D(G, f, 0.1, physics_mode=False)
            a.advance(3)
            b = runner.Run2D(G, f, 0.1, physics_mode=True)
            b.advance(3)
            free.append(a.state.clone())
            field.append(b.state.double().cpu().numpy())
            del a, b
    finally:
        for k, v in old.items():
            lib.vp_tune_set(L.KNOBS[k], v)
    for i in range(1, len(free)):
        assert torch.equal(free[0], free[i]), i
        assert np.max(np.abs(field[i] - field[0])) / np.max(np.abs(field[0])) <= 1e-6, i
    assert not torch.equal(free[0], f)


@pytest.mark.parametrize("shape", [(128, 64, 32, 128), (64, 64, 32, 64)])
@pytest.mark.parametrize("passes", [1, 2])
def test_fp32_xdrift_column_pairs_equal_single_columns(device, shape, passes):
    """The float32 X drift moving pairs of vy columns as 8-byte elements
    gives the bits of the one-column-per-thread kernel."""
    v2, runner, L = _mods()
    G, f0 = _setup(shape, 6)
    f = torch.from_numpy(f0).float().to(device)
    sh = L.i64x4(shape)
    lib = L.lib()
    ws = torch.empty(lib.vp_fused_workspace(sh, 4), dtype=torch.uint8, device=device)
    s = torch.cuda.current_stream().cuda_stream
    L.check(lib.vp_fused_tables_f32(sh, G.vx.data_ptr(), G.vy.data_ptr(), 0.05, G.dx, G.dy, ws.data_ptr(),
                                    ws.numel(), s), "tables")
    outs = []
    old = lib.vp_tune_get(L.KNOBS["f32_pairs"])
    try:
        for mode in (2, 1):
            lib.vp_tune_set(L.KNOBS["f32_pairs"], mode)
            o = torch.empty_like(f)
            L.check(lib.vp_fused_xdrift_f32(f.data_ptr(), o.data_ptr(), sh, passes, ws.data_ptr(), ws.numel(), s),
                    "xdrift")
            outs.append(o)
    finally:
        lib.vp_tune_set(L.KNOBS["f32_pairs"], old)
    assert torch.equal(outs[0], outs[1])
    assert not torch.equal(outs[0], f)


@pytest.mark.parametrize("shape", [(64, 128, 32, 128), (64, 64, 32, 64)])
def test_fp32_ydrift_rho_single_columns_and_pairs(device, shape):
    """The float32 Y drift + rho with one vy column per thread (row sums over
    groups of 64 lanes) and with pairs of columns (groups of 32): the same f
    bits, and delta_rho of both within 1e-13 of max rho of an 80-bit sum of
    the drifted f (a lane group merged wrongly would be off by whole partial
    sums)."""
    v2, runner, L = _mods()
    G, f0 = _setup(shape, 7)
    f = torch.from_numpy(f0).float().to(device)
    sh = L.i64x4(shape)
    lib = L.lib()
    s = torch.cuda.current_stream().cuda_stream
    nx, ny = shape[0], shape[1]
    kx, ky, iks = (t.to(torch.float64).contiguous() for t in (G.kx, G.ky, G.inv_ksq))
    outs = []
    old = lib.vp_tune_get(L.KNOBS["f32_pairs"])
    try:
        for mode in (2, 1):
            lib.vp_tune_set(L.KNOBS["f32_pairs"], mode)
            ws = torch.empty(lib.vp_fused_workspace(sh, 4), dtype=torch.uint8, device=device)
            L.check(lib.vp_fused_tables_f32(sh, G.vx.data_ptr(), G.vy.data_ptr(), 0.05, G.dx, G.dy, ws.data_ptr(),
                                            ws.numel(), s), "tables")
            o = torch.empty_like(f)
            L.check(lib.vp_fused_ydrift_f32(f.data_ptr(), o.data_ptr(), sh, 1, ws.data_ptr(), ws.numel(), s),
                    "ydrift")
            fields = torch.empty((4, nx, ny), dtype=torch.float64, device=device)
            L.check(lib.vp_fused_poisson_f32(sh, G.dvx, G.dvy, kx.data_ptr(), ky.data_ptr(), iks.data_ptr(),
                                             fields[0].data_ptr(), fields[1].data_ptr(), fields[2].data_ptr(),
                                             fields[3].data_ptr(), ws.data_ptr(), ws.numel(), s), "poisson")
            torch.cuda.synchronize()
            outs.append((o, fields[2].clone()))
    finally:
        lib.vp_tune_set(L.KNOBS["f32_pairs"], old)
    assert torch.equal(outs[0][0], outs[1][0])
    fh = outs[0][0].double().cpu().numpy().reshape(nx * ny, -1)
    S = fh.astype(np.longdouble).sum(axis=1)
    ref = (S * np.longdouble(G.dvx) * np.longdouble(G.dvy) - 1).astype(np.float64).reshape(nx, ny)
    for mode, (_, drho) in zip((2, 1), outs):
        err = np.max(np.abs(drho.cpu().numpy() - ref))
        assert err <= 1e-13 * np.max(np.abs(ref + 1.0)), (mode, err)


@pytest.mark.parametrize("shape,axis,periodic", [((1, 1, 1000, 300), 2, True), ((2, 1, 40, 2000), 3, False),
                                                 ((600, 2, 3, 70), 0, True)])
def test_fp32_long_line_sweeps(device, shape, axis, periodic):
    """Float32 sweeps of lines longer than 256 (segmented kernels): within
    float32 rounding of the float64 oracle on the same (rounded) input, and
    the same bits for every segment length (which also moves segments to the
    global-memory fallback)."""
    v2, runner, L = _mods()
    lib = L.lib()
    rng = np.random.default_rng(sum(shape))
    f32 = torch.from_numpy(rng.random(shape)).float().to(device)
    batch = tuple(s for i, s in enumerate(shape) if i != axis)
    vel = torch.from_numpy(rng.uniform(-1, 1, size=batch) * 40.0).to(device)
    dt, dx = 0.1, 0.25
    outs = []
    old = {k: lib.vp_tune_get(L.KNOBS[k]) for k in ("seg_rows", "seg_contig")}
    try:
        for rows, contig in ((0, 0), (16, 256), (128, 4096)):
            lib.vp_tune_set(L.KNOBS["seg_rows"], rows)
            lib.vp_tune_set(L.KNOBS["seg_contig"], contig)
            outs.append(v2.advect_1d(f32, vel, axis, dt, dx, periodic=periodic))
    finally:
        for k, v in old.items():
            lib.vp_tune_set(L.KNOBS[k], v)
    assert outs[0].dtype == torch.float32
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
    ref = vo.advect_axis(f32.double().cpu().numpy(), vel.cpu().numpy(), axis, dt, dx, periodic=periodic)
    err = np.max(np.abs(outs[0].double().cpu().numpy() - ref)) / np.max(np.abs(ref))
    assert err <= 2e-6, err


def test_fp32_rejects_in_place_kick_and_small_shapes(device):
    v2, runner, L = _mods()
    lib = L.lib()
    assert lib.vp_run_2d2v_f32_workspace(L.i64x4((32, 64, 64, 64))) == 0
    assert lib.vp_run_2d2v_f32_workspace(L.i64x4((64, 64, 64, 32))) == 0
    shape = (64, 64, 32, 64)
    f = torch.zeros(shape, dtype=torch.float32, device=device)
    E = torch.zeros(shape[:2], dtype=torch.float64, device=device)
    rc = lib.vp_fused_kick_f32(f.data_ptr(), f.data_ptr(), L.i64x4(shape), E.data_ptr(), E.data_ptr(), 0.1, 0.1,
                               0.1, torch.cuda.current_stream().cuda_stream)
    assert rc == -4  # VP_ERR_UNSUPPORTED: FP32 kick runs out of place only


@pytest.mark.parametrize("shape", [(4, 4, 128, 128), (8, 4, 64, 128), (4, 8, 128, 64)])
def test_fp32_kick_kernels_agree(device, shape):
    """The float32 row-split kick (128 vy columns, 2/4/8/16 waves per plane,
    4/8 prefetched rows) and
    the two-wave row-streaming kick give the same bits, with multi-cell and
    out-of-domain shifts in the field and displacements of 8e-17 and 1e-15
    cells (two and seventeen rows of a plane whose departure rounds differently
    from the rest)."""
    v2, runner, L = _mods()
    lib = L.lib()
    nx, ny, nvx, nvy = shape
    rng = np.random.default_rng(5)
    f = torch.from_numpy(rng.random(shape)).float().to(device)
    dt, dvx, dvy = 0.1, 12.0 / (nvx - 1), 12.0 / (nvy - 1)
    choices = np.array([0.0, 1e-18, 8e-17, 1e-15, 0.4, -0.7, 2.0, 5.3, -5.3, 80.0, -80.0])
    Ex = torch.from_numpy(rng.choice(choices, size=(nx, ny)) * dvx / dt).to(device)
    Ey = torch.from_numpy(rng.choice(choices, size=(nx, ny)) * dvy / dt).to(device)
    sh = L.i64x4(shape)
    s = torch.cuda.current_stream().cuda_stream
    outs = []
    old = {k: lib.vp_tune_get(L.KNOBS[k]) for k in ("kick_kernel", "kick_prefetch")}
    try:
        for kk, pf in ((0, 2), (2, 1), (3, 3), (4, 1), (4, 3), (5, 1), (6, 1), (6, 3)):
            lib.vp_tune_set(L.KNOBS["kick_kernel"], kk)
            lib.vp_tune_set(L.KNOBS["kick_prefetch"], pf)
            o = torch.empty_like(f)
            L.check(lib.vp_fused_kick_f32(f.data_ptr(), o.data_ptr(), sh, Ex.data_ptr(), Ey.data_ptr(), dt, dvx, dvy,
                                          s), "vp_fused_kick_f32")
            outs.append(o)
    finally:
        for k, v in old.items():
            lib.vp_tune_set(L.KNOBS[k], v)
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
    # and against the generic float32 sweeps (same arithmetic)
    a = v2.advect_1d(f, Ex[:, :, None], 2, dt, dvx, periodic=False)
    a = v2.advect_1d(a, Ey[:, :, None], 3, dt, dvy, periodic=False)
    assert torch.equal(outs[0], a)


# ---------------------------------------------------------------------------
# Deviation storage: float32 g = f - feq, feq = fx(vx) fy(vy) from float64
# profiles of the (x, y) mean of f (fx: its sums over vy, fy: its sums over
# vx over its total).
# ---------------------------------------------------------------------------
def _split(L, f64, shape, dvx, dvy):
    lib = L.lib()
    dev = f64.device
    sh = L.i64x4(shape)
    feq = torch.empty(shape[2] + shape[3], dtype=torch.float64, device=dev)
    rs = torch.empty(1, dtype=torch.float64, device=dev)
    g = torch.empty(shape, dtype=torch.float32, device=dev)
    ws = torch.empty(lib.vp_dev_split_workspace(sh), dtype=torch.uint8, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    L.check(lib.vp_dev_split_f32(f64.data_ptr(), sh, dvx, dvy, feq.data_ptr(), rs.data_ptr(), g.data_ptr(),
                                 ws.data_ptr(), ws.numel(), s), "vp_dev_split_f32")
    return feq, rs, g


def test_deviation_split_and_join(device):
    v2, runner, L = _mods()
    shape = (8, 6, 32, 128)
    nvx = shape[2]
    rng = np.random.default_rng(21)
    f = rng.random(shape) + 1.0
    dvx, dvy = 0.3, 0.2
    feq, rs, g = _split(L, torch.from_numpy(f).to(device), shape, dvx, dvy)
    M = f.reshape(-1, shape[2] * shape[3]).mean(axis=0).reshape(shape[2:])
    prof = feq.cpu().numpy()
    fx, fy = prof[:nvx], prof[nvx:]
    np.testing.assert_allclose(fx, M.sum(axis=1), rtol=1e-14)
    np.testing.assert_allclose(fy, M.sum(axis=0) / M.sum(), rtol=1e-14)
    E = fx[:, None] * fy[None, :]  # feq as every kernel forms it (one rounded product)
    assert abs(float(rs) - ((E.sum() * dvx) * dvy - 1.0)) <= 1e-12 * E.sum() * dvx * dvy
    np.testing.assert_array_equal(g.cpu().numpy(), (f - E[None, None]).astype(np.float32))
    out32 = torch.empty(shape, dtype=torch.float32, device=device)
    out64 = torch.empty(shape, dtype=torch.float64, device=device)
    L.check(L.lib().vp_dev_join_f32(g.data_ptr(), feq.data_ptr(), L.i64x4(shape), out32.data_ptr(), out64.data_ptr(),
                                    torch.cuda.current_stream().cuda_stream), "vp_dev_join_f32")
    h = E[None, None] + g.cpu().numpy().astype(np.float64)
    np.testing.assert_array_equal(out64.cpu().numpy(), h)
    np.testing.assert_array_equal(out32.cpu().numpy(), h.astype(np.float32))
    # the rounding of the stored deviation only: |f - join| <= half an ulp of g
    assert np.max(np.abs(out64.cpu().numpy() - f)) <= np.max(np.abs(f - E[None, None])) * 2.0 ** -24 * 1.01


def test_deviation_profiles_of_a_separable_mean(device):
    """A mean a(vx) b(vy) (every initial condition of the reference) is its
    own rank-one part: fx fy reproduces it to rounding, so g holds only the
    (x, y) perturbation."""
    v2, runner, L = _mods()
    shape = (8, 4, 32, 128)
    x = np.arange(shape[0]) / shape[0] * 2 * math.pi
    vx = np.linspace(-8, 8, shape[2])
    vy = np.linspace(-8, 8, shape[3])
    a = 0.5 * (np.exp(-0.5 * (vx - 2.4) ** 2) + np.exp(-0.5 * (vx + 2.4) ** 2))
    b = np.exp(-0.5 * vy ** 2) / math.sqrt(2 * math.pi)
    pert = 1 + 0.001 * np.cos(x)
    f = pert[:, None, None, None] * np.ones(shape[1])[None, :, None, None] * (a[:, None] * b[None, :])[None, None]
    feq, rs, g = _split(L, torch.from_numpy(f).to(device), shape, 0.1, 0.1)
    prof = feq.cpu().numpy()
    E = prof[:shape[2], None] * prof[None, shape[2]:]
    M = f.mean(axis=(0, 1))
    assert np.max(np.abs(E - M)) <= 4e-15 * np.max(M)
    gg = g.cpu().numpy().astype(np.float64)
    dev_ref = f - M[None, None]
    assert np.max(np.abs(gg - dev_ref)) <= 1e-6 * np.max(np.abs(dev_ref)) + 1e-15 * np.max(M)


@pytest.mark.parametrize("shape", [(4, 4, 128, 128), (4, 8, 32, 128)])
def test_deviation_kick_matches_float64_kick_of_the_sum(device, shape):
    """The deviation kick stores float32(K32(g) + (Kvx fx)(Kvy fy) - fx fy):
    against the float64 oracle kick of fx fy + g minus fx fy, within the
    float32 arithmetic of the kick of g (64 float32 ulps of max |g|) for
    every launch shape, with multi-cell and out-of-domain shifts; with g = 0
    the correction alone, float64 but for the final rounding."""
    v2, runner, L = _mods()
    lib = L.lib()
    nx, ny, nvx, nvy = shape
    rng = np.random.default_rng(8)
    fx = np.exp(-0.5 * np.linspace(-4, 4, nvx) ** 2)
    fy = np.exp(-0.5 * np.linspace(-4, 4, nvy) ** 2) / 2.5
    E = fx[:, None] * fy[None, :]
    dt, dvx, dvy = 0.1, 8.0 / (nvx - 1), 8.0 / (nvy - 1)
    choices = np.array([0.0, 1e-18, 8e-17, 1e-15, 0.4, -0.7, 2.0, 5.3, -5.3, 80.0, -80.0])
    Ex = rng.choice(choices, size=(nx, ny)) * dvx / dt
    Ey = rng.choice(choices, size=(nx, ny)) * dvy / dt
    s = torch.cuda.current_stream().cuda_stream
    prof = torch.from_numpy(np.concatenate([fx, fy])).to(device)
    tEx, tEy = torch.from_numpy(Ex).to(device), torch.from_numpy(Ey).to(device)
    for scale in (1e-3, 0.0):
        g = ((rng.random(shape) - 0.5) * scale).astype(np.float32)
        h = E[None, None] + g.astype(np.float64)
        ref = vo.advect_axis(h, Ex[:, :, None], 2, dt, dvx, periodic=False)
        ref = vo.advect_axis(ref, Ey[:, :, None], 3, dt, dvy, periodic=False)
        expect = ref - E[None, None]
        tg = torch.from_numpy(g).to(device)
        out = torch.empty_like(tg)
        old = {k: lib.vp_tune_get(L.KNOBS[k]) for k in ("kick_kernel", "kick_prefetch")}
        try:
            for kk, pf in ((0, 2), (0, 1), (5, 3), (3, 1), (6, 0)):
                lib.vp_tune_set(L.KNOBS["kick_kernel"], kk)
                lib.vp_tune_set(L.KNOBS["kick_prefetch"], pf)
                L.check(lib.vp_fused_kick_dev_f32(tg.data_ptr(), out.data_ptr(), L.i64x4(shape), prof.data_ptr(),
                                                  tEx.data_ptr(), tEy.data_ptr(), dt, dvx, dvy, s),
                        "vp_fused_kick_dev_f32")
                o = out.cpu().numpy().astype(np.float64)
                ulp = np.abs(expect).astype(np.float32).astype(np.float64) * 2.0 ** -23
                tol = 64 * 2.0 ** -24 * float(np.max(np.abs(g))) + 1e-15 * np.max(E)
                assert np.all(np.abs(o - expect) <= tol + ulp), (scale, kk, pf, float(np.max(np.abs(o - expect))))
        finally:
            for k, v in old.items():
                lib.vp_tune_set(L.KNOBS[k], v)


def test_deviation_trajectory_vs_fp64_oracle(device):
    """Landau run at 64 x 64 x 32 x 128, 10 steps: the deviation form tracks
    the float64 oracle to 1e-6 in the field-energy history (the plain
    float32 form only to 1e-5 here), and mass to 1e-9."""
    v2, runner, _ = _mods()
    G, f0 = _setup(_TRAJ_SHAPE, 4)
    run = runner.Run2D(G, torch.from_numpy(f0).to(device), 0.1, physics_mode=True, storage="deviation")
    energy = torch.zeros(_TRAJ_STEPS, dtype=torch.float64, device=device)
    run.advance(_TRAJ_STEPS, energy=energy)
    f_ref, W, drho = _oracle_trajectory()
    f_dev = run.state_f64().cpu().numpy()
    err_f = np.max(np.abs(f_dev - f_ref)) / np.max(np.abs(f_ref))
    err_w = np.max(np.abs(energy.cpu().numpy() - W)) / max(W)
    assert err_f <= 2e-6, err_f
    assert err_w <= 1e-6, err_w
    assert np.max(np.abs(run.delta_rho.cpu().numpy() - drho)) <= 1e-6
    assert abs(f_dev.sum() - f0.sum()) / f0.sum() <= 1e-9
    assert run.state.dtype == torch.float32


def test_vlasov2d_float32_driver_deviation_storage(device, tmp_path):
    v2, runner, _ = _mods()
    kw = dict(Nx=64, Ny=64, Nvx=32, Nvy=128, Lx=4 * math.pi, Ly=4 * math.pi, Vx_max=6.0, Vy_max=6.0, nbiter=20,
              diag_steps=10, dt=0.1, physics_mode=True, epsilon=0.05, device=device)
    assert runner.deviation_supported((64, 64, 32, 128))
    a = v2.Vlasov(dtype="float64", out_dir=str(tmp_path / "a"), **kw)
    b = v2.Vlasov(dtype="float32", fp32_storage="deviation", out_dir=str(tmp_path / "b"), **kw)
    c = v2.Vlasov(dtype="float32", fp32_storage="plain", out_dir=str(tmp_path / "c"), **kw)
    sa, sb, sc = (s.run(s.initialize()) for s in (a, b, c))
    assert sb.variables.fn.dtype == torch.float32
    eb = np.max(np.abs(b.energy_history - a.energy_history)) / np.max(a.energy_history)
    ec = np.max(np.abs(c.energy_history - a.energy_history)) / np.max(a.energy_history)
    assert eb <= 1e-6, eb
    assert eb < ec


# ---------------------------------------------------------------------------
# 1D1V FP32 storage (vp_step_1d1v_f32)
# ---------------------------------------------------------------------------

def _landau_1d(nx, nv, L=4 * math.pi, V=6.0, eps=0.05):
    x = np.arange(nx) * (L / nx)
    v = np.linspace(-V, V, nv)
    return (1 + eps * np.cos(2 * math.pi / L * x))[:, None] * np.exp(-0.5 * v ** 2)[None, :] / math.sqrt(2 * math.pi)


@pytest.mark.parametrize("nx,nv", [(64, 96), (128, 256)])
def test_fp32_1d_step_vs_fp64_oracle(device, nx, nv):
    """One float32 step against the float64 oracle step of the same (float32)
    values: f within a few float32 roundings per sweep, the float64 fields
    within the float32 rounding of the first sweep's output."""
    from vlasov_amd import vlasov1D1V as v1
    g = v1.Grid.create(nx, nv, 4 * math.pi, 6.0, device=device)
    f32 = torch.from_numpy(_landau_1d(nx, nv)).float().to(device)
    out, rho, phi, E = v1.step_with_field(f32, 0.1, g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv, True)
    assert out.dtype == torch.float32 and E.dtype == rho.dtype == phi.dtype == torch.float64
    args = (g.v.cpu().numpy(), g.inv_kx.cpu().numpy(), g.inv_kx_sq.cpu().numpy(), g.dx, g.dv)
    ref, ref_rho, _, ref_E = vo.vlasov_step_1d(f32.double().cpu().numpy(), 0.1, *args)
    scale = np.abs(ref).max()
    assert np.abs(out.double().cpu().numpy() - ref).max() <= 32 * 2.0 ** -24 * scale
    # delta rho = rho - 1 cancels: compare against the density itself
    assert np.abs(rho.cpu().numpy() - ref_rho).max() <= 16 * 2.0 ** -24
    assert np.abs(E.cpu().numpy() - ref_E).max() <= 1e-6 * np.abs(ref_E).max()
    assert torch.isfinite(phi).all()


def test_fp32_1d_free_streaming_matches_generic_sweeps(device):
    """physics_mode off: the step is X(dt/2), V with E = 0, X(dt/2), the float32
    sweeps of vp_advect_f32 bit for bit."""
    from vlasov_amd import vlasov1D1V as v1
    g = v1.Grid.create(96, 64, 4 * math.pi, 6.0, device=device)
    f32 = torch.from_numpy(_landau_1d(96, 64)).float().to(device)
    out, rho, phi, E = v1.step_with_field(f32, 0.1, g.v, g.inv_kx, g.inv_kx_sq, g.dx, g.dv, False)
    assert not E.any() and not rho.any() and not phi.any()
    a = v1.advect_rows(f32.t().contiguous(), g.v, 0.05, g.dx).t()
    z = torch.zeros(96, dtype=torch.float64, device=device)
    b = v1.advect_rows(a.contiguous(), z, 0.1, g.dv, periodic=False)
    c = v1.advect_rows(b.t().contiguous(), g.v, 0.05, g.dx).t()
    assert torch.equal(out, c)


def test_vlasov1d_float32_driver_tracks_float64(device, tmp_path):
    from vlasov_amd import vlasov1D1V as v1
    kw = dict(Nx=128, Nv=128, L=4 * math.pi, V_max=6.0, nbiter=100, diag_steps=50, dt=0.1, physics_mode=True,
              epsilon=0.01, device=device)
    a = v1.Vlasov(dtype="float64", out_dir=str(tmp_path / "a"), **kw)
    b = v1.Vlasov(dtype="float32", out_dir=str(tmp_path / "b"), **kw)
    sa = a.run(a.initialize())
    sb = b.run(b.initialize())
    assert sb.variables.fn.dtype == torch.float32 and sa.variables.fn.dtype == torch.float64
    err = np.max(np.abs(b.energy_history - a.energy_history)) / np.max(a.energy_history)
    assert err <= 1e-5, err
    ma = np.array([m[2] for m in a.moment_history])
    mb = np.array([m[2] for m in b.moment_history])
    assert np.max(np.abs(mb - ma)) / ma[0] <= 1e-6
    with pytest.raises(ValueError):
        v1.Vlasov(dtype="float32", solver=1, out_dir=str(tmp_path / "c"), **kw)
