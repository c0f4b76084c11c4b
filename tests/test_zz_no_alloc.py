"""The allocation-free 1D1V step (vp_step_1d1v_f64_noalloc) against the
workspace form vp_step_1d1v_f64 it wraps.

Both run the same kernels; the allocation-free form takes its scratch from a
static device array that keeps the previous call's contents, so the checks
run two problem sizes in a row (the second smaller), chained steps, and the
field-free branch, each bit for bit against the workspace form.
"""
import math

import pytest
import torch

pytestmark = pytest.mark.gpu


def _step(lib, L, mode, f, G, dt, physics_mode, s):
    nx, nv = f.shape
    out = torch.empty_like(f)
    E = torch.empty(nx, dtype=torch.float64, device=f.device)
    rho = torch.empty_like(E)
    phi = torch.empty_like(E)
    args = (f.data_ptr(), out.data_ptr(), nx, nv, dt, G.v.data_ptr(), G.inv_kx.data_ptr(), G.inv_kx_sq.data_ptr(),
            G.dx, G.dv, int(physics_mode), E.data_ptr(), rho.data_ptr(), phi.data_ptr())
    if mode == "ws":
        ws = torch.empty(lib.vp_step_1d1v_workspace(nx, nv), dtype=torch.uint8, device=f.device)
        L.check(lib.vp_step_1d1v_f64(*args, ws.data_ptr(), ws.numel(), s), "vp_step_1d1v_f64")
    else:
        L.check(lib.vp_step_1d1v_f64_noalloc(*args, s), "vp_step_1d1v_f64_noalloc")
    return out, E, rho, phi


@pytest.mark.parametrize("physics_mode", [True, False])
def test_noalloc_step_equals_workspace_step(device, physics_mode):
    from vlasov_amd import _lib, configs
    from vlasov_amd import vlasov1D1V as v1

    lib = _lib.lib()
    s = torch.cuda.current_stream().cuda_stream
    # two calls in a row, the second on a smaller problem (the scratch keeps
    # the larger call's contents), then a 3-step chain on the smaller one
    for nx, nv in ((512, 1024), (128, 128)):
        print(f"[noalloc] nx={nx} nv={nv} physics_mode={physics_mode}", flush=True)
        assert lib.vp_step_1d1v_workspace(nx, nv) <= 9 << 20
        G = v1.Grid.create(nx, nv, 4 * math.pi, 6.0, device=device, dtype=torch.float64)
        f = configs.landau_1d(G, eps=0.05)
        a = _step(lib, _lib, "ws", f, G, 0.1, physics_mode, s)
        b = _step(lib, _lib, "noalloc", f, G, 0.1, physics_mode, s)
        torch.cuda.synchronize()
        for x, y in zip(a, b):
            assert torch.equal(x, y)
    fa = fb = f
    for _ in range(3):
        fa = _step(lib, _lib, "ws", fa, G, 0.1, physics_mode, s)[0]
        fb = _step(lib, _lib, "noalloc", fb, G, 0.1, physics_mode, s)[0]
    torch.cuda.synchronize()
    assert torch.equal(fa, fb)
    if physics_mode:
        assert not torch.equal(fa, f)


def test_noalloc_step_rejects_problems_beyond_the_scratch(device):
    from vlasov_amd import _lib

    lib = _lib.lib()
    nx, nv = 1024, 2048  # 16 MiB of f: beyond the 9 MiB static scratch
    assert lib.vp_step_1d1v_workspace(nx, nv) > 9 << 20
    t = torch.zeros(8, dtype=torch.float64, device=device)
    p = t.data_ptr()
    rc = lib.vp_step_1d1v_f64_noalloc(p, p, nx, nv, 0.1, p, p, p, 0.1, 0.1, 1, p, p, p,
                                      torch.cuda.current_stream().cuda_stream)
    assert rc == -5
    assert b"scratch" in lib.vp_error_string(rc)
    assert torch.equal(t, torch.zeros_like(t))  # nothing launched
