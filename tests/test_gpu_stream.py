"""The STREAM-style HBM calibration kernels (vp_stream_f64) give correct results.

The bench's measured copy ceiling comes from kind 0; kinds 1 (read, one
partial sum per workgroup) and 2 (write) are the other two STREAM patterns.
The reference's STREAM tests check the same properties of its kernels
(simulations/stream/pytorch/tests/test_kernels.py: copied values, source left
unchanged, set fills the value).  Integer-valued data make every partial sum
exact, so the read kernel's sums are compared bit for bit with the
grid-stride partition recomputed here.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SENTINEL = -12345.0


def _lib():
    from vlasov_amd import _lib
    return _lib, _lib.lib()


@pytest.mark.parametrize("unroll", [1, 2, 4, 8])
@pytest.mark.parametrize("nontemporal", [0, 1])
@pytest.mark.parametrize("n,grid", [(2, 1), (1 << 20, 1024), (1000002, 37), (3 << 16, 5000)])
def test_copy_read_write(device, unroll, nontemporal, n, grid):
    L, lib = _lib()
    s = torch.cuda.current_stream().cuda_stream
    a_host = (np.arange(n + 8) % 997).astype(np.float64) - 300.0
    a = torch.from_numpy(a_host).to(device)
    b = torch.full((n + 8,), SENTINEL, dtype=torch.float64, device=device)
    L.check(lib.vp_stream_f64(0, a.data_ptr(), b.data_ptr(), n, unroll, nontemporal, grid, s), "copy")
    bh = b.cpu().numpy()
    np.testing.assert_array_equal(bh[:n], a_host[:n])
    assert np.all(bh[n:] == SENTINEL)
    np.testing.assert_array_equal(a.cpu().numpy(), a_host)  # source unchanged

    part = torch.full((grid + 4,), SENTINEL, dtype=torch.float64, device=device)
    L.check(lib.vp_stream_f64(1, a.data_ptr(), part.data_ptr(), n, unroll, nontemporal, grid, s), "read")
    ph = part.cpu().numpy()
    # 16-byte element i (doubles 2i, 2i+1) belongs to workgroup ((i // 256) // unroll) % grid
    pairs = a_host[:n].reshape(-1, 2).sum(axis=1)
    owner = ((np.arange(n // 2) // 256) // unroll) % grid
    want = np.bincount(owner, weights=pairs, minlength=grid)
    np.testing.assert_array_equal(ph[:grid], want)
    assert np.all(ph[grid:] == SENTINEL)
    assert float(ph[:grid].sum()) == float(a_host[:n].sum())

    w = torch.full((n + 8,), SENTINEL, dtype=torch.float64, device=device)
    L.check(lib.vp_stream_f64(2, None, w.data_ptr(), n, unroll, nontemporal, grid, s), "write")
    wh = w.cpu().numpy()
    assert np.all(wh[:n] == 1.0) and np.all(wh[n:] == SENTINEL)


def test_copy_of_the_bench_size(device):
    """The bench's calibration launch: two 128^4 float64 arrays, one 16-byte
    element per thread, no grid-stride loop (grid = n / 2 / 256)."""
    L, lib = _lib()
    n = 128 ** 4
    a = torch.randn(n, dtype=torch.float64, device=device)
    b = torch.empty_like(a)
    L.check(lib.vp_stream_f64(0, a.data_ptr(), b.data_ptr(), n, 1, 1, n // 2 // 256,
                              torch.cuda.current_stream().cuda_stream), "copy")
    assert torch.equal(a, b)
