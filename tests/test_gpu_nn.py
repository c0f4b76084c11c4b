"""GPU checks of the surrogate's kernels against float64 NumPy references.

The GEMM rounds both operands to bfloat16 and accumulates in float32, so
its reference is the float64 product of the bfloat16-rounded operands
(agreement to float32 accumulation error); the transposed and batched forms
are the same kernel with other strides.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _bf16(x):
    return torch.from_numpy(np.asarray(x, dtype=np.float32)).to(torch.bfloat16).to(torch.float64).numpy()


def _lib():
    from vlasov_amd import _lib
    return _lib


@pytest.mark.parametrize("batch,M,N,K", [(1, 64, 64, 16), (3, 70, 33, 50), (2, 1, 200, 129), (5, 130, 64, 512)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_matches_bf16_reference(device, batch, M, N, K, ta, tb):
    L = _lib()
    rng = np.random.default_rng(batch * 1000 + M + N + K)
    A = rng.standard_normal((batch, K, M) if ta else (batch, M, K)).astype(np.float32)
    B = rng.standard_normal((batch, N, K) if tb else (batch, K, N)).astype(np.float32)
    C0 = rng.standard_normal((batch, M, N)).astype(np.float32)
    Ar = _bf16(A).transpose(0, 2, 1) if ta else _bf16(A)
    Br = _bf16(B).transpose(0, 2, 1) if tb else _bf16(B)
    ref = 0.5 * np.einsum("bik,bkj->bij", Ar, Br) - 2.0 * C0
    tA, tB = torch.from_numpy(A).to(device), torch.from_numpy(B).to(device)
    tC = torch.from_numpy(C0).to(device)
    sA = (M * K, 1, M) if ta else (M * K, K, 1)
    sB = (K * N, 1, K) if tb else (K * N, N, 1)
    s = torch.cuda.current_stream().cuda_stream
    L.check(L.lib().vp_nn_gemm(batch, M, N, K, 0.5, tA.data_ptr(), *sA, tB.data_ptr(), *sB, -2.0, tC.data_ptr(),
                               M * N, N, 1, s), "gemm")
    out = tC.cpu().numpy()
    scale = np.abs(Ar).max() * np.abs(Br).max() * K
    assert np.max(np.abs(out - ref)) <= 1e-6 * scale + 1e-5


def test_gemm_beta_zero_ignores_garbage(device):
    L = _lib()
    A = torch.ones((64, 32), device=device)
    B = torch.ones((32, 16), device=device)
    C = torch.full((64, 16), float("nan"), device=device)
    L.check(L.lib().vp_nn_gemm(1, 64, 16, 32, 1.0, A.data_ptr(), 0, 32, 1, B.data_ptr(), 0, 16, 1, 0.0, C.data_ptr(),
                               0, 16, 1, torch.cuda.current_stream().cuda_stream), "gemm")
    assert torch.equal(C, torch.full_like(C, 32.0))


def test_elementwise_kernels(device):
    L = _lib()
    lib = L.lib()
    s = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(0)
    x = rng.standard_normal(1000).astype(np.float32) * 3
    dy = rng.standard_normal(1000).astype(np.float32)
    tx, tdy = torch.from_numpy(x).to(device), torch.from_numpy(dy).to(device)
    y, dx = torch.empty_like(tx), torch.empty_like(tx)
    L.check(lib.vp_nn_gelu(tx.data_ptr(), y.data_ptr(), 1000, s), "gelu")
    L.check(lib.vp_nn_gelu_bwd(tx.data_ptr(), tdy.data_ptr(), dx.data_ptr(), 1000, s), "gelu_bwd")
    from scipy.special import erf
    xd = x.astype(np.float64)
    np.testing.assert_allclose(y.cpu().numpy(), 0.5 * xd * (1 + erf(xd / np.sqrt(2))), rtol=2e-6, atol=2e-6)
    d = 0.5 * (1 + erf(xd / np.sqrt(2))) + xd * np.exp(-0.5 * xd * xd) / np.sqrt(2 * np.pi)
    np.testing.assert_allclose(dx.cpu().numpy(), dy * d, rtol=2e-6, atol=2e-6)
    # bias add / grad
    Y = torch.from_numpy(rng.standard_normal((300, 70)).astype(np.float32)).to(device)
    b = torch.from_numpy(rng.standard_normal(70).astype(np.float32)).to(device)
    Y0 = Y.clone()
    L.check(lib.vp_nn_bias_add(Y.data_ptr(), b.data_ptr(), 300, 70, s), "bias_add")
    assert torch.allclose(Y, Y0 + b[None, :])
    db = torch.empty(70, device=device)
    L.check(lib.vp_nn_bias_grad(Y0.data_ptr(), db.data_ptr(), 300, 70, s), "bias_grad")
    np.testing.assert_allclose(db.cpu().numpy(), Y0.cpu().numpy().astype(np.float64).sum(0), rtol=1e-4, atol=1e-4)
    # mse
    p = torch.from_numpy(rng.standard_normal(1000).astype(np.float32)).to(device)
    tgt = torch.from_numpy(rng.standard_normal(1000).astype(np.float32)).to(device)
    g = torch.empty_like(p)
    part = torch.empty(4, device=device)
    L.check(lib.vp_nn_mse(p.data_ptr(), tgt.data_ptr(), g.data_ptr(), part.data_ptr(), 1000, s), "mse")
    diff = p.cpu().numpy().astype(np.float64) - tgt.cpu().numpy()
    assert abs(float(part.sum()) - float(np.sum(diff ** 2))) <= 1e-3
    np.testing.assert_allclose(g.cpu().numpy(), 2 * diff / 1000, rtol=1e-5, atol=1e-8)
    # adam: one step from m = v = 0 moves every weight by lr * sign(g) (up to eps)
    w = torch.zeros(50, device=device)
    gw = torch.from_numpy(rng.standard_normal(50).astype(np.float32)).to(device)
    m, v = torch.zeros_like(w), torch.zeros_like(w)
    L.check(lib.vp_nn_adam(w.data_ptr(), gw.data_ptr(), m.data_ptr(), v.data_ptr(), 50, 0.01, 0.9, 0.999, 1e-8, 1, s),
            "adam")
    np.testing.assert_allclose(w.cpu().numpy(), -0.01 * np.sign(gw.cpu().numpy()), rtol=1e-4)


def test_weighted_squared_error(device):
    L = _lib()
    lib = L.lib()
    s = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(4)
    n = 1537
    p = rng.standard_normal(n).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    w = rng.random(n).astype(np.float32)
    tp, tt, tw = (torch.from_numpy(a).to(device) for a in (p, t, w))
    grad = torch.empty_like(tp)
    part = torch.empty((n + 255) // 256, device=device)
    L.check(lib.vp_nn_wmse(tp.data_ptr(), tt.data_ptr(), tw.data_ptr(), grad.data_ptr(), part.data_ptr(), n, s),
            "wmse")
    d = p.astype(np.float64) - t
    np.testing.assert_allclose(float(part.sum()), np.sum(w * d * d), rtol=1e-5)
    np.testing.assert_allclose(grad.cpu().numpy(), 2 * w * d, rtol=1e-6, atol=1e-7)
    # w = 1/n: the gradient of vp_nn_mse, whose partial sums are n times these
    tw.fill_(1.0 / n)
    g2, part2 = torch.empty_like(tp), torch.empty_like(part)
    L.check(lib.vp_nn_wmse(tp.data_ptr(), tt.data_ptr(), tw.data_ptr(), grad.data_ptr(), part.data_ptr(), n, s),
            "wmse")
    L.check(lib.vp_nn_mse(tp.data_ptr(), tt.data_ptr(), g2.data_ptr(), part2.data_ptr(), n, s), "mse")
    np.testing.assert_allclose(grad.cpu().numpy(), g2.cpu().numpy(), rtol=1e-6)
    np.testing.assert_allclose(float(part.sum()) * n, float(part2.sum()), rtol=1e-5)
