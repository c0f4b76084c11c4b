"""GPU forward/backward of the Fourier-operator surrogate against the float64
CPU oracle (oracle/fno_oracle.py, itself pinned by finite differences).

The GPU multiplies BF16-rounded operands with FP32 accumulation, so the
comparison is to BF16 precision: relative Frobenius error <= 2e-2 for the
output and every parameter gradient but the lift weight P (<= 6e-2, its
gradient is a cancelling sum of random-sign inputs).
"""
import numpy as np
import pytest
import torch

from oracle import fno_oracle as fo

pytestmark = pytest.mark.gpu


def _setup(B=4, T=50, C=16, M=8, Hd=32, L=3, seed=7):
    from vlasov_amd import surrogate as sg
    p = fo.init_params(sg.CIN, C, M, Hd, L, seed=seed)
    for k in p:  # spectral weights large enough to matter in the check
        if k.startswith("Wr") or k.startswith("Wi"):
            p[k] = p[k] * C * C * 0.3
    rng = np.random.default_rng(seed + 1)
    X0 = rng.standard_normal((B, T, sg.CIN))
    target = rng.standard_normal((B, T))
    net = sg.FNO(T, width=C, modes=M, hidden=Hd, layers=L, params=p)
    return sg, p, X0, target, net


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_and_gradients_match_oracle(device):
    sg, p, X0, target, net = _setup()
    mats = fo.dft_mats(X0.shape[1], net.M)
    y_ref, cache = fo.forward(p, X0, mats)
    loss_ref, _ = fo.mse(y_ref, target)

    tX = torch.from_numpy(X0).float().to(device)
    tt = torch.from_numpy(target).float().to(device)
    y = net.forward(tX)
    dy = torch.empty_like(y)
    sse = net.mse(y, tt, dy)
    net.backward(dy)
    assert _rel(y.cpu().numpy(), y_ref) <= 2e-2
    assert abs(float(sse) / y.numel() - loss_ref) <= 2e-2 * loss_ref
    # the backward pass is compared on the same output gradient (the GPU's),
    # so residual cancellation in dy does not amplify forward rounding
    dy_gpu = dy.double().cpu().numpy()
    np.testing.assert_allclose(dy_gpu, 2 * (y.double().cpu().numpy() - target) / y.numel(), rtol=1e-5, atol=1e-9)
    g_ref = fo.backward(p, cache, dy_gpu, mats)
    errs = {k: _rel(net.g[k].cpu().numpy(), gr) for k, gr in g_ref.items()}
    print("relative gradient errors", {k: round(v, 4) for k, v in errs.items()})
    for k, err in errs.items():
        # dP = X0^T dH0 sums 200 products of random-sign BF16-rounded inputs:
        # the cancellation lifts its relative error above the others (~4e-2)
        assert err <= (6e-2 if k == "P" else 2e-2), (k, err)


def test_adam_training_reduces_loss(device):
    sg, p, X0, target, net = _setup(B=8, T=40, C=16, M=6, Hd=16, L=2, seed=3)
    tX = torch.from_numpy(X0).float().to(device)
    # a learnable target: a smooth function of the inputs
    tt = torch.tanh(tX[..., 0] + 0.5 * tX[..., 1]).contiguous()
    dy = torch.empty((tX.shape[0], tX.shape[1]), device=device)
    losses = []
    for _ in range(200):
        y = net.forward(tX)
        losses.append(float(net.mse(y, tt, dy)))
        net.backward(dy)
        net.adam(3e-3)
    assert losses[-1] < 0.2 * losses[0], (losses[0], losses[-1])


def test_f0_encoding_matches_oracle(device):
    """vp_nn_f0_moments on the runs' initial distributions against the float64
    oracle (itself pinned by closed forms); decoding recovers (A, u, s)."""
    from vlasov_amd import surrogate as sg
    params = sg.sample_params(16, 5)
    f0 = sg.f0_batch(params, 128, 256, device)
    m = sg.encode_f0(f0, sg.vmax_of(params)).cpu().numpy()
    ref = fo.f0_moments(f0.cpu().numpy(), sg.vmax_of(params))
    np.testing.assert_allclose(m, ref, rtol=1e-12, atol=1e-14)
    A, u, s = fo.decode_f0_moments(m)
    np.testing.assert_allclose(A, params[:, 1], rtol=1e-6)
    np.testing.assert_allclose(s, params[:, 3], rtol=2e-3)
    big = params[:, 2] > 0.5
    np.testing.assert_allclose(u[big], params[big, 2], rtol=2e-3)


def test_surrogate_routing_at_2_percent(device):
    """Two small members trained on the batched solver's histories learn
    (held-out error far below the untrained networks'), their disagreement
    tracks the error, and the routed query at tol = 0.02 serves part of a
    held-out request from the network with >= 85 % of the served answers
    within 2 % (the rule is calibrated at 95 % on other held-out runs), the
    rest from the solver; judged queries keep the solver's answers; queries
    outside the training box always go to the solver."""
    from vlasov_amd import surrogate as sg
    kw = dict(nx=64, nv=128)
    steps = 100
    ptr, pcal, pte = sg.sample_params(4096, 1), sg.sample_params(512, 3), sg.sample_params(512, 2)
    Wtr, Mtr = sg.solve(ptr, steps, return_moments=True, **kw)
    Wcal = sg.solve(pcal, steps, **kw)
    Wte = sg.solve(pte, steps, **kw)
    sur = sg.Surrogate(steps=steps, width=48, modes=16, hidden=64, layers=4, members=2)
    sur.mu, sur.sd = float(sur._log_target(Wtr).mean()), float(sur._log_target(Wtr).std())
    e_untrained = float(sg.rel_l2(sur.predict(pte), Wte).median())
    sur.fit(ptr, Wtr, Mtr, iters=6000, batch=128, lr=3e-3)
    sur.calibrate(pcal, Wcal)
    Wp, spread = sur.predict(pte, with_spread=True)
    e = sg.rel_l2(Wp, Wte).cpu().numpy()
    sp = spread.cpu().numpy()
    print("median error", float(np.median(e)), "untrained", e_untrained,
          "corr", float(np.corrcoef(np.log(sp), np.log(e))[0, 1]))
    assert float(np.median(e)) < e_untrained / 5
    assert np.corrcoef(np.log(sp), np.log(e))[0, 1] > 0.5
    W, info = sur.query(pte, tol=0.02, judge_fraction=0.05, **kw)
    served = info.pop("served_mask")
    print("query", info)
    assert info["served"] > 0 and info["served"] == int(served.sum())
    assert info["served"] + info["solved"] == len(pte)
    eq = sg.rel_l2(W, Wte).cpu().numpy()
    assert np.mean(eq[served] <= 0.02) >= 0.85
    # every answer not served by the network is the solver's, bit for bit
    assert torch.equal(W[torch.from_numpy(~served).to(W.device)], Wte[torch.from_numpy(~served).to(W.device)])
    np.testing.assert_allclose(W[torch.from_numpy(served).to(W.device)].cpu().numpy(),
                               Wp[torch.from_numpy(served).to(W.device)].cpu().numpy(), rtol=1e-12)
    # a tolerance no calibration run meets: everything goes to the solver
    W, info = sur.query(pte[:64], tol=1e-9, **kw)
    assert info["served"] == 0 and torch.equal(W, Wte[:64])
    # parameters outside the training box always go to the solver
    out = pte[:3].copy()
    out[:, 0] = 0.9
    W, info = sur.query(out, tol=1e9, **kw)
    assert info["outside_box"] == 3 and info["served"] == 0
    assert torch.equal(W, sg.solve(out, steps, **kw))
