"""GPU: the selector network's forward, backward and Adam kernels against
the CPU oracle (oracle/mlp_oracle.py). Bars: forward |dy| <= 1e-5 max|y|
(FP32 accumulation vs float64), gradients <= 1e-4 max|g| (FP32 atomics in
any order), BF16 copies bit-exact."""

import numpy as np
import pytest

from oracle import mlp_oracle as O

pytestmark = pytest.mark.gpu

F = 14


def _setup(n=1500, seed=0):
    import torch
    from felizjueves_amd import native
    from felizjueves_amd.mlp import MLP
    rng = np.random.default_rng(seed)
    net = MLP(F, seed=seed)
    X = rng.normal(0, 1, (n, F)).astype(np.float32)
    T = rng.normal(-9, 2, n).astype(np.float32)
    return torch, native, net, X, T


def test_bf16_copy_bit_exact():
    torch, native, net, X, T = _setup()
    torch.cuda.synchronize()
    got = net.Wb.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    ref = O.bf16(net.params()).view(np.uint32)
    assert np.array_equal(got, ref)


def test_forward_vs_oracle():
    torch, native, net, X, T = _setup()
    y = net.forward_dev(native.to_dev(X))
    torch.cuda.synchronize()
    ref, _ = O.forward(net.params(), X, F)
    err = np.abs(y.cpu().numpy() - ref).max()
    print("forward max|dy|", err, flush=True)
    assert err <= 1e-5 * np.abs(ref).max()


def test_backward_vs_oracle():
    torch, native, net, X, T = _setup(n=3000, seed=1)
    net.grad(native.to_dev(X), native.to_dev(T))
    torch.cuda.synchronize()
    loss, g = O.grad(net.params(), X, T, F)
    gd = net.G.cpu().numpy()
    err = np.abs(gd - g).max()
    print("loss", net.loss.item(), loss, "grad max|dg|", err, "max|g|", np.abs(g).max(), flush=True)
    assert abs(net.loss.item() - loss) <= 1e-5 * loss
    assert err <= 1e-4 * np.abs(g).max()


def test_quantile_backward_vs_oracle():
    torch, native, net, X, T = _setup(n=3000, seed=4)
    # shift the targets to straddle the predictions (residuals of both signs)
    y0, _ = O.forward(net.params(), X, F)
    T = (y0 + np.random.default_rng(5).normal(0, 1, len(y0))).astype(np.float32)
    for q in (0.99, 0.3):
        net.grad(native.to_dev(X), native.to_dev(T), quantile=q)
        torch.cuda.synchronize()
        loss, g = O.grad(net.params(), X, T, F, quantile=q)
        gd = net.G.cpu().numpy()
        err = np.abs(gd - g).max()
        print("q", q, "loss", net.loss.item(), loss, "grad max|dg|", err, flush=True)
        assert abs(net.loss.item() - loss) <= 1e-5 * loss
        assert err <= 1e-4 * np.abs(g).max()


def test_adam_step_vs_numpy():
    torch, native, net, X, T = _setup(n=800, seed=2)
    P0 = net.params().astype(np.float64)
    Xd, Td = native.to_dev(X), native.to_dev(T)
    net.grad(Xd, Td)
    torch.cuda.synchronize()
    g = net.G.cpu().numpy().astype(np.float64)
    net.step(1e-3)
    torch.cuda.synchronize()
    m = 0.1 * g
    v = 0.001 * g * g
    ref = P0 - 1e-3 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    P1 = net.params()
    assert np.abs(P1 - ref).max() <= 1e-6
    got = net.Wb.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    assert np.array_equal(got, O.bf16(P1).view(np.uint32))


def test_training_fits_a_smooth_target():
    """Minibatch Adam with the device kernels fits y = sum sin(x) to a small
    fraction of its variance."""
    torch, native, net, X, T = _setup(n=20000, seed=3)
    y = np.sin(X[:, :4]).sum(1) + 0.5 * X[:, 5] * X[:, 6]
    hist = net.fit(X, y, epochs=30, batch=1024, lr=3e-3)
    pred = net.predict(X)
    resid = np.mean((pred - y) ** 2) / np.var(y)
    print("loss", hist[0], "->", hist[-1], "relative residual", resid, flush=True)
    assert resid < 0.05
