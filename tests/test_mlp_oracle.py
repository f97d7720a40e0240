"""Pins of the selector network's CPU oracle (CPU only): BF16 rounding
against torch's bfloat16 conversion, and the MSE and pinball-loss gradients
against torch autograd in float64 on the same BF16 effective weights."""

import numpy as np
import torch

from oracle import mlp_oracle as M

F = 14


def _params(seed=0):
    rng = np.random.default_rng(seed)
    n = M.H * F + M.H + M.H * M.H + M.H + M.H + 1
    return (rng.normal(0, 0.3, n)).astype(np.float32)


def test_bf16_matches_torch_round_to_nearest_even():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-30, 1000),
                        rng.normal(0, 1e30, 1000)]).astype(np.float32)
    # exact ties: low 16 bits 0x8000, with even and odd upper halves
    ties = (np.arange(0x3f800000, 0x3f800000 + 64 * 65536, 65536, dtype=np.uint32) | 0x8000)
    x = np.concatenate([x, ties.view(np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16(x).view(np.uint32), ref.view(np.uint32))


def _autograd(P, X, T, quantile):
    W1, b1, W2, b2, w3, b3 = M.unpack(P, F)
    ts = [torch.tensor(M.bf16(W1), dtype=torch.float64, requires_grad=True),
          torch.tensor(b1, dtype=torch.float64, requires_grad=True),
          torch.tensor(M.bf16(W2), dtype=torch.float64, requires_grad=True),
          torch.tensor(b2, dtype=torch.float64, requires_grad=True),
          torch.tensor(M.bf16(w3), dtype=torch.float64, requires_grad=True),
          torch.tensor(float(b3), dtype=torch.float64, requires_grad=True)]
    x = torch.tensor(X)
    h1 = torch.relu(x @ ts[0].T + ts[1])
    h2 = torch.relu(h1 @ ts[2].T + ts[3])
    y = h2 @ ts[4] + ts[5]
    # pinball loss written as max(q u, (q - 1) u), u = t - y: a form of its
    # own, not the oracle's branch on the sign of the residual
    u = torch.tensor(T) - y
    L = torch.maximum(quantile * u, (quantile - 1.0) * u).mean()
    L.backward()
    return L.item(), np.concatenate([t.grad.numpy().ravel() for t in ts])


def test_quantile_grad_matches_autograd():
    P = _params(6)
    rng = np.random.default_rng(7)
    X = rng.normal(0, 1, (400, F))
    T = rng.normal(0, 1, 400)
    for q in (0.1, 0.5, 0.99):
        loss, g = M.grad(P, X, T, F, quantile=q)
        L, ref = _autograd(P, X, T, q)
        assert abs(loss - L) < 1e-12 * max(1.0, L)
        assert np.abs(g - ref).max() < 1e-12 * np.abs(ref).max()


def test_quantile_half_is_half_absolute_error():
    """q = 1/2: the pinball loss is half the mean absolute error."""
    P = _params(8)
    rng = np.random.default_rng(9)
    X = rng.normal(0, 1, (100, F))
    T = rng.normal(0, 1, 100)
    y, _ = M.forward(P, X, F)
    loss, _ = M.grad(P, X, T, F, quantile=0.5)
    assert abs(loss - 0.5 * np.mean(np.abs(y - T))) < 1e-13


def test_grad_matches_autograd():
    P = _params(1)
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (300, F))
    T = rng.normal(0, 1, 300)
    loss, g = M.grad(P, X, T, F)
    W1, b1, W2, b2, w3, b3 = M.unpack(P, F)
    ts = [torch.tensor(M.bf16(W1), dtype=torch.float64, requires_grad=True),
          torch.tensor(b1, dtype=torch.float64, requires_grad=True),
          torch.tensor(M.bf16(W2), dtype=torch.float64, requires_grad=True),
          torch.tensor(b2, dtype=torch.float64, requires_grad=True),
          torch.tensor(M.bf16(w3), dtype=torch.float64, requires_grad=True),
          torch.tensor(float(b3), dtype=torch.float64, requires_grad=True)]
    x = torch.tensor(X)
    h1 = torch.relu(x @ ts[0].T + ts[1])
    h2 = torch.relu(h1 @ ts[2].T + ts[3])
    y = h2 @ ts[4] + ts[5]
    L = ((y - torch.tensor(T)) ** 2).mean()
    L.backward()
    ref = np.concatenate([t.grad.numpy().ravel() for t in ts])
    assert abs(loss - L.item()) < 1e-12 * max(1.0, L.item())
    assert np.abs(g - ref).max() < 1e-12 * np.abs(ref).max()


def test_bias_gradient_finite_difference():
    """Biases are FP32 (not rounded): central differences of the loss."""
    P = _params(4)
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (200, F))
    T = rng.normal(0, 1, 200)
    _, g = M.grad(P, X, T, F)
    o_b1 = M.H * F
    o_b3 = o_b1 + M.H + M.H * M.H + M.H + M.H
    for i in (o_b1, o_b1 + 7, o_b3):
        h = 1e-6
        Pp, Pm = P.astype(np.float64).copy(), P.astype(np.float64).copy()
        Pp[i] += h
        Pm[i] -= h
        # the oracle reads float32 parameters: evaluate the loss in float64
        # with the perturbed bias applied after unpacking
        def loss_at(Q):
            W1, b1, W2, b2, w3, b3 = M.unpack(P, F)
            b1 = Q[o_b1:o_b1 + M.H]
            b3 = Q[o_b3]
            W1, W2, w3 = (M.bf16(w).astype(np.float64) for w in (W1, W2, w3))
            h1 = np.maximum(X @ W1.T + b1, 0)
            h2 = np.maximum(h1 @ W2.T + b2, 0)
            return np.mean((h2 @ w3 + b3 - T) ** 2)
        fd = (loss_at(Pp) - loss_at(Pm)) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])
