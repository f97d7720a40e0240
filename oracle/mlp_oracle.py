"""CPU oracle of the selector network (TEST INFRASTRUCTURE: imported only by
tests/): the forward pass and the MSE-loss gradients of the library's
fj_mlp_forward / fj_mlp_backward, in float64 numpy, with the same BF16
rounding of the weights (round to nearest even) and FP32 biases.

Parameter layout: W1[64][F] b1[64] W2[64][64] b2[64] w3[64] b3.
"""

from __future__ import annotations

import numpy as np

H = 64


def bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest-even BF16 -> float32 (the kernels' fj_to_bf)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(P: np.ndarray, F: int):
    P = np.asarray(P, dtype=np.float32)
    o = 0
    W1 = P[o:o + H * F].reshape(H, F); o += H * F
    b1 = P[o:o + H]; o += H
    W2 = P[o:o + H * H].reshape(H, H); o += H * H
    b2 = P[o:o + H]; o += H
    w3 = P[o:o + H]; o += H
    b3 = P[o]
    return W1, b1, W2, b2, w3, b3


def forward(P, X, F):
    W1, b1, W2, b2, w3, b3 = unpack(P, F)
    W1, W2, w3 = (bf16(w).astype(np.float64) for w in (W1, W2, w3))
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    y = h2 @ w3 + float(b3)
    return y, (X, a1, h1, a2, h2, W2, w3)


def grad(P, X, T, F, quantile=None):
    """(loss, dL/dP) for L = mean (y - t)^2 (or, with `quantile` q, the
    pinball loss mean of q (t - y) if t >= y else (q - 1)(t - y)), the
    straight-through gradient of the FP32 master weights (d/dW of the BF16
    copy, as the kernel)."""
    y, (X, a1, h1, a2, h2, W2, w3) = forward(P, X, F)
    n = len(y)
    r = y - np.asarray(T, dtype=np.float64)
    if quantile is None:
        dy = 2.0 * r / n
        loss = float(np.mean(r * r))
    else:
        dy = np.where(r <= 0, -quantile, 1.0 - quantile) / n
        loss = float(np.mean(np.where(r <= 0, -quantile * r, (1.0 - quantile) * r)))
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    d2 = (dy[:, None] * w3[None, :]) * (a2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g
