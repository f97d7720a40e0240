"""The precision selector's network on the device: a 2 x 64 ReLU MLP with
BF16 weights and FP32 accumulation, trained and evaluated by the library's
own kernels (csrc/fj_mlp.hip): fj_mlp_forward, fj_mlp_backward (MSE loss,
gradients of the FP32 master weights), fj_mlp_adam.

Inputs are standardised on the host with the training set's mean and
standard deviation (kept with the weights); the target is log10 of the
measured FP32-vs-FP64 error of a work item.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import native

H = 64


def _call(name, *args):
    fn = getattr(native.lib(), name)
    fn.restype = C.c_int
    vals = [C.c_void_p(native.stream_handle())]
    for a in args:
        vals.append(C.c_void_p(a.data_ptr()) if hasattr(a, "data_ptr") else a)
    native.check(fn(*vals), name)


def nparams(F: int) -> int:
    return H * F + H + H * H + H + H + 1


def init_params(F: int, seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases, in the library's parameter layout."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0.0, np.sqrt(2.0 / F), (H, F))
    W2 = rng.normal(0.0, np.sqrt(2.0 / H), (H, H))
    w3 = rng.normal(0.0, np.sqrt(1.0 / H), H)
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), w3, [0.0]]).astype(np.float32)


class MLP:
    def __init__(self, F: int, seed: int = 0, params: np.ndarray | None = None,
                 mean: np.ndarray | None = None, std: np.ndarray | None = None):
        import torch
        self.F = F
        self.np = nparams(F)
        dev = native.device()
        p = init_params(F, seed) if params is None else np.asarray(params, dtype=np.float32)
        if p.shape != (self.np,):
            raise ValueError(f"parameter block of {p.shape}, want ({self.np},)")
        self.P = native.to_dev(p)
        self.Wb = torch.zeros(self.np, dtype=torch.int16, device=dev)
        self.M = torch.zeros(self.np, dtype=torch.float32, device=dev)
        self.V = torch.zeros(self.np, dtype=torch.float32, device=dev)
        self.G = torch.zeros(self.np, dtype=torch.float32, device=dev)
        self.loss = torch.zeros(1, dtype=torch.float32, device=dev)
        self.t = 0
        self.mean = np.zeros(F) if mean is None else np.asarray(mean, dtype=np.float64)
        self.std = np.ones(F) if std is None else np.asarray(std, dtype=np.float64)
        _call("fj_mlp_round", C.c_int(self.np), self.P, self.Wb)

    def normalise(self, X: np.ndarray) -> np.ndarray:
        return ((np.asarray(X, dtype=np.float64) - self.mean) / self.std).astype(np.float32)

    def forward_dev(self, Xd):
        import torch
        Y = torch.empty(Xd.shape[0], dtype=torch.float32, device=Xd.device)
        _call("fj_mlp_forward", C.c_int(Xd.shape[0]), C.c_int(self.F), Xd, self.Wb, self.P, Y)
        return Y

    def predict(self, X: np.ndarray) -> np.ndarray:
        """Predicted log10 FP32 error for raw features X [n, F]."""
        import torch
        if len(X) == 0:
            return np.zeros(0)
        Xd = native.to_dev(self.normalise(X))
        Y = self.forward_dev(Xd)
        torch.cuda.synchronize()
        return Y.cpu().numpy().astype(np.float64)

    def grad(self, Xd, Td, quantile: float | None = None):
        """Zero G and loss, then accumulate dL/dP and L on the device (mean
        squared error, or the pinball loss of `quantile`)."""
        self.G.zero_()
        self.loss.zero_()
        if quantile is None:
            _call("fj_mlp_backward", C.c_int(Xd.shape[0]), C.c_int(self.F), Xd, Td, self.Wb, self.P,
                  self.G, self.loss)
        else:
            _call("fj_mlp_backward_quantile", C.c_int(Xd.shape[0]), C.c_int(self.F), Xd, Td, self.Wb,
                  self.P, self.G, self.loss, C.c_float(quantile))

    def step(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> None:
        self.t += 1
        _call("fj_mlp_adam", C.c_int(self.np), self.P, self.G, self.M, self.V, self.Wb,
              C.c_float(lr), C.c_float(b1), C.c_float(b2), C.c_float(eps), C.c_int(self.t))

    def fit(self, X: np.ndarray, y: np.ndarray, epochs: int = 60, batch: int = 4096,
            lr: float = 2e-3, seed: int = 0, log=None, quantile: float | None = None) -> list[float]:
        """Minibatch Adam on (standardised X, y) with the device kernels
        (mean squared error, or the pinball loss of `quantile`)."""
        import torch
        self.mean = X.mean(axis=0)
        self.std = X.std(axis=0) + 1e-6
        Xd = native.to_dev(self.normalise(X))
        Td = native.to_dev(np.asarray(y, dtype=np.float32))
        n = Xd.shape[0]
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            perm = native.to_dev(rng.permutation(n).astype(np.int64))
            tot = 0.0
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                xb, tb = Xd[idx].contiguous(), Td[idx].contiguous()
                self.grad(xb, tb, quantile)
                self.step(lr * (0.5 ** (ep // max(1, epochs // 3))))
                tot += float(self.loss.item()) * len(idx)
            hist.append(tot / n)
            if log:
                log(ep, hist[-1])
        torch.cuda.synchronize()
        return hist

    def params(self) -> np.ndarray:
        return self.P.cpu().numpy()

    def save(self, path: Path) -> None:
        np.savez(path, P=self.params(), mean=self.mean, std=self.std, F=self.F)

    @classmethod
    def load(cls, path: Path) -> "MLP":
        with np.load(path, allow_pickle=False) as z:
            return cls(int(z["F"]), params=z["P"], mean=z["mean"], std=z["std"])
