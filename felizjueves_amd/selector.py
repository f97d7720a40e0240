"""Precision selection for the mixed build: per-work-item features, the
FP32-vs-FP64 error probe that generates training data, and the rules.

A work item is one tile of a FockBuilder launch: one bra pair and up to
`threads` ket pairs (and optionally a chunk of the ket primitive loop). Its
FP32 error is measured, not modelled: the probe runs the same work items in
both tiers, each item into its own F buffer (one FjSys entry per item, all
sharing D and Dsh), and compares. The FP64 result is the reference.
"""

from __future__ import annotations

import numpy as np

from . import native
from .features import N_FEATURES, tile_features
from .fock import FP32, FP64, FockBuilder

def item_features(fb: FockBuilder, L, w) -> np.ndarray:
    """The tile features (features.tile_features) of work item
    w = (bra, first, sys|chunk, end) of launch L."""
    b = int(w[0])
    ks = np.unique(L.ket_list[w[1]:w[3]] & 0x1ffff)
    _, dmax = fb._bounds_dmax(L.bra, L.ket, b)
    return tile_features(L.cls, L.bra.feat, L.ket.feat, b, ks, np.array([0]), dmax[ks])[0]


def probe_errors(fb: FockBuilder, D: np.ndarray, max_items: int = 4096, seed: int = 0):
    """Run (a sample of) the FP64 work items of `fb` in both tiers, each into
    its own F, and return (features [n, N_FEATURES], err [n], dE [n]) with
    err = max |F32 - F64| of the item's closed contribution and dE its
    energy error |sum D (dF + dF^T)| / 2."""
    import torch
    rng = np.random.default_rng(seed)
    if not fb._dev:
        fb.to_device()
    Dd = native.to_dev(D)
    n = fb.basis.n_ao
    items = [(L, i) for L in fb.launches if L.precision == FP64 for i in range(L.work.shape[0])]
    if len(items) > max_items:
        sel = rng.choice(len(items), size=max_items, replace=False)
        items = [items[i] for i in sorted(sel)]
    feats, errs, des = [], [], []
    native.density_blocks(fb.shell_ao[0], fb.shell_size[0], Dd, fb.Dsh)
    batch = 512
    for s0 in range(0, len(items), batch):
        chunk = items[s0:s0 + batch]
        F = {t: torch.zeros((len(chunk), n, n), dtype=torch.float64, device=Dd.device)
             for t in (FP64, FP32)}
        for tier in (FP64, FP32):
            systems = [(Dd, fb.Dsh, F[tier][j]) for j in range(len(chunk))]
            sys = native.sys_table(systems)
            # group the chunk's items by launch: one call per launch and tier
            by_launch: dict = {}
            for j, (L, i) in enumerate(chunk):
                by_launch.setdefault(id(L), (L, []))[1].append((j, i))
            for L, lst in by_launch.values():
                w = L.work[[i for _, i in lst]].copy()
                w[:, 2] = (w[:, 2] & ~0xffff) | np.array([j for j, _ in lst], dtype=np.int32)
                wd = native.to_dev(w)
                native.call_fx(L.cls, L.threads, tier, L.bra.side, L.ket.side, wd, w.shape[0],
                               sys, L.diag, fb.tau, None, L.ket_list_dev, fb._workspace(0))
        torch.cuda.synchronize()
        d = (F[FP32] - F[FP64])
        d = d + d.transpose(1, 2)
        err = d.abs().amax(dim=(1, 2)).cpu().numpy()
        de = 0.5 * (d * Dd[None]).sum(dim=(1, 2)).abs().cpu().numpy()
        for j, (L, i) in enumerate(chunk):
            feats.append(item_features(fb, L, L.work[i]))
        errs.append(err)
        des.append(de)
    return np.array(feats), np.concatenate(errs), np.concatenate(des)


def probe_builder(basis, D, tau: float = 1e-12, threads: int = 64) -> FockBuilder:
    """An all-FP64 builder whose work items are whole tiles (no primitive
    loop chunks), so each item's error is that of one tile."""
    return FockBuilder(basis, D, tau=tau, threads=threads, lane_flops=1e30)


def training_systems(n: int, seed0: int = 1000):
    """Seeded training molecules: three-atom systems of the small recipe and
    random clusters of 4-14 atoms of the working-size recipe (seeds from
    seed0, disjoint from the evaluation systems' seeds)."""
    from .systems import config1, config2
    out = []
    for i in range(n):
        s = seed0 + i
        if i % 2 == 0:
            out.append(config1(seed=s))
        else:
            n_atoms = 4 + (i * 7919) % 11
            out.append(config2(seed=s, n_atoms=n_atoms))
    return out


LOG_FLOOR = -30.0


def log_error(err: np.ndarray) -> np.ndarray:
    return np.log10(np.maximum(err, 10.0 ** LOG_FLOOR))


class LearnedSelector:
    """FP32 for a tile iff the network's predicted log10 error plus
    `margin` is below log10(budget) and the tile's features lie inside the
    training data's range (else the tile keeps the bound rule's FP64).

    margin: the `quantile` of (measured - predicted) on held-out tiles, so
    a tile the network sends to FP32 exceeds the budget with about that
    probability."""

    def __init__(self, net, budget: float, margin: float, lo: np.ndarray, hi: np.ndarray,
                 classes: set):
        self.net = net
        self.budget = float(budget)
        self.margin = float(margin)
        self.lo, self.hi = np.asarray(lo), np.asarray(hi)
        self.classes = set(classes)

    def inside(self, X: np.ndarray) -> np.ndarray:
        ok = np.all((X[:, 4:] >= self.lo[4:]) & (X[:, 4:] <= self.hi[4:]), axis=1)
        cls = [tuple(int(v) for v in r) for r in X[:, :4]]
        return ok & np.array([c in self.classes for c in cls], dtype=bool)

    def predict(self, X: np.ndarray) -> np.ndarray:
        return self.net.predict(X)

    def __call__(self, X: np.ndarray) -> np.ndarray:
        if not len(X):
            return np.zeros(0, dtype=bool)
        return (self.predict(X) + self.margin < np.log10(self.budget)) & self.inside(X)

    def save(self, path) -> None:
        from pathlib import Path
        path = Path(path)
        self.net.save(path.with_suffix(".net.npz"))
        np.savez(path.with_suffix(".sel.npz"), budget=self.budget, margin=self.margin, lo=self.lo,
                 hi=self.hi, classes=np.array(sorted(self.classes), dtype=np.int32).reshape(-1, 4))

    @classmethod
    def load(cls, path) -> "LearnedSelector":
        from pathlib import Path
        from .mlp import MLP
        path = Path(path)
        net = MLP.load(path.with_suffix(".net.npz"))
        with np.load(path.with_suffix(".sel.npz"), allow_pickle=False) as z:
            return cls(net, float(z["budget"]), float(z["margin"]), z["lo"], z["hi"],
                       {tuple(int(v) for v in r) for r in z["classes"]})


def collect(systems, max_items: int = 20000, seed: int = 0, log=None):
    """Probe every system's tiles (up to max_items each): (X, log10 err)."""
    Xs, ys = [], []
    for i, s in enumerate(systems):
        fb = probe_builder(s.basis, s.D)
        X, err, _ = probe_errors(fb, s.D, max_items=max_items, seed=seed + i)
        Xs.append(X)
        ys.append(log_error(err))
        if log:
            log(i, s, len(X))
    return np.concatenate(Xs), np.concatenate(ys)


def train_selector(X: np.ndarray, y: np.ndarray, budget: float = 1e-11, quantile: float = 0.999,
                   epochs: int = 60, seed: int = 0, log=None,
                   fit_quantile: float | None = 0.99) -> LearnedSelector:
    """Train the network on the device (90/10 split), set the margin from
    the held-out residuals of the tiles predicted within three decades of
    the budget (the ones the decision depends on), return the selector.

    With `fit_quantile` q the network is fitted with the pinball loss of q,
    so it predicts the q-quantile of a tile's log error rather than its
    mean; the margin then only covers the gap from q to `quantile` (a few
    tenths of a decade instead of several decades for the mean fit), and
    far more tiles fall under the budget at the same held-out risk.
    `fit_quantile=None` fits the mean (squared error)."""
    from .mlp import MLP
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(X))
    n_val = max(1, len(X) // 10)
    va, tr = idx[:n_val], idx[n_val:]
    net = MLP(N_FEATURES, seed=seed)
    net.fit(X[tr], y[tr], epochs=epochs, batch=4096, lr=2e-3, seed=seed, log=log,
            quantile=fit_quantile)
    pred = net.predict(X[va])
    r = y[va] - pred
    near = pred < np.log10(budget) + 3.0
    margin = float(np.quantile(r[near] if near.sum() >= 1000 else r, quantile))
    classes = {tuple(int(v) for v in row) for row in np.unique(X[:, :4], axis=0)}
    sel = LearnedSelector(net, budget, margin, X.min(0), X.max(0), classes)
    sel.val_rmse = float(np.sqrt(np.mean(r * r)))
    return sel


def main() -> None:
    """Generate training data on the device (FP32 vs FP64 per tile of seeded
    molecules), train the network with the library's kernels, save."""
    import argparse
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=200)
    ap.add_argument("--max-items", type=int, default=20000)
    ap.add_argument("--budget", type=float, default=1e-11)
    ap.add_argument("--out", required=True)
    ap.add_argument("--data", default=None, help="also save the (features, log error) pairs here")
    ap.add_argument("--from-data", default=None, help="train on pairs saved by --data instead of probing")
    ap.add_argument("--fit-quantile", type=float, default=0.99,
                    help="pinball-loss quantile of the fit (<= 0: squared error)")
    a = ap.parse_args()
    t = time.time()
    if a.from_data:
        with np.load(a.from_data, allow_pickle=False) as z:
            X, y = z["X"], z["y"]
    else:
        X, y = collect(training_systems(a.systems), max_items=a.max_items,
                       log=lambda i, s, n: print(f"[selector] probe {i} {s.name} tiles {n} "
                                                 f"{time.time() - t:.0f}s", flush=True))
    if a.data:
        np.savez(a.data, X=X, y=y)
    print(f"[selector] {len(X)} pairs in {time.time() - t:.0f}s", flush=True)
    sel = train_selector(X, y, budget=a.budget,
                         fit_quantile=a.fit_quantile if a.fit_quantile > 0 else None,
                         log=lambda e, l: print(f"[selector] epoch {e} loss {l:.4f}", flush=True))
    print(f"[selector] val rmse {sel.val_rmse:.3f} margin {sel.margin:.3f}", flush=True)
    sel.save(a.out)


if __name__ == "__main__":
    main()
