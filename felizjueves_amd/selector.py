"""Precision selection for the mixed build: per-work-item features, the
FP32-vs-FP64 error probe that generates training data, and the rules.

A work item is one tile of a FockBuilder launch: one bra pair and up to
`threads` ket pairs (and optionally a chunk of the ket primitive loop). Its
FP32 error is measured, not modelled: the probe runs the same work items in
both tiers, each item into its own F buffer (one FjSys entry per item, all
sharing D and Dsh), and compares. The FP64 result is the judge.
"""

from __future__ import annotations

import math

import numpy as np

from . import native
from .fock import FP32, FP64, FockBuilder

N_FEATURES = 14


def item_features(fb: FockBuilder, L, w) -> np.ndarray:
    """Features of work item w = (bra, first, sys|chunk, end) of launch L:
    [la, lb, lc, ld, log10 min/max bra exponent sum, log10 min/max ket
    exponent sum, |AB|, max |CD|, min |PQ| (pair centres), log10 G_ab,
    log10 max G_cd, log10 max |D| over the tile's six blocks]."""
    b = int(w[0])
    ks = L.ket_list[w[1]:w[3]]
    bra, ket = L.bra, L.ket
    pb = bra.pairs[b]
    zb = pb.rec[:, 6]
    zk = np.concatenate([ket.pairs[k].rec[:, 6] for k in ks])
    basis = fb.basis
    A, B = basis.shell_centre[pb.a], basis.shell_centre[pb.b]
    ab = float(np.linalg.norm(A - B))
    cd = max(float(np.linalg.norm(basis.shell_centre[ket.pairs[k].a] -
                                  basis.shell_centre[ket.pairs[k].b])) for k in ks)
    P = (pb.rec[:, :3] * zb[:, None]).sum(0) / zb.sum()
    pq = min(float(np.linalg.norm(P - (ket.pairs[k].rec[:, :3] * ket.pairs[k].rec[:, 6:7]).sum(0)
                                  / ket.pairs[k].rec[:, 6].sum())) for k in ks)
    Gk = max(ket.G[k] for k in ks)
    Dsh = fb.Dsh_plan
    sa, sb = pb.a, pb.b
    dm = 0.0
    for k in ks:
        sc, sd = ket.pairs[k].a, ket.pairs[k].b
        dm = max(dm, Dsh[sa, sb], Dsh[sc, sd], Dsh[sa, sc], Dsh[sa, sd], Dsh[sb, sc], Dsh[sb, sd])
    lg = lambda x: math.log10(max(x, 1e-300))  # noqa: E731
    return np.array([*L.cls, lg(zb.min()), lg(zb.max()), lg(zk.min()), lg(zk.max()), ab, cd, pq,
                     lg(bra.G[b]), lg(Gk), lg(dm)], dtype=np.float64)


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
    native.density_blocks(fb.shell_ao, fb.shell_size, Dd, fb.Dsh)
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
                native.call_fx(L.cls, fb.threads, tier, L.bra.side, L.ket.side, wd, w.shape[0],
                               sys, L.diag, fb.tau, None, L.ket_list_dev)
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
