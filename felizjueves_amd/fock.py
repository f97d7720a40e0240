"""The native Fock builder: ``G = J - K/2`` on one MI355X.

Host planning, once per basis (and density screen):

* every canonical shell pair, its primitive-pair bounds, pruning and staged
  records (`pairs.kept_pairs`, the upstream rules);
* per pair class (la, lb) ONE ragged pair table on the device -- pairs sorted
  by (degree, G) descending, each pair's records contiguous, an offset
  array, no degree-uniform batching (so a class pair is one launch, not one
  per (degree, magnitude) bin pair);
* per class pair (X <= Y) a work list of tiles (bra pair, first ket): one
  workgroup each, ``threads`` consecutive kets; tiles whose every quartet
  fails the six-block screen ``G_ab G_cd dmax < tau`` are dropped here, and
  (mixed mode) a tile whose largest bound is below the FP32 threshold goes
  to the FP32-tier kernel instead of the FP64 one.

Per build (device only): zero F, the density block maxima, one launch per
(class pair, tier) on one stream, close F <- F + F^T.

The screen, weights and degeneracy are the upstream ones
(`eri/fock/kernel_f.py:1-165`), so with the same pruning the result equals
the upstream pair-stream walk up to the eq. 29 pre-launch screen, which
only removes quartets with G_ab G_cd < tau.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import native
from . import pairs as P
from .basis import Basis
from .features import side_arrays, tile_features
from .hgp import cost

FP64, FP32 = 0, 1

# The FP32 tier's error rule (the fallback of the learned selector): the
# FP32 VRR's error is not bounded by the final integrals' Schwarz bound
# G_ab G_cd dmax. The HRR's all-on-one-centre boundary (e0|f0) exceeds the
# final (ab|cd) by about |PA|^|e| for separated centres, and tight, diffuse
# primitive pairs mix magnitudes. So the estimate multiplies the bound by
# (1 + |AB|)^(la+lb) (1 + |CD|)^(lc+ld) sqrt(zeta_max/zeta_min) per side and by
# FP32_EST_SCALE; measured FP32-vs-FP64 errors of 6,469 tiles (3, 10 and
# 20-atom clusters) all lie at least 10^3.4 below this estimate.
FP32_EST_SCALE = 1e-3


def fp32_amplification(basis, p) -> float:
    """(1 + |A - B|)^(la + lb) sqrt(zeta_max / zeta_min) of one pair."""
    d = float(np.linalg.norm(basis.shell_centre[p.a] - basis.shell_centre[p.b]))
    z = p.rec[:, 6]
    return (1.0 + d) ** (p.la + p.lb) * math.sqrt(z.max() / z.min())
FOCK_LMAX = 2  # the top class built: (dd|dd)


@dataclass
class SideTable:
    cls: tuple
    pairs: list               # FPair, in table order
    G: np.ndarray
    deg: np.ndarray
    sh: np.ndarray            # [n, 2]
    dev: dict = field(default_factory=dict)
    side: native.FjSide = None

    @property
    def n(self) -> int:
        return len(self.pairs)


@dataclass
class Launch:
    cls: tuple
    bra: SideTable
    ket: SideTable
    diag: int
    precision: int
    work: np.ndarray          # [n, 2] int32
    work_dev: object = None
    ket_list: np.ndarray = None
    ket_list_dev: object = None
    n_quartets: int = 0       # computed (post-screen) contracted quartets
    n_prim: int = 0           # primitive quartets of those
    flops: float = 0.0        # algorithmic FLOPs of those


class FockBuilder:
    def __init__(self, basis: Basis, D: np.ndarray, *, tau: float = 1e-12, threads: int = 64,
                 mixed: bool = False, fp32_error: float = 1e-11, selector=None,
                 lane_flops: float = 1.0e6):
        if basis.l_max > FOCK_LMAX:
            raise ValueError(f"shells up to l = {basis.l_max}: the fused F kernels are built through "
                             f"l = {FOCK_LMAX} ((dd|dd)); f classes are not built yet")
        self.basis = basis
        self.tau = float(tau)
        self.threads = int(threads)
        self.mixed = bool(mixed)
        self.fp32_error = float(fp32_error)
        self.lane_flops = float(lane_flops)
        t0 = time.time()
        kept, self.deleted, self.eps, self.g_max, self.n_prim_total, _ = P.kept_pairs(
            basis, D, self.tau, 1)
        big = [p for p in kept if p.degree > P.MAX_PRIM_PAIRS]
        if big:
            raise ValueError(f"{len(big)} pairs keep more than {P.MAX_PRIM_PAIRS} primitive "
                             "pairs, the staging cap of the F kernels")
        by_cls: dict = {}
        for p in kept:
            by_cls.setdefault((p.la, p.lb), []).append(p)
        self.tables = {}
        for c, ps in sorted(by_cls.items()):
            ps = sorted(ps, key=lambda p: (-p.degree, -p.G))
            self.tables[c] = SideTable(
                c, ps, np.array([p.G for p in ps]), np.array([p.degree for p in ps]),
                np.array([(p.a, p.b) for p in ps], dtype=np.int64))
            self.tables[c].amp = np.array([fp32_amplification(basis, p) for p in ps])
            self.tables[c].feat = side_arrays(basis, ps)
        self.Dsh_plan = P.density_blocks(basis, D)
        # selector(features [n, N_FEATURES]) -> bool [n]: True moves an FP64
        # tile of the bound rule to the FP32 tier (mixed mode only)
        self.selector = selector
        self.launches = self._plan()
        self.plan_seconds = time.time() - t0
        self._dev = False

    # -- planning ---------------------------------------------------------
    def _bounds(self, bra: SideTable, ket: SideTable, b: int) -> np.ndarray:
        return self._bounds_dmax(bra, ket, b)[0]

    def _bounds_dmax(self, bra: SideTable, ket: SideTable, b: int):
        """(G_ab G_cd dmax, dmax) of bra pair b against every ket pair."""
        Dsh = self.Dsh_plan
        sa, sb = bra.sh[b]
        sc, sd = ket.sh[:, 0], ket.sh[:, 1]
        dmax = np.maximum.reduce([
            np.full(ket.n, 2.0 * Dsh[sa, sb]), 2.0 * Dsh[sc, sd], 0.5 * Dsh[sa, sc],
            0.5 * Dsh[sa, sd], 0.5 * Dsh[sb, sc], 0.5 * Dsh[sb, sd]])
        return bra.G[b] * (ket.G * dmax), dmax

    def _tiles(self) -> list:
        """Per class pair: the live (bra, ket) quartets of each bra, split by
        the bound rule into the FP32 set (mixed mode) and the rest, each cut
        into tiles of `threads` kets in table order (degree, then G,
        descending). Returns [(X, Y, [(b, kets, tier)], features of the FP64
        tiles or None)]."""
        T = self.threads
        classes = sorted(self.tables)
        want_feats = self.mixed and self.selector is not None
        out = []
        for i, X in enumerate(classes):
            for Y in classes[i:]:
                bra, ket = self.tables[X], self.tables[Y]
                diag = X == Y
                tiles, feats = [], []
                for b in range(bra.n):
                    bound, dmax = self._bounds_dmax(bra, ket, b)
                    live = bound >= self.tau
                    if diag:
                        live[:b] = False
                    ks = np.nonzero(live)[0]
                    if not len(ks):
                        continue
                    if self.mixed:
                        est = bound[ks] * bra.amp[b] * ket.amp[ks] * FP32_EST_SCALE
                        lo = est < self.fp32_error
                        parts = ((FP64, ks[~lo]), (FP32, ks[lo]))
                    else:
                        parts = ((FP64, ks),)
                    for tier, kt in parts:
                        if not len(kt):
                            continue
                        starts = np.arange(0, len(kt), T)
                        for t0 in starts:
                            tiles.append((b, kt[t0:t0 + T], tier))
                        if want_feats and tier == FP64:
                            feats.append(tile_features((*X, *Y), bra.feat, ket.feat, b, kt, starts,
                                                       dmax[kt]))
                f = np.concatenate(feats) if feats else None
                out.append((X, Y, tiles, f))
        return out

    def _plan(self) -> list:
        """Tiles -> tiers (bound rule, then the learned selector on the
        FP64 tiles) -> one work item per tile (or per primitive-loop chunk)
        per (class pair, tier) launch."""
        T = self.threads
        groups = self._tiles()
        if self.mixed and self.selector is not None:
            allf = [f for *_, f in groups if f is not None]
            if allf:
                to32 = self.selector(np.concatenate(allf))
                k = 0
                for X, Y, tiles, f in groups:
                    if f is None:
                        continue
                    sel = to32[k:k + len(f)]
                    k += len(f)
                    j = 0
                    for t, (b, seg, tier) in enumerate(tiles):
                        if tier == FP64:
                            if sel[j]:
                                tiles[t] = (b, seg, FP32)
                            j += 1
        out = []
        for X, Y, tiles, _ in groups:
            bra, ket = self.tables[X], self.tables[Y]
            cls = (*X, *Y)
            c = cost(cls)
            diag = int(X == Y)
            work = {FP64: [], FP32: []}
            cost_w = {FP64: [], FP32: []}
            lists = {FP64: [], FP32: []}
            n_list = {FP64: 0, FP32: 0}
            stats = {FP64: [0, 0, 0.0], FP32: [0, 0, 0.0]}
            for b, seg, tr in tiles:
                kab = int(bra.deg[b])
                # a lane's primitive loop is cut into chunks of at most
                # lane_flops / per_prim primitive quartets, so no work item
                # runs much longer than the others
                chunk = max(1, min(255, max(8, int(self.lane_flops // c.per_prim)) // kab))
                st = n_list[tr]
                lists[tr].append(seg)
                n_list[tr] += len(seg)
                kmax = int(ket.deg[seg].max())
                if kmax <= chunk:
                    work[tr].append((b, st, 0, st + len(seg)))
                    cost_w[tr].append(kab * kmax)
                else:
                    for kc in range(0, kmax, chunk):
                        work[tr].append((b, st, (kc << 16) | (chunk << 24), st + len(seg)))
                        cost_w[tr].append(kab * min(chunk, kmax - kc))
                nprim = int(kab * ket.deg[seg].sum())
                s_ = stats[tr]
                s_[0] += len(seg)
                s_[1] += nprim
                s_[2] += float(c.per_prim * nprim + c.per_quartet * len(seg))
            for tier in (FP64, FP32):
                if work[tier]:
                    # costliest work items first: the dispatcher starts them early
                    order = np.argsort(-np.array(cost_w[tier]), kind="stable")
                    L = Launch(cls, bra, ket, diag, tier, np.array(work[tier], dtype=np.int32)[order])
                    L.ket_list = np.concatenate(lists[tier]).astype(np.int32)
                    L.n_quartets, L.n_prim, L.flops = stats[tier]
                    out.append(L)
        # biggest launches first: their tails overlap the small ones'
        out.sort(key=lambda L: -L.flops)
        return out

    # -- device -------------------------------------------------------------
    def to_device(self) -> None:
        import torch
        dev = native.device()
        b = self.basis
        for t in self.tables.values():
            recs = np.concatenate([p.rec for p in t.pairs], axis=0)
            off = np.concatenate([[0], np.cumsum(t.deg)]).astype(np.int32)
            sep = np.array([b.shell_centre[p.a] - b.shell_centre[p.b] for p in t.pairs])
            ao = np.array([(b.shell_ao[p.a], b.shell_ao[p.b]) for p in t.pairs], dtype=np.int32)
            t.dev = {"prim": native.to_dev(recs), "off": native.to_dev(off),
                     "sep": native.to_dev(sep), "G": native.to_dev(t.G),
                     "sh": native.to_dev(t.sh.astype(np.int32)), "ao": native.to_dev(ao)}
            d = t.dev
            t.side = native.FjSide(d["prim"].data_ptr(), d["off"].data_ptr(), 0,
                                   d["sep"].data_ptr(), d["G"].data_ptr(), d["sh"].data_ptr(),
                                   d["ao"].data_ptr(), t.n)
        for L in self.launches:
            L.work_dev = native.to_dev(L.work)
            L.ket_list_dev = native.to_dev(L.ket_list)
        self.shell_ao = native.to_dev(b.shell_ao.astype(np.int32))
        self.shell_size = native.to_dev(b.shell_size.astype(np.int32))
        self.Dsh = torch.zeros((b.n_sh, b.n_sh), dtype=torch.float64, device=dev)
        self.F = torch.zeros((b.n_ao, b.n_ao), dtype=torch.float64, device=dev)
        self._sys_for = None
        self._dev = True

    def _sys(self, D_dev):
        """The one-entry system table for density D_dev (rebuilt when the
        density buffer changes)."""
        key = D_dev.data_ptr()
        if self._sys_for != key:
            self._D_ref = D_dev
            self.sys = native.sys_table([(D_dev, self.Dsh, self.F)])
            self._sys_for = key
        return self.sys

    def launch(self, L: "Launch", sys, stream=None) -> None:
        native.call_fx(L.cls, self.threads, L.precision, L.bra.side, L.ket.side, L.work_dev,
                       L.work.shape[0], sys, L.diag, self.tau, stream, L.ket_list_dev)

    def build(self, D_dev, stream=None, close: bool = True):
        """G for the device density D_dev (n_ao x n_ao float64), on `stream`.
        Returns the builder's F tensor (overwritten by the next build)."""
        if not self._dev:
            self.to_device()
        sys = self._sys(D_dev)
        self.F.zero_()
        native.density_blocks(self.shell_ao, self.shell_size, D_dev, self.Dsh, stream)
        for L in self.launches:
            self.launch(L, sys, stream)
        if close:
            native.close(self.F, stream)
        return self.F

    # -- graph capture --------------------------------------------------------
    def capture(self, D_dev, stream=None):
        """Capture one whole build (zero F, density block maxima, every
        launch, close) into a HIP graph on a side stream; `replay()` then
        reruns it with one launch. `D_dev` must keep its address: to build
        for a new density, copy it into `D_dev` first. F is cleared by
        torch's fill kernel (no memset node in the graph)."""
        import torch
        if not self._dev:
            self.to_device()
        self.build(D_dev)           # warm: tables resident, code objects loaded
        torch.cuda.synchronize()
        self._graph_D = D_dev
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.build(D_dev)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()
        return self.F

    # -- counts -------------------------------------------------------------
    @property
    def n_quartets(self) -> int:
        return sum(L.n_quartets for L in self.launches)

    @property
    def n_prim_quartets(self) -> int:
        return sum(L.n_prim for L in self.launches)

    @property
    def flops(self) -> float:
        return sum(L.flops for L in self.launches)

    def fp32_share(self) -> float:
        q = self.n_quartets
        return sum(L.n_quartets for L in self.launches if L.precision == FP32) / q if q else 0.0


def energy(D: np.ndarray, G: np.ndarray) -> float:
    """E_2e = 1/2 sum_mn D_mn G_mn."""
    return 0.5 * float(np.sum(D * G))
