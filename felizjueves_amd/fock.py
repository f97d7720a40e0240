"""The native Fock builder: ``G = J - K/2`` on one MI355X.

Host planning, once per basis (and density screen):

* every canonical shell pair, its primitive-pair bounds, pruning and staged
  records (`pairs.kept_pairs`, the upstream rules);
* per pair class (la, lb) ONE ragged pair table on the device -- pairs sorted
  by (degree, G) descending, each pair's records contiguous, an offset
  array, no degree-uniform batching (so a class pair is one launch, not one
  per (degree, magnitude) bin pair);
* per class pair (X <= Y) a work list of tiles (bra pair, first ket): one
  workgroup each, ``threads`` consecutive kets. Which quartets the plan
  covers depends on the pair bounds alone: those with G_ab G_cd < tau are
  dropped (the upstream pre-launch screen, `eri/fock/launch_f.py:448-466`,
  per quartet). The density the builder is planned with only orders the
  rest: kets whose six-block screen G_ab G_cd dmax < tau fails under it are
  tiled apart (cold items, listed last), and in mixed mode the bound rule
  (or the learned selector) groups the likely FP32 kets together.

Per build (device only): zero F, the density block maxima, (mixed mode)
the tier of every work item from the CURRENT density (`fj_tier_select`:
the bound rule and the selector's network), one launch per class pair,
close F <- F + F^T. The kernels apply the six-block density screen per
quartet with the current density (`eri/fock/kernel_f.py:150-182`), so a
plan stays exact for any density: the screen, weights and degeneracy are
the upstream ones (`eri/fock/kernel_f.py:1-165`), and the result equals the
upstream pair-stream walk up to its binned pre-launch screen (eq. 29),
which drops a subset of the quartets with G_ab G_cd < tau.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import native
from . import pairs as P
from .basis import Basis
from .features import N_FEATURES, items_features, side_arrays, tile_features
from .hgp import cost

FP64, FP32 = 0, 1
MIXED = 2   # one launch for both tiers: items [0, n_fp64) FP64, the rest FP32
# Planning only: kets that pass the density-free screen G_ab G_cd >= tau but
# not the planning density's six-block screen. They are tiled apart from the
# rest (whole kets per lane, FP64 work items listed last, counted as no
# work): under the planning density every lane of such an item fails the
# kernel's density screen at once, and under a density that has grown
# there, they compute what the kernel's screen lets through.
COLD = 3

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
FOCK_LMAX = 3  # through (ff|ff): HGP kernels through (dd|dd), the f classes in loop form

# The primitive-loop split (see FockBuilder._lane_flops): a lane's share of
# a quartet's primitive loop is at most LANE_FLOPS_MAX algorithmic FLOPs, and
# less when the whole build would not give LANES_TARGET lanes (256 CUs x 4
# SIMDs x 4 wavefronts x 64 lanes) that much work each.
LANE_FLOPS_MAX = 1.0e6
LANES_TARGET = 256 * 4 * 4 * 64
# Splitting a quartet's primitive loop costs FP64 accuracy: every part goes
# through its own HRR and its own F updates, and F then sums many more,
# larger, partly cancelling terms (on a 10-atom cluster, single-primitive
# parts move G by ~1e-10 where whole quartets agree with the oracle to
# ~1e-11). So builds of at least SPLIT_MAX_FLOPS, which fill the chip
# anyway, are split only down to LANE_FLOPS_MAX per lane.
SPLIT_MAX_FLOPS = 4.0e9
# Launch workspace per stream (the f classes keep their contracted boundary
# there): at most this many bytes.
WS_MAX_BYTES = 8 << 30
KET_INDEX_LIMIT = 1 << 17  # ket-list entries: pair index in bits 0-16
# Below this many shell pairs a mixed builder with a learned selector plans
# in Python, so the selector also groups its FP32 tiles at plan time.
SELECTOR_PLAN_PAIRS = 50_000


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
    work: np.ndarray          # [n, 4] int32 items (bra, first ket entry, system | bra chunk, end entry)
    work_dev: object = None
    ket_list: np.ndarray = None
    ket_list_dev: object = None
    threads: int = 64         # lanes per work item
    n_quartets: int = 0       # computed (post-screen) contracted quartets
    n_prim: int = 0           # primitive quartets of those
    flops: float = 0.0        # algorithmic FLOPs of those
    n_fp64: int = 0           # MIXED: work items [0, n_fp64) run FP64, the rest FP32 (as planned)
    n_quartets32: int = 0     # MIXED: the FP32 items' quartets (as planned)
    n_cold: int = 0           # the last n_cold FP64 items hold the cold kets (see COLD)
    tier_dev: object = None   # MIXED: per-item tiers chosen on the device per build


WAVE = 64             # lanes of a wavefront
FILL_MIN_PRIMS = 4    # a lane is halved to fill idle lanes only from this many primitive pairs


def fill_wave(k: np.ndarray, lo: np.ndarray, cnt: np.ndarray):
    """Lanes (ket, first primitive pair, count), sorted by count descending,
    of one bra pair and tier: the last wavefront's idle lanes are filled by
    halving the lane with the most primitive pairs (the first such, ceil
    half kept in place, the rest appended) while that count is at least
    FILL_MIN_PRIMS, then the lanes are stably re-sorted by count. A
    wavefront runs as long as its longest lane; the halves run in lanes
    that would idle, so the same wavefronts finish sooner."""
    n = len(cnt)
    target = -(-n // WAVE) * WAVE
    if n == target or n == 0:
        return k, lo, cnt
    k, lo, cnt = list(k), list(lo), list(cnt)
    while len(cnt) < target:
        i = max(range(len(cnt)), key=lambda x: (cnt[x], -x))
        c = cnt[i]
        if c < FILL_MIN_PRIMS:
            break
        h = (c + 1) // 2
        cnt[i] = h
        k.append(k[i])
        lo.append(lo[i] + h)
        cnt.append(c - h)
    k, lo, cnt = np.array(k), np.array(lo, dtype=np.int64), np.array(cnt, dtype=np.int64)
    order = np.argsort(-cnt, kind="stable")
    return k[order], lo[order], cnt[order]


# A pair's primitive pairs are kept together (built on one centre) while the
# estimated FP64 amplification of the ones on the far side stays below this
# (relative error ~ 1e-16 x this); otherwise the pair is split by centre.
ORIENT_AMP_MAX = 10.0
# pairs whose shells are all below this angular momentum stay as they are.
# Orienting s/p pairs too changes little in FP64 (10-atom cluster: the same
# G to 1e-14) but matters for the FP32 tier, where their one-unit HRR still
# amplifies FP32 rounding: the 40-atom mixed build was 1.9e-8 from FP64
# with them left as they are, 2.1e-11 oriented; leaving them saved launches
# (45 -> 36 class pairs) but no time in the graph build.
ORIENT_L_MIN = 0


def orient_pairs(basis: Basis, pairs: list) -> list:
    """Build every pair on the centre its primitive pairs sit near. The
    HGP boundary is built on a pair's first centre and the HRR moves lb
    units of angular momentum across |AB|: a primitive pair whose product
    centre P lies near B has moments about A larger than its result by
    about (1 + |PA| sqrt(2 zeta))^lb, and those cancel in the HRR, so its
    rounding error is amplified by that much (against an MMD reference,
    a 10-atom cluster's G was 1.3e-10 off built on A only, 1.7e-12
    oriented). Per pair, with size_j = |U_j| (2 zeta_j)^(-(la+lb)/2) the
    scale of primitive pair j's contribution and amp_j its amplification
    in the other orientation: keep the pair (a, b) if the size-weighted
    amplification of its near-B primitive pairs is below ORIENT_AMP_MAX,
    else take it whole as (b, a) (class (lb, la), records (P, P - B, zeta,
    U)) if that of its near-A ones is, else split it into the two parts.
    The parts sum to the pair (the integral is linear in its primitive
    pairs), and a bra part (b, a) reaches F[b][a] instead of F[a][b], which
    the closing F + F^T makes the same."""
    import copy
    out = []
    for p in pairs:
        if p.a == p.b or max(p.la, p.lb) < ORIENT_L_MIN:
            out.append(p)
            continue
        A = basis.shell_centre[p.a]
        B = basis.shell_centre[p.b]
        if not np.any(A != B):
            out.append(p)
            continue
        P_ = p.rec[:, 0:3]
        z = p.rec[:, 6]
        dA = np.sqrt(((P_ - A) ** 2).sum(1))
        dB = np.sqrt(((P_ - B) ** 2).sum(1))
        near_b = dB < dA
        size = np.abs(p.rec[:, 7]) * (2.0 * z) ** (-0.5 * (p.la + p.lb))
        tot = size.sum()
        amp_on_a = (1.0 + dA * np.sqrt(2.0 * z)) ** p.lb     # built on A
        amp_on_b = (1.0 + dB * np.sqrt(2.0 * z)) ** p.la     # built on B
        err_a = (size * amp_on_a)[near_b].sum() / tot
        err_b = (size * amp_on_b)[~near_b].sum() / tot

        def on_b(q, sel):
            r = copy.copy(q)
            r.a, r.b, r.la, r.lb = q.b, q.a, q.lb, q.la
            r.prims = q.prims[sel]
            rec = q.rec[sel].copy()
            rec[:, 3:6] = rec[:, 0:3] - B
            r.rec = np.ascontiguousarray(rec)
            return r
        if err_a <= ORIENT_AMP_MAX:
            out.append(p)
        elif err_b <= ORIENT_AMP_MAX:
            out.append(on_b(p, np.ones(len(z), dtype=bool)))
        else:
            q = copy.copy(p)
            q.prims = p.prims[~near_b]
            q.rec = np.ascontiguousarray(p.rec[~near_b])
            out.append(q)
            out.append(on_b(p, near_b))
    return out


def merge_tier_launches(launches: list, every=False) -> list:
    """Each class pair's FP64 and FP32 launches as one MIXED launch: the
    FP64 work items (in their order) then the FP32 ones, the FP32 items'
    ket-list ranges moved past the FP64 list; biggest launches first.
    every: a class pair with items of one tier only is a MIXED launch too
    (so that every launch of a mixed build takes per-build tiers); a
    predicate of the class pair limits that to the class pairs it accepts."""
    by = {}
    for L in launches:
        by.setdefault(L.cls, {})[L.precision] = L
    out = []
    for cls, d in by.items():
        if (every(cls) if callable(every) else every) and len(d) == 1:
            (tier, a), = d.items()
            M = Launch(cls, a.bra, a.ket, a.diag, MIXED, a.work)
            M.threads, M.ket_list = a.threads, a.ket_list
            M.n_fp64 = a.work.shape[0] if tier == FP64 else 0
            M.n_cold = a.n_cold
            M.n_quartets, M.n_prim, M.flops = a.n_quartets, a.n_prim, a.flops
            M.n_quartets32 = a.n_quartets if tier == FP32 else 0
            out.append(M)
        elif FP64 in d and FP32 in d:
            a, b = d[FP64], d[FP32]
            wb = b.work.copy()
            off = len(a.ket_list)
            wb[:, 1] += off
            wb[:, 3] += off
            M = Launch(cls, a.bra, a.ket, a.diag, MIXED, np.concatenate([a.work, wb]).astype(np.int32))
            M.threads = a.threads
            M.ket_list = np.concatenate([a.ket_list, b.ket_list]).astype(np.int32)
            M.n_fp64 = a.work.shape[0]
            M.n_cold = a.n_cold
            M.n_quartets = a.n_quartets + b.n_quartets
            M.n_quartets32 = b.n_quartets
            M.n_prim = a.n_prim + b.n_prim
            M.flops = a.flops + b.flops
            out.append(M)
        else:
            out.extend(d.values())
    out.sort(key=lambda L: -L.flops)
    return out


class FockBuilder:
    """G = J - K/2 for one system, or (FockBuilder.batch) for many
    independent systems in the same launches: every class pair of every
    system is one launch per tier, each work item naming its system."""

    def __init__(self, basis: Basis, D: np.ndarray, *, tau: float = 1e-12, threads=64,
                 mixed: bool = False, fp32_error: float = 1e-11, selector=None,
                 lane_flops=None, fp32_bound: float | None = None, planner: str = "auto",
                 merge_tiers: bool = True, orient: bool = True, fp32_classes=None):
        self._setup([(basis, D)], tau=tau, threads=threads, mixed=mixed, fp32_error=fp32_error,
                     selector=selector, lane_flops=lane_flops, fp32_bound=fp32_bound,
                     planner=planner, merge_tiers=merge_tiers, orient=orient,
                     fp32_classes=fp32_classes)

    @classmethod
    def batch(cls, systems, **kw) -> "FockBuilder":
        """One builder for the systems [(basis, D), ...]: tables hold every
        system's pairs (grouped by system), kets pair only with bras of
        their own system, F/D/Dsh per system."""
        self = cls.__new__(cls)
        self._setup(list(systems), **kw)
        return self

    def _setup(self, systems, *, tau: float = 1e-12, threads=64, mixed: bool = False,
               fp32_error: float = 1e-11, selector=None, lane_flops=None,
               fp32_bound: float | None = None, planner: str = "auto",
               merge_tiers: bool = True, orient: bool = True, fp32_classes=None):
        """threads, lane_flops: one value for every class pair, or a dict
        {(la, lb, lc, ld): value} (the autotuner's choice per class pair;
        missing ones take the default 64 and the automatic split).
        fp32_bound: instead of the amplified estimate, the plain rule "FP32
        iff G_ab G_cd dmax < fp32_bound" (the rule the working-size mixed
        configuration is specified with; not accurate enough for the FP64
        G bars, kept for comparison).
        merge_tiers: a class pair with tiles in both tiers runs as ONE
        launch (fj_fxm_*: each work item in its tier) instead of one per
        tier, so the FP32 tiles add no launches and no launch tails.
        orient: every shell pair's primitive pairs are built on the centre
        nearer their product centre (`orient_pairs`): a pair may run as
        (b, a), in class (lb, la), and a mixed pair as two partial pairs.
        n_quartets then counts the partial quartets; shell_quartets() the
        contracted shell quartets.
        fp32_classes: the class pairs (la, lb, lc, ld) a mixed build may
        send to the FP32 tier (the autotuner's choice: where the mixed
        launch is measured faster than the FP64 one); the others run FP64
        launches. None: every class pair."""
        self.merge_tiers = bool(merge_tiers)
        self.orient = bool(orient)
        self.fp32_classes = None if fp32_classes is None else {tuple(int(x) for x in c)
                                                               for c in fp32_classes}
        self.fp32_bound = fp32_bound
        if not systems:
            raise ValueError("no systems")
        if len(systems) > 0xFFFF:
            raise ValueError("a work item names at most 65535 systems")
        for basis, _ in systems:
            if basis.l_max > FOCK_LMAX:
                raise ValueError(f"shells up to l = {basis.l_max}: the F kernels are built through "
                                 f"l = {FOCK_LMAX}")
        self.systems = systems
        self.basis = systems[0][0]
        self.tau = float(tau)
        if isinstance(threads, dict):
            self.class_threads = {tuple(k): int(v) for k, v in threads.items()}
            self.threads = 64
        else:
            self.class_threads = {}
            self.threads = int(threads)
        self.mixed = bool(mixed)
        self.fp32_error = float(fp32_error)
        if isinstance(lane_flops, dict):
            self.class_lane_flops = {tuple(k): float(v) for k, v in lane_flops.items()}
            self.lane_flops = None
        else:
            self.class_lane_flops = {}
            self.lane_flops = None if lane_flops is None else float(lane_flops)
        t0 = time.time()
        by_cls: dict = {}
        self.deleted, self.n_prim_total = [], 0
        self.Dsh_plans = []
        for m, (basis, D) in enumerate(systems):
            kept, deleted, self.eps, self.g_max, n_prim, _ = P.kept_pairs(basis, D, self.tau, 1)
            self.deleted += deleted
            self.n_prim_total += n_prim
            big = [p for p in kept if p.degree > P.MAX_PRIM_PAIRS]
            if big:
                raise ValueError(f"{len(big)} pairs keep more than {P.MAX_PRIM_PAIRS} primitive "
                                 "pairs, the staging cap of the F kernels")
            if self.orient:
                kept = orient_pairs(basis, kept)
            for p in kept:
                by_cls.setdefault((p.la, p.lb), []).append((m, p))
            self.Dsh_plans.append(P.density_blocks(basis, D))
        self.Dsh_plan = self.Dsh_plans[0]
        self.tables = {}
        for c, mps in sorted(by_cls.items()):
            if len(mps) >= KET_INDEX_LIMIT:
                raise ValueError(f"{len(mps)} shell pairs of class {c}: a ket-list entry holds "
                                 f"pair indices below {KET_INDEX_LIMIT}")
            mps = sorted(mps, key=lambda mp: (mp[0], -mp[1].degree, -mp[1].G))
            ps = [p for _, p in mps]
            mol = np.array([m for m, _ in mps], dtype=np.int64)
            t = SideTable(c, ps, np.array([p.G for p in ps]), np.array([p.degree for p in ps]),
                          np.array([(p.a, p.b) for p in ps], dtype=np.int64))
            t.mol = mol
            t.mol_range = {}
            for m in np.unique(mol):
                idx = np.nonzero(mol == m)[0]
                t.mol_range[int(m)] = (int(idx[0]), int(idx[-1]) + 1)
            t.amp = np.array([fp32_amplification(systems[m][0], p) for m, p in mps])
            feats = [side_arrays(systems[m][0], [p for mm, p in mps if mm == m])
                     for m in sorted(t.mol_range)]
            t.feat = {k: np.concatenate([f[k] for f in feats]) for k in feats[0]}
            self.tables[c] = t
        # selector(features [n, N_FEATURES]) -> bool [n]: True moves an FP64
        # tile of the bound rule to the FP32 tier (mixed mode only)
        self.selector = selector
        # the native planner, unless the learned selector should also group
        # the tiles it sends to FP32 at plan time (the Python planner, the
        # same plan otherwise; "auto" does that below SELECTOR_PLAN_PAIRS
        # shell pairs). Either way the device picks every mixed work item's
        # tier per build, with the network when there is one.
        n_pairs = sum(t.n for t in self.tables.values())
        if planner == "auto":
            planner = ("python" if (mixed and selector is not None and n_pairs < SELECTOR_PLAN_PAIRS)
                       else "native")
        if planner not in ("python", "native"):
            raise ValueError(f"planner = {planner!r}: 'native', 'python' or 'auto'")
        self.planner = planner
        self.launches = self._plan() if self.planner == "python" else self._plan_native()
        if self.merge_tiers:
            self.launches = merge_tier_launches(self.launches, every=self.mixed_for)
        self.plan_seconds = time.time() - t0
        self._dev = False

    # -- planning ---------------------------------------------------------
    def mixed_for(self, cls) -> bool:
        """Whether class pair cls may put work in the FP32 tier."""
        return self.mixed and (self.fp32_classes is None or tuple(cls) in self.fp32_classes)

    def _bounds(self, bra: SideTable, ket: SideTable, b: int) -> np.ndarray:
        return self._bounds_dmax(bra, ket, b)[0]

    def _bounds_range(self, bra: SideTable, ket: SideTable, b: int):
        """(lo, G_ab G_cd dmax, dmax) of bra pair b against the ket pairs
        lo .. lo + n of its own system."""
        m = int(bra.mol[b])
        lo, hi = ket.mol_range.get(m, (0, 0))
        Dsh = self.Dsh_plans[m]
        sa, sb = bra.sh[b]
        sc, sd = ket.sh[lo:hi, 0], ket.sh[lo:hi, 1]
        dmax = np.maximum.reduce([
            np.full(hi - lo, 2.0 * Dsh[sa, sb]), 2.0 * Dsh[sc, sd], 0.5 * Dsh[sa, sc],
            0.5 * Dsh[sa, sd], 0.5 * Dsh[sb, sc], 0.5 * Dsh[sb, sd]])
        return lo, bra.G[b] * (ket.G[lo:hi] * dmax), dmax

    def _bounds_dmax(self, bra: SideTable, ket: SideTable, b: int):
        """(G_ab G_cd dmax, dmax) of bra pair b against every ket pair (zero
        for the kets of other systems)."""
        lo, bound, dmax = self._bounds_range(bra, ket, b)
        fb, fd = np.zeros(ket.n), np.zeros(ket.n)
        fb[lo:lo + len(bound)] = bound
        fd[lo:lo + len(dmax)] = dmax
        return fb, fd

    def _tiles(self) -> list:
        """Per class pair: the live (bra, ket) quartets of each bra, split by
        the bound rule into the FP32 set (mixed mode) and the rest, each cut
        into tiles of `threads` kets in table order (degree, then G,
        descending). Returns [(X, Y, [(b, kets, tier)], features of the FP64
        tiles or None)]."""
        T = self.threads
        classes = sorted(self.tables)
        out = []
        for i, X in enumerate(classes):
            for Y in classes[i:]:
                bra, ket = self.tables[X], self.tables[Y]
                diag = X == Y
                want_feats = self.mixed_for((*X, *Y)) and self.selector is not None
                tiles, feats = [], []
                for b in range(bra.n):
                    lo, bound, dmax = self._bounds_range(bra, ket, b)
                    # the density-free screen decides which quartets the
                    # plan covers; the planning density only sorts them
                    live = bra.G[b] * ket.G[lo:lo + len(bound)] >= self.tau
                    if diag:
                        live[:max(0, b - lo)] = False
                    kl = np.nonzero(live)[0]
                    if not len(kl):
                        continue
                    hot = bound[kl] >= self.tau
                    cold = kl[~hot] + lo
                    kl = kl[hot]
                    ks = kl + lo
                    bound, dmax = bound[kl], dmax[kl]
                    if self.mixed_for((*X, *Y)):
                        if self.fp32_bound is not None:
                            small = bound < self.fp32_bound
                        else:
                            est = bound * bra.amp[b] * ket.amp[ks] * FP32_EST_SCALE
                            small = est < self.fp32_error
                        parts = ((FP64, ks[~small], dmax[~small]), (FP32, ks[small], dmax[small]),
                                 (COLD, cold, None))
                    else:
                        parts = ((FP64, ks, dmax), (COLD, cold, None))
                    for tier, kt, dm in parts:
                        if not len(kt):
                            continue
                        starts = np.arange(0, len(kt), T)
                        for t0 in starts:
                            tiles.append((b, kt[t0:t0 + T], tier))
                        if want_feats and tier == FP64:
                            feats.append(tile_features((*X, *Y), bra.feat, ket.feat, b, kt, starts,
                                                       dm))
                f = np.concatenate(feats) if feats else None
                out.append((X, Y, tiles, f))
        return out

    def _plan_native(self) -> list:
        """The same plan as _plan, from the native planner (planner.py):
        a counting pass for the automatic split, then one call per class
        pair."""
        from . import planner as NP
        sides = {c: NP.Side(t) for c, t in self.tables.items()}
        systems = NP.Systems(self.Dsh_plans)
        classes = sorted(self.tables)
        pairs = [(X, Y) for i, X in enumerate(classes) for Y in classes[i:]]
        common = dict(tau=self.tau, fp32_error=self.fp32_error,
                      fp32_bound=self.fp32_bound, est_scale=FP32_EST_SCALE,
                      max_prim=P.MAX_PRIM_PAIRS)
        if self.lane_flops is None:
            total = 0.0
            for X, Y in pairs:
                c = cost((*X, *Y))
                r, _ = NP.plan_pairs(sides[X], sides[Y], self.tables[Y], systems, diag=int(X == Y),
                                     threads=64, lane_flops=0.0, per_prim=c.per_prim,
                                     per_quartet=c.per_quartet, mixed=self.mixed_for((*X, *Y)),
                                     **common)
                total += r[0][2][2] + r[1][2][2]
            auto_lf = LANE_FLOPS_MAX if total >= SPLIT_MAX_FLOPS else \
                float(min(LANE_FLOPS_MAX, max(1.0, total / LANES_TARGET)))
        else:
            auto_lf = self.lane_flops
        self.lane_flops_used = auto_lf
        out = []
        for X, Y in pairs:
            cls = (*X, *Y)
            c = cost(cls)
            T = self.class_threads.get(cls, self.threads)
            lf = self.class_lane_flops.get(cls, auto_lf)
            r, n_cold = NP.plan_pairs(sides[X], sides[Y], self.tables[Y], systems,
                                      diag=int(X == Y), threads=T, lane_flops=lf,
                                      per_prim=c.per_prim, per_quartet=c.per_quartet,
                                      mixed=self.mixed_for(cls), **common)
            for tier, (w, k, stats) in zip((FP64, FP32), r):
                if w is None:
                    continue
                L = Launch(cls, self.tables[X], self.tables[Y], int(X == Y), tier, w)
                L.threads = T
                L.ket_list = k
                L.n_quartets, L.n_prim, L.flops = stats
                L.n_cold = n_cold if tier == FP64 else 0
                out.append(L)
        out.sort(key=lambda L: -L.flops)
        return out

    def _lane_flops(self, groups) -> float:
        """The primitive-loop split: `lane_flops` if given; else, for a build
        below SPLIT_MAX_FLOPS, its FLOPs spread over LANES_TARGET lanes (a
        small system is split down to single primitive quartets per lane so
        that its few quartets still fill the chip), and LANE_FLOPS_MAX for
        larger builds."""
        if self.lane_flops is not None:
            return self.lane_flops
        total = 0.0
        for X, Y, tiles, _ in groups:
            c = cost((*X, *Y))
            bra, ket = self.tables[X], self.tables[Y]
            for b, seg, t in tiles:
                if t != COLD:
                    total += c.per_prim * bra.deg[b] * ket.deg[seg].sum() + c.per_quartet * len(seg)
        if total >= SPLIT_MAX_FLOPS:
            return LANE_FLOPS_MAX
        return float(min(LANE_FLOPS_MAX, max(1.0, total / LANES_TARGET)))

    def _plan(self) -> list:
        """Tiles -> tiers (bound rule, then the learned selector on the
        FP64 tiles) -> one work item per tile (or per primitive-loop chunk)
        per (class pair, tier) launch."""
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
        auto_lf = self._lane_flops(groups)
        self.lane_flops_used = auto_lf
        out = []
        for X, Y, tiles, _ in groups:
            bra, ket = self.tables[X], self.tables[Y]
            cls = (*X, *Y)
            c = cost(cls)
            lane_flops = self.class_lane_flops.get(cls, auto_lf)
            T = self.class_threads.get(cls, self.threads)
            diag = int(X == Y)
            work = {FP64: [], FP32: []}
            cost_w = {FP64: [], FP32: []}
            lists = {FP64: [], FP32: []}
            n_list = {FP64: 0, FP32: 0}
            stats = {FP64: [0, 0, 0.0], FP32: [0, 0, 0.0]}
            # every tile of one bra pair and tier is expanded into lanes
            # together: (ket, primitive chunk), chunks of one ket equal in
            # size, lanes sorted by primitive count so that each work item's
            # lanes loop the same number of times
            by_bra: dict = {}
            n_cold = 0
            for b, seg, tr in tiles:
                by_bra.setdefault((b, tr), []).append(seg)
            for (b, tr), segs in by_bra.items():
                seg = np.concatenate(segs)
                kab = int(bra.deg[b])
                kcd = ket.deg[seg].astype(np.int64)
                if tr == COLD:
                    # whole kets, whole bra, no cost: under the planning
                    # density every lane exits at the kernel's screen
                    st = n_list[FP64]
                    lists[FP64].append(seg.astype(np.int64))
                    n_list[FP64] += len(seg)
                    m = int(bra.mol[b])
                    for t0 in range(0, len(seg), T):
                        work[FP64].append((b, st + t0, m, st + min(len(seg), t0 + T)))
                        cost_w[FP64].append(-1)   # below every counted item: listed last
                        n_cold += 1
                    continue
                # split the primitive loop so that no lane's share exceeds
                # lane_flops: bra primitive pairs per work item (cb), then
                # ket primitive pairs per lane (ck)
                cb = kab if kab * c.per_prim <= lane_flops else max(1, int(lane_flops // c.per_prim))
                ck = max(1, min(P.MAX_PRIM_PAIRS, int(lane_flops // (cb * c.per_prim))))
                n_ch = -(-kcd // ck)                       # chunks per ket
                size = -(-kcd // n_ch)                     # equal chunk size
                rep = np.repeat(np.arange(len(seg)), n_ch)
                j = np.arange(len(rep)) - np.repeat(np.cumsum(n_ch) - n_ch, n_ch)
                lo = j * size[rep]
                cnt = np.minimum(size[rep], kcd[rep] - lo)
                order = np.argsort(-cnt, kind="stable")
                rep, lo, cnt = fill_wave(rep[order], lo[order], cnt[order])
                whole = (lo == 0) & (cnt == kcd[rep])
                enc = np.where(whole, 0, np.where(lo + cnt >= kcd[rep], 0, cnt))
                lanes = seg[rep].astype(np.int64) | (np.where(whole, 0, lo) << 17) | (enc << 24)
                lane_prims = cnt
                st = n_list[tr]
                lists[tr].append(lanes)
                n_list[tr] += len(lanes)
                m = int(bra.mol[b])
                for t0 in range(0, len(lanes), T):
                    t1 = min(len(lanes), t0 + T)
                    pmax = int(lane_prims[t0])
                    for blo in range(0, kab, cb):
                        bn = 0 if cb >= kab else min(cb, kab - blo)
                        work[tr].append((b, st + t0, m | (blo << 16) | (bn << 24), st + t1))
                        cost_w[tr].append(min(cb, kab - blo) * pmax)
                nprim = int(kab * kcd.sum())
                s_ = stats[tr]
                s_[0] += len(seg)
                s_[1] += nprim
                s_[2] += float(c.per_prim * nprim + c.per_quartet * len(seg))
            for tier in (FP64, FP32):
                if work[tier]:
                    # costliest work items first: the dispatcher starts them early
                    order = np.argsort(-np.array(cost_w[tier]), kind="stable")
                    L = Launch(cls, bra, ket, diag, tier, np.array(work[tier], dtype=np.int32)[order])
                    L.threads = T
                    L.ket_list = np.concatenate(lists[tier]).astype(np.int32)
                    L.n_quartets, L.n_prim, L.flops = stats[tier]
                    L.n_cold = n_cold if tier == FP64 else 0
                    out.append(L)
        # biggest launches first: their tails overlap the small ones'
        out.sort(key=lambda L: -L.flops)
        return out

    # -- device -------------------------------------------------------------
    def to_device(self) -> None:
        import torch
        dev = native.device()
        for t in self.tables.values():
            recs = np.concatenate([p.rec for p in t.pairs], axis=0)
            off = np.concatenate([[0], np.cumsum(t.deg)]).astype(np.int32)
            sep, ao = [], []
            for m, p in zip(t.mol, t.pairs):
                bs = self.systems[m][0]
                sep.append(bs.shell_centre[p.a] - bs.shell_centre[p.b])
                ao.append((bs.shell_ao[p.a], bs.shell_ao[p.b]))
            t.dev = {"prim": native.to_dev(recs), "off": native.to_dev(off),
                     "sep": native.to_dev(np.array(sep)), "G": native.to_dev(t.G),
                     "sh": native.to_dev(t.sh.astype(np.int32)),
                     "ao": native.to_dev(np.array(ao, dtype=np.int32))}
            d = t.dev
            t.side = native.FjSide(d["prim"].data_ptr(), d["off"].data_ptr(), 0,
                                   d["sep"].data_ptr(), d["G"].data_ptr(), d["sh"].data_ptr(),
                                   d["ao"].data_ptr(), t.n)
        for t in self.tables.values():
            t.dev["amp"] = native.to_dev(np.asarray(t.amp, dtype=np.float64))
        for L in self.launches:
            L.work_dev = native.to_dev(L.work)
            L.ket_list_dev = native.to_dev(L.ket_list)
            if self.mixed and L.precision == MIXED:
                self._tier_setup(L)
        # per system: shell offsets/sizes, Dsh and F (views of flat buffers)
        self.shell_ao, self.shell_size = [], []
        n_ao = [b.n_ao for b, _ in self.systems]
        n_sh = [b.n_sh for b, _ in self.systems]
        self._F_flat = torch.zeros(sum(n * n for n in n_ao), dtype=torch.float64, device=dev)
        self._Dsh_flat = torch.zeros(sum(n * n for n in n_sh), dtype=torch.float64, device=dev)
        self.Fs, self.Dshs = [], []
        fo = so = 0
        for (b, _), na, ns in zip(self.systems, n_ao, n_sh):
            self.shell_ao.append(native.to_dev(b.shell_ao.astype(np.int32)))
            self.shell_size.append(native.to_dev(b.shell_size.astype(np.int32)))
            self.Fs.append(self._F_flat[fo:fo + na * na].view(na, na))
            self.Dshs.append(self._Dsh_flat[so:so + ns * ns].view(ns, ns))
            fo += na * na
            so += ns * ns
        self.F, self.Dsh = self.Fs[0], self.Dshs[0]
        self._sys_for = None
        self._dev = True

    def _tier_setup(self, L: "Launch") -> None:
        """The device side of a mixed launch's per-build tier choice: the
        tier array (initially the plan's), and the fj_tier_select record:
        the bound rule's parameters and, with a learned selector, every work
        item's density-free features, standardised, and whether they lie in
        the network's training range."""
        n = L.work.shape[0]
        plan = np.zeros(n, dtype=np.int32)
        plan[L.n_fp64:] = 1
        L.tier_dev = native.to_dev(plan)
        sel = native.FjTierSel()
        sel.work, sel.ket_list = L.work_dev.data_ptr(), L.ket_list_dev.data_ptr()
        sel.n_items, sel.diag, sel.tau = n, int(L.diag), self.tau
        sel.amp_bra, sel.amp_ket = L.bra.dev["amp"].data_ptr(), L.ket.dev["amp"].data_ptr()
        sel.est_scale, sel.fp32_error = FP32_EST_SCALE, self.fp32_error
        sel.fp32_bound = -1.0 if self.fp32_bound is None else float(self.fp32_bound)
        keep = [L.tier_dev]
        ls = self.selector if hasattr(self.selector, "net") else None
        if ls is not None:
            net = ls.net
            X = items_features(L.cls, L.bra.feat, L.ket.feat, L.work, L.ket_list)
            inside = np.all((X[:, 4:13] >= ls.lo[4:13]) & (X[:, 4:13] <= ls.hi[4:13]), axis=1)
            inside &= tuple(int(v) for v in L.cls) in ls.classes
            inside[L.n_fp64 - L.n_cold:L.n_fp64] = False      # cold items
            feat = native.to_dev(net.normalise(X))
            ins = native.to_dev(inside.astype(np.uint8))
            keep += [feat, ins]
            sel.feat, sel.inside, sel.n_feat = feat.data_ptr(), ins.data_ptr(), N_FEATURES
            sel.d_mean, sel.d_std = float(net.mean[N_FEATURES - 1]), float(net.std[N_FEATURES - 1])
            sel.d_lo, sel.d_hi = float(ls.lo[N_FEATURES - 1]), float(ls.hi[N_FEATURES - 1])
            sel.Wb, sel.P = net.Wb.data_ptr(), net.P.data_ptr()
            sel.margin, sel.log_budget = float(ls.margin), float(np.log10(ls.budget))
            keep.append(net)
        sel.tier = L.tier_dev.data_ptr()
        L.tier_sel, L.tier_keep = sel, keep

    def _tier_jobs(self):
        """(device job table, n_jobs, n_items, network) of the mixed
        launches' per-build tier choice (built once), or None."""
        if getattr(self, "_tjobs", None) is None:
            jobs, first = [], 0
            for L in self.launches:
                if L.tier_dev is not None:
                    jobs.append((L.bra.side, L.ket.side, L.tier_sel, first))
                    first += L.work.shape[0]
            if not jobs:
                self._tjobs = ()
            else:
                net = self.selector.net if hasattr(self.selector, "net") else None
                self._tjobs = (native.tier_jobs(jobs), len(jobs), first, net)
        return self._tjobs or None

    def _sys(self, D_devs):
        """The system table for the device densities D_devs (rebuilt when a
        density buffer changes)."""
        key = tuple(D.data_ptr() for D in D_devs)
        if self._sys_for != key:
            self._D_ref = list(D_devs)
            self.sys = native.sys_table(list(zip(D_devs, self.Dshs, self.Fs)))
            self._sys_for = key
        return self.sys

    def _workspace(self, i: int):
        """Launch workspace i (one per stream the build uses): the largest
        any launch asks for, capped at WS_MAX_BYTES (a workspace class then
        runs on fewer persistent workgroups)."""
        import torch
        ws = getattr(self, "_ws", None)
        if ws is None:
            ws = self._ws = []
        need = min(WS_MAX_BYTES, max([native.workspace_bytes(L.cls, L.threads, L.work.shape[0])
                                      for L in self.launches] + [0]))
        while len(ws) <= i:
            ws.append(torch.empty(need // 8, dtype=torch.float64, device=native.device())
                      if need else None)
        return ws[i]

    def launch(self, L: "Launch", sys, stream=None, ws=None) -> None:
        if L.precision == MIXED:
            native.call_fxm(L.cls, L.threads, L.bra.side, L.ket.side, L.work_dev, L.work.shape[0],
                            L.n_fp64, sys, L.diag, self.tau, stream, L.ket_list_dev, ws,
                            tier=L.tier_dev)
            return
        native.call_fx(L.cls, L.threads, L.precision, L.bra.side, L.ket.side, L.work_dev,
                       L.work.shape[0], sys, L.diag, self.tau, stream, L.ket_list_dev, ws)

    def _stream_assignment(self, n: int) -> list:
        """Stream of each launch: the launches, costliest first, each to the
        stream with the least work so far (by their FLOP counts)."""
        load = [0.0] * n
        out = []
        for L in self.launches:
            j = min(range(n), key=lambda i: load[i])
            load[j] += L.flops
            out.append(j)
        return out

    def _side_streams(self, n: int) -> list:
        import torch
        if len(getattr(self, "_streams", [])) != n:
            self._streams = [torch.cuda.Stream() for _ in range(n)]   # default priority
        return self._streams

    def build(self, D_dev, stream=None, close: bool = True, n_streams: int = 1):
        """G for the device density D_dev (n_ao x n_ao float64; for a batch
        builder, a list with one per system), on `stream`. With n_streams > 1
        the class-pair launches are dealt (costliest first, each to the least
        loaded) to that many side streams that fork from and join back into
        `stream`,
        so launches too small to fill the chip overlap. Returns the F tensor
        (the list of them for a batch; overwritten by the next build)."""
        import torch
        if self.mixed and not self.merge_tiers:
            raise ValueError("a mixed build chooses its tiers per build only with merged tiers "
                             "(merge_tiers=True); an unmerged mixed plan is for inspection")
        if not self._dev:
            self.to_device()
        D_devs = list(D_dev) if isinstance(D_dev, (list, tuple)) else [D_dev]
        if len(D_devs) != len(self.systems):
            raise ValueError(f"{len(D_devs)} densities for {len(self.systems)} systems")
        sys = self._sys(D_devs)
        main = stream if stream is not None else torch.cuda.current_stream()
        with torch.cuda.stream(main):
            self._F_flat.zero_()
        for m, D in enumerate(D_devs):
            native.density_blocks(self.shell_ao[m], self.shell_size[m], D, self.Dshs[m], main)
        if self.mixed:
            # every mixed launch's tiers, from this density, in one launch
            jobs = self._tier_jobs()
            if jobs is not None:
                native.tier_select_jobs(jobs[0], jobs[1], jobs[2], sys, jobs[3], main)
        self._built = True
        if n_streams <= 1:
            ws = self._workspace(0)
            for L in self.launches:
                self.launch(L, sys, main, ws)
        else:
            side = self._side_streams(n_streams)
            wss = [self._workspace(j) for j in range(n_streams)]
            for s in side:
                s.wait_stream(main)
            for L, j in zip(self.launches, self._stream_assignment(n_streams)):
                self.launch(L, sys, side[j], wss[j])
            for s in side:
                main.wait_stream(s)
        if close:
            for F in self.Fs:
                native.close(F, main)
        return self.Fs if isinstance(D_dev, (list, tuple)) else self.F

    # -- graph capture --------------------------------------------------------
    def capture(self, D_dev, stream=None, n_streams: int = 1):
        """Capture one whole build (zero F, density block maxima, the
        per-build tiers of a mixed build, every launch, close) into a HIP
        graph on a side stream; `replay()` then reruns it with one launch.
        The graph reads the density from D_dev's address: to build for a new
        density, copy it into `D_dev` and replay (the plan covers every
        quartet the pair bounds keep, so any density is built exactly). F
        is cleared by torch's fill kernel (no memset node in the graph)."""
        import torch
        if not self._dev:
            self.to_device()
        self.build(D_dev, n_streams=n_streams)   # warm: tables resident, code loaded
        torch.cuda.synchronize()
        self._graph_D = D_dev
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.build(D_dev, stream=s, n_streams=n_streams)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()
        return self.Fs if isinstance(self._graph_D, (list, tuple)) else self.F

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

    def shell_quartets(self) -> int:
        """Contracted shell quartets this build computes (the screened
        canonical quartets, however the pairs were oriented or split)."""
        if not self.orient:
            return int(self.n_quartets)
        if not hasattr(self, "_shell_quartets"):
            one = FockBuilder.__new__(FockBuilder)
            one._setup(self.systems, tau=self.tau, orient=False, merge_tiers=False)
            self._shell_quartets = int(one.n_quartets)
        return self._shell_quartets

    @staticmethod
    def item_quartets(L: "Launch") -> np.ndarray:
        """Per work item, the contracted quartets it starts (the lanes that
        hold a ket's first primitive chunk, in its first bra chunk; cold
        items count none), so that the counts of a launch sum to its
        n_quartets."""
        first = ((L.ket_list.astype(np.int64) >> 17) & 0x7f) == 0
        cs = np.concatenate([[0], np.cumsum(first)])
        q = (cs[L.work[:, 3]] - cs[L.work[:, 1]]).astype(np.int64)
        q[((L.work[:, 2].astype(np.int64) >> 16) & 0xff) != 0] = 0
        if L.n_cold:
            end = L.n_fp64 if L.precision == MIXED else L.work.shape[0]
            q[end - L.n_cold:end] = 0
        return q

    def fp32_share(self) -> float:
        """The share of the computed quartets in the FP32 tier: with the
        tiers the device chose in the last build (mixed launches), else as
        planned."""
        tot = f32 = 0
        for L in self.launches:
            q = self.item_quartets(L)
            if L.precision == MIXED:
                if L.tier_dev is not None and getattr(self, "_built", False):
                    t = L.tier_dev.cpu().numpy()
                else:
                    t = (np.arange(len(q)) >= L.n_fp64).astype(np.int32)
                f32 += int(q[t == 1].sum())
            elif L.precision == FP32:
                f32 += int(q.sum())
            tot += int(q.sum())
        return f32 / tot if tot else 0.0


def energy(D: np.ndarray, G: np.ndarray) -> float:
    """E_2e = 1/2 sum_mn D_mn G_mn."""
    return 0.5 * float(np.sum(D * G))
