"""HGP recurrences as DAGs, for the HIP generator and the FLOP model.

VRR (upstream `hgp/vrr.py:1-60,152-178`; eqs 25-26 of Head-Gordon & Pople as
the upstream repository reads them). A node ``(e, f, m)`` is the primitive
``[e0|f0]^(m)``. A non-leaf node is reduced along ONE axis ``d``: the first
(x, y, z) nonzero component of ``f`` when ``f != 0`` (eq 26), else of ``e``
(eq 25):

    [e0|f+1_d,0]^(m) = QC_d [e0|f0]^(m) + WQ_d [e0|f0]^(m+1)
                     + f_d/(2q) ([e0|f-1_d,0]^(m) - rho/q [e0|f-1_d,0]^(m+1))
                     + e_d/(2(p+q)) [e-1_d,0|f0]^(m+1)
    [e+1_d,0|00]^(m) = PA_d [e0|00]^(m) + WP_d [e0|00]^(m+1)
                     + e_d/(2p) ([e-1_d,0|00]^(m) - rho/p [e-1_d,0|00]^(m+1))

with ``WP = -q/(p+q) PQ`` and ``WQ = +p/(p+q) PQ`` (``PQ = P - Q``; the
ket sign is the derived one, upstream `hgp/recurrence.py:53-61`) and the
seed ``[00|00]^(m) = U_P U_Q (2/sqrt(pi)) sqrt(rho) F_m(T)``,
``T = rho |PQ|^2`` (`hgp/recurrence.py:62-94`). The targets are every
``(e, f)`` at m = 0 with ``la <= |e| <= la+lb``, ``lc <= |f| <= lc+ld``
(`hgp/vrr.py:116-124`); the DAG is their closure (dead nodes pruned).

HRR (upstream `hgp/hrr.py:1-60,101-120`), on CONTRACTED values:

    (a, b+1_k | c d) = (a+1_k, b | c d) + AB_k (a b | c d)
    (a b | c, d+1_k) = (a b | c+1_k, d) + CD_k (a b | c d)

reduced on the ket first (so the bra HRR runs first in build order), axis =
first nonzero component of ``d`` (else of ``b``). ``AB = A - B``,
``CD = C - D``.

The same reduction rules and the same axis choice as upstream, so the node
and statement counts are the upstream ones (2,330 VRR nodes, 961 boundary
values and 6,254 HRR statements at (dd|dd)); tests hold these against the
upstream FLOP table's class list.
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache

from .basis import cartesians

Cart = tuple[int, int, int]
ZERO: Cart = (0, 0, 0)


def _range(lo: int, hi: int) -> tuple[Cart, ...]:
    return tuple(c for n in range(lo, hi + 1) for c in cartesians(n))


def _sub(v: Cart, d: int, k: int = 1) -> Cart:
    return tuple(c - k * (i == d) for i, c in enumerate(v))  # type: ignore


def _add(v: Cart, d: int, k: int = 1) -> Cart:
    return tuple(c + k * (i == d) for i, c in enumerate(v))  # type: ignore


# ---------------------------------------------------------------------------
# VRR

@dataclass(frozen=True)
class VTerm:
    coef: str     # "PA" "WP" "i2p" "i2p_r" | "QC" "WQ" "i2q" "i2q_r" "i2pq"
    mult: int     # integer multiplier (the e_d / f_d factor, signed)
    child: tuple  # (e, f, m)


def vrr_reduce(node) -> tuple[tuple[int, tuple[VTerm, ...]]]:
    """(axis, terms) of a non-leaf node, () for the leaf [00|00]^(m)."""
    e, f, m = node
    if f != ZERO:
        d = next(i for i, c in enumerate(f) if c)
        g = _sub(f, d)
        t = [VTerm("QC", 1, (e, g, m)), VTerm("WQ", 1, (e, g, m + 1))]
        if g[d]:
            h = _sub(g, d)
            t += [VTerm("i2q", g[d], (e, h, m)), VTerm("i2q_r", -g[d], (e, h, m + 1))]
        if e[d]:
            t.append(VTerm("i2pq", e[d], (_sub(e, d), g, m + 1)))
        return d, tuple(t)
    if e != ZERO:
        d = next(i for i, c in enumerate(e) if c)
        g = _sub(e, d)
        t = [VTerm("PA", 1, (g, f, m)), VTerm("WP", 1, (g, f, m + 1))]
        if g[d]:
            h = _sub(g, d)
            t += [VTerm("i2p", g[d], (h, f, m)), VTerm("i2p_r", -g[d], (h, f, m + 1))]
        return d, tuple(t)
    return ()


def _vrr_key(node):
    e, f, m = node
    return (f != ZERO, sum(f), sum(e), -m, tuple(-c for c in e), tuple(-c for c in f))


@dataclass(frozen=True)
class VrrDag:
    cls: tuple[int, int, int, int]
    targets: tuple          # (e, f) at m = 0, e-major
    order: tuple            # nodes in a topological build order
    terms: dict             # node -> (axis, terms) or ()

    @property
    def n_nodes(self) -> int:
        return len(self.order)

    @property
    def m_max(self) -> int:
        return max(n[2] for n in self.order)


@lru_cache(maxsize=None)
def vrr_dag(cls) -> VrrDag:
    la, lb, lc, ld = cls
    targets = tuple((e, f) for e in _range(la, la + lb) for f in _range(lc, lc + ld))
    terms: dict = {}
    todo = [(e, f, 0) for e, f in targets]
    while todo:
        n = todo.pop()
        if n in terms:
            continue
        r = vrr_reduce(n)
        terms[n] = r
        if r:
            todo.extend(t.child for t in r[1])
    order = tuple(sorted(terms, key=_vrr_key))
    return VrrDag(tuple(cls), targets, order, terms)


# ---------------------------------------------------------------------------
# HRR

@lru_cache(maxsize=None)
def hrr_dag(cls):
    """(order, terms, boundary): order is a topological build order of every
    HRR node (boundary nodes (e,0|f,0) included); terms[node] is
    (side, axis, moved, lowered) or () for a boundary node."""
    la, lb, lc, ld = cls
    targets = tuple((a, b, c, d) for a in cartesians(la) for b in cartesians(lb)
                    for c in cartesians(lc) for d in cartesians(ld))
    terms: dict = {}
    todo = list(targets)
    while todo:
        n = todo.pop()
        if n in terms:
            continue
        a, b, c, d = n
        if d != ZERO:
            k = next(i for i, x in enumerate(d) if x)
            dm = _sub(d, k)
            r = ("ket", k, (a, b, _add(c, k), dm), (a, b, c, dm))
        elif b != ZERO:
            k = next(i for i, x in enumerate(b) if x)
            bm = _sub(b, k)
            r = ("bra", k, (_add(a, k), bm, c, d), (a, bm, c, d))
        else:
            r = ()
        terms[n] = r
        if r:
            todo += [r[2], r[3]]
    # build order: boundary first, then the bra HRR (d = 0) by |b|, then the
    # ket HRR by |d|; within a level by |a| descending (children first)
    def key(n):
        a, b, c, d = n
        return (sum(d), sum(b), -sum(c), -sum(a), tuple(-x for x in a),
                tuple(-x for x in b), tuple(-x for x in c), tuple(-x for x in d))
    order = tuple(sorted(terms, key=key))
    boundary = frozenset((n[0], n[2]) for n, r in terms.items() if not r)
    return order, terms, targets, boundary


# ---------------------------------------------------------------------------
# FLOP model (upstream convention, `hgp/cost.py:1-60`): one +,-,*,/ = 1 FLOP,
# no FMA credit; a VRR node costs one multiply per term, one more when
# |mult| > 1, and n - 1 adds; an HRR node costs one multiply and one add.

@dataclass(frozen=True)
class Cost:
    vrr_nodes: int
    boundary: int
    hrr_statements: int
    hrr_interior: int
    n_eri: int
    vrr_flops: int       # per primitive quartet (VRR steps + seed multiplies)
    hrr_flops: int       # per contracted quartet
    boys_order: int

    @property
    def per_prim(self) -> int:
        """FLOPs per primitive quartet: the VRR steps and seed multiplies, one
        add per boundary value accumulated, and the quartet geometry (16 for
        (ss|ss), 26 otherwise: the upstream table's counts)."""
        return self.vrr_flops + self.boundary + (16 if self.boys_order == 0 else 26)

    @property
    def per_quartet(self) -> int:
        """FLOPs per contracted quartet: the HRR and the 12-FLOP-per-ERI
        six-block predigest."""
        return self.hrr_flops + 12 * self.n_eri

    def flops(self, k_ab: int, k_cd: int) -> int:
        """Algorithmic FLOPs of one contracted quartet of degrees k_ab, k_cd."""
        return self.per_prim * k_ab * k_cd + self.per_quartet


@lru_cache(maxsize=None)
def cost(cls) -> Cost:
    v = vrr_dag(tuple(cls))
    vf = 0
    for n in v.order:
        r = v.terms[n]
        if not r:
            vf += 1  # the seed multiply S * F_m
            continue
        ts = r[1]
        vf += sum(1 + (abs(t.mult) > 1) for t in ts) + len(ts) - 1
    order, terms, targets, boundary = hrr_dag(tuple(cls))
    interior = sum(1 for n in order if terms[n])
    return Cost(v.n_nodes, len(boundary), len(order), interior, len(targets),
                vf, 2 * interior, sum(cls))


def builder_classes(lmax: int) -> tuple[tuple[int, int, int, int], ...]:
    """The classes the native builder launches: pair classes (la, lb) in
    both orders (a shell pair's primitive pairs are built on the centre
    nearer their product centre, so a (p, d) pair may run as (d, p)), bra
    pair class <= ket pair class: 45 for lmax = 2. The canonical ones
    (fock_classes) also carry the upstream-ABI entry."""
    pairs = [(a, b) for a in range(lmax + 1) for b in range(lmax + 1)]
    return tuple(p + q for i, p in enumerate(pairs) for q in pairs[i:])


def canonical(cls) -> bool:
    return cls[0] <= cls[1] and cls[2] <= cls[3]


def fock_classes(lmax: int) -> tuple[tuple[int, int, int, int], ...]:
    """Every AO-order class (la lb|lc ld), la <= lb, lc <= ld, bra pair <=
    ket pair (upstream `hgp/vrr.py:89-91` for lmax = 2: 21 classes)."""
    pairs = [(a, b) for a in range(lmax + 1) for b in range(a, lmax + 1)]
    return tuple(p + q for i, p in enumerate(pairs) for q in pairs[i:])


# ---------------------------------------------------------------------------
# Target groups for the staged kernels: the boundary split by bra component
# e, each group's VRR being the closure of its targets (nodes shared between
# groups are recomputed). Chosen so that each group's live set fits the
# register file.

def closure(dag: VrrDag, targets) -> frozenset:
    seen: set = set()
    todo = list(targets)
    while todo:
        n = todo.pop()
        if n in seen:
            continue
        seen.add(n)
        r = dag.terms[n]
        if r:
            todo.extend(t.child for t in r[1])
    return frozenset(seen)


def peak_live(dag: VrrDag, nodes) -> int:
    """Most values live at once executing `nodes` in build order, each
    target consumed (accumulated) right after it is formed, each other node
    dead after its last reader."""
    order = [n for n in dag.order if n in nodes]
    pos = {n: i for i, n in enumerate(order)}
    last: dict = {}
    for n in order:
        r = dag.terms[n]
        if r:
            for t in r[1]:
                last[t.child] = pos[n]
    alive: set = set()
    peak = 0
    for i, n in enumerate(order):
        alive.add(n)
        r = dag.terms[n]
        if r:
            for t in r[1]:
                if last.get(t.child) == i:
                    alive.discard(t.child)
        peak = max(peak, len(alive))
        if n not in last:
            alive.discard(n)
    return peak


@lru_cache(maxsize=None)
def e_groups(cls, live_max: int = 100):
    """Split the boundary's bra components into contiguous groups whose
    closures (plus their accumulators) keep at most `live_max` values live.
    Returns a tuple of tuples of e; one group means the fused kernel."""
    dag = vrr_dag(tuple(cls))
    es = []
    for e, _ in dag.targets:
        if e not in es:
            es.append(e)
    nF = len(dag.targets) // len(es)

    def live(group) -> int:
        tg = [(e, f, 0) for e, f in dag.targets if e in group]
        return peak_live(dag, closure(dag, tg)) + len(tg)

    if live(es) <= live_max:
        return (tuple(es),)
    for size in range(len(es) - 1, 0, -1):
        groups = [tuple(es[i:i + size]) for i in range(0, len(es), size)]
        if all(live(g) <= live_max for g in groups):
            return tuple(groups)
    del nF
    return tuple((e,) for e in es)


@lru_cache(maxsize=None)
def ef_groups(cls, live_max: int = 120):
    """Split the boundary targets (e, f), in build order, into consecutive
    groups whose closures (plus their accumulators) keep at most `live_max`
    values live: finer than e_groups, for classes where even one bra
    component's targets need more. Cached on disk (the search takes seconds
    for (ff|ff))."""
    import json
    from pathlib import Path
    cache = Path(__file__).resolve().parent / "_build" / "ef_groups.json"
    key = f"{''.join(map(str, cls))}:{live_max}"
    try:
        saved = json.loads(cache.read_text())
    except (OSError, ValueError):
        saved = {}
    if key in saved:
        return tuple(tuple((tuple(e), tuple(f)) for e, f in g) for g in saved[key])
    dag = vrr_dag(tuple(cls))

    def live(tg) -> int:
        return peak_live(dag, closure(dag, [(e, f, 0) for e, f in tg])) + len(tg)

    groups, cur = [], []
    for t in dag.targets:
        if cur and live(cur + [t]) > live_max:
            groups.append(tuple(cur))
            cur = [t]
        else:
            cur.append(t)
    groups.append(tuple(cur))
    saved[key] = [[[list(e), list(f)] for e, f in g] for g in groups]
    cache.parent.mkdir(exist_ok=True)
    cache.write_text(json.dumps(saved))
    return tuple(groups)
