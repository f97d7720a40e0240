"""The fused Fock build's pair stream and staged records (host side).

Restates upstream `eri/fock/launch_f.py:1-470` (the plan),
`ao/pairs.py:238-305,528-560` (the bound and pruning),
`eri/fock/kernel_f.py:529-628` (records, `pack`, `density_blocks`,
`close`) and `eri/fock/walk_f.py:177-210` (pair records), vectorised with
numpy.

THE BOUND. Per primitive pair of an ordered shell pair (a, b),
``G_[ab] = max_{cart pair} sum_p |E~^{ab}_p| I_p(zeta)`` with ``E~`` the
dressed Hermite coefficients ``c_a c_b (pi/zeta)^{3/2} E`` and
``I_p = sqrt((-1)^|p| [2p]^(0)(PQ = 0, theta^2 = zeta/2))``. At PQ = 0 the
raising recursion leaves only its ``(r_d - 1)`` terms, so in closed form
``I_p^2 = (2/sqrt(pi)) sqrt(zeta/2) zeta^|p| prod_d (2 p_d - 1)!! / (2|p| + 1)``.
The contracted ``G`` is the sum over ALL primitive pairs, before pruning.

PRUNING. Primitive pairs sorted by bound descending (stable); the longest
tail whose sum is ``< eps = tau / (2 Gmax)`` is deleted; a pair with nothing
left is deleted whole.

BINS, BATCHES, LAUNCHES. Bin key ``(la, lb, degree, mag)``, eq. 29
magnitudes ``tau^(i/n)``, ``mag`` = the largest i >= 1 with ``G < m_i``
else 0 (the upper-bound assignment); pairs inside a batch descending by
``G * dmax`` (stable); batch pairs ``I <= J`` launched unless
``mag_I + mag_J >= n`` (screened, checked ``Gmax_I Gmax_J < tau``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np

from .basis import Basis, cartesians, n_cart

REC = 8  # P(3) X(3) zeta U
MAX_PRIM_PAIRS = 128


# ---------------------------------------------------------------------------
# Hermite E coefficients, vectorised over primitive pairs (upstream
# mmd/hermite.py:45-79):  E^{00}_0 = exp(zeta X_PA X_PB);
#   E^{i+1,j}_t = E^{ij}_{t-1}/(2 zeta) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}
#   E^{i,j+1}_t = E^{ij}_{t-1}/(2 zeta) + X_PB E^{ij}_t + (t+1) E^{ij}_{t+1}

def e_axis(la: int, lb: int, zeta: np.ndarray, pa: np.ndarray, pb: np.ndarray) -> np.ndarray:
    """E[n, i, j, t] for i <= la, j <= lb, t <= la + lb (zero past i + j)."""
    n = zeta.shape[0]
    T = la + lb + 1
    E = np.zeros((n, la + 1, lb + 1, T + 1))
    h = 0.5 / zeta
    E[:, 0, 0, 0] = np.exp(zeta * pa * pb)
    for i in range(la + 1):
        if i > 0:
            prev = E[:, i - 1, 0]
            E[:, i, 0, :T] = pa[:, None] * prev[:, :T]
            E[:, i, 0, 1:T] += h[:, None] * prev[:, :T - 1]
            E[:, i, 0, :T - 1] += np.arange(1, T)[None, :] * prev[:, 1:T]
        for j in range(1, lb + 1):
            prev = E[:, i, j - 1]
            E[:, i, j, :T] = pb[:, None] * prev[:, :T]
            E[:, i, j, 1:T] += h[:, None] * prev[:, :T - 1]
            E[:, i, j, :T - 1] += np.arange(1, T)[None, :] * prev[:, 1:T]
    return E[..., :T]


def hermite_order(L: int) -> tuple[tuple[int, int, int], ...]:
    """Every t with |t| <= L: |t| ascending, then tx, then ty descending
    (upstream mmd/hermite.py:80-99)."""
    return tuple((tx, ty, n - tx - ty) for n in range(L + 1)
                 for tx in range(n, -1, -1) for ty in range(n - tx, -1, -1))


def _dfact_odd(n: int) -> int:
    out = 1
    for k in range(1, 2 * n, 2):
        out *= k
    return out


def hermite_bound_sq(zeta: np.ndarray, p) -> np.ndarray:
    """I_p^2 = (-1)^|p| [2p]^(0)(PQ = 0, theta^2 = zeta/2), closed form."""
    s = sum(p)
    df = _dfact_odd(p[0]) * _dfact_odd(p[1]) * _dfact_odd(p[2])
    return (2.0 / math.sqrt(math.pi)) * np.sqrt(zeta / 2.0) * zeta ** s * df / (2 * s + 1)


@dataclass
class PrimPairs:
    """Every primitive pair (i outer, j inner) of a list of ordered shell
    pairs of one class (la, lb)."""

    la: int
    lb: int
    pair: np.ndarray     # [n] index into the class's shell-pair list
    i: np.ndarray        # [n] primitive index in shell a
    j: np.ndarray        # [n] primitive index in shell b
    P: np.ndarray        # [n, 3]
    PA: np.ndarray       # [n, 3]
    zeta: np.ndarray     # [n]
    U: np.ndarray        # [n]  c_a c_b (pi/zeta)^{3/2} exp(-ab/zeta |AB|^2)
    bound: np.ndarray    # [n]  G_[ab]


def primitive_pairs(basis: Basis, shell_pairs: list[tuple[int, int]]) -> PrimPairs:
    """Primitive-pair data and bounds for ordered shell pairs of ONE class."""
    la = int(basis.shell_l[shell_pairs[0][0]])
    lb = int(basis.shell_l[shell_pairs[0][1]])
    rows_pair, rows_ia, rows_ib = [], [], []
    for k, (a, b) in enumerate(shell_pairs):
        na, nb = int(basis.shell_nprim[a]), int(basis.shell_nprim[b])
        ii, jj = np.meshgrid(np.arange(na), np.arange(nb), indexing="ij")
        rows_pair.append(np.full(na * nb, k))
        rows_ia.append(ii.ravel())
        rows_ib.append(jj.ravel())
    pair = np.concatenate(rows_pair)
    i = np.concatenate(rows_ia)
    j = np.concatenate(rows_ib)
    sp = np.array(shell_pairs)
    sa, sb = sp[pair, 0], sp[pair, 1]
    A = basis.shell_centre[sa]
    B = basis.shell_centre[sb]
    al = basis.exps[basis.shell_pstart[sa] + i]
    be = basis.exps[basis.shell_pstart[sb] + j]
    ca = basis.coefs[basis.shell_pstart[sa] + i]
    cb = basis.coefs[basis.shell_pstart[sb] + j]
    zeta = al + be
    P = (al[:, None] * A + be[:, None] * B) / zeta[:, None]
    PA = P - A
    PB = P - B
    ab2 = ((A - B) ** 2).sum(1)
    pref = ca * cb * (math.pi / zeta) ** 1.5
    U = pref * np.exp(-al * be / zeta * ab2)
    # the bound
    L = la + lb
    Ex = [e_axis(la, lb, zeta, PA[:, d], PB[:, d]) for d in range(3)]
    herm = hermite_order(L)
    Ip = np.stack([np.sqrt(hermite_bound_sq(zeta, t)) for t in herm], axis=1)  # [n, nh]
    best = np.zeros(zeta.shape[0])
    for a in cartesians(la):
        for b in cartesians(lb):
            s = np.zeros(zeta.shape[0])
            for h, t in enumerate(herm):
                if any(t[d] > a[d] + b[d] for d in range(3)):
                    continue
                e = Ex[0][:, a[0], b[0], t[0]] * Ex[1][:, a[1], b[1], t[1]] * Ex[2][:, a[2], b[2], t[2]]
                s += np.abs(pref * e) * Ip[:, h]
            best = np.maximum(best, s)
    return PrimPairs(la, lb, pair, i, j, P, PA, zeta, U, best)


def prune(bounds: np.ndarray, eps: float) -> np.ndarray:
    """Indices kept, bound-descending (stable); the longest tail summing to
    less than eps is dropped (upstream ao/pairs.py:545-560)."""
    order = np.argsort(-bounds, kind="stable")
    if eps <= 0.0:
        return order
    cut = len(order)
    tail = 0.0
    vals = bounds[order]
    # grow the tail while its (exact) sum stays below eps
    while cut:
        t = math.fsum(vals[cut - 1:])
        if t < eps:
            cut -= 1
        else:
            break
    del tail
    return order[:cut]


# ---------------------------------------------------------------------------
# eq. 29


def magnitudes(tau: float, n: int) -> tuple[float, ...]:
    return tuple(tau ** (i / n) for i in range(n))


def magnitude_index(G: float, mags) -> int:
    for i in range(len(mags) - 1, 0, -1):
        if G < mags[i]:
            return i
    return 0


def prelaunch_screened(i: int, j: int, n: int) -> bool:
    return Fraction(i, n) + Fraction(j, n) >= 1


# ---------------------------------------------------------------------------
# the plan


@dataclass
class FPair:
    a: int
    b: int
    la: int
    lb: int
    n_prim: int
    prims: np.ndarray        # kept primitive-pair indices (i * K_b + j), bound order
    G: float
    dmax: float
    key: float
    mag: int
    # staged record data for the kept primitive pairs, in `prims` order
    rec: np.ndarray = field(repr=False, default=None)  # [degree, 8]

    @property
    def degree(self) -> int:
        return len(self.prims)


@dataclass
class FBatch:
    la: int
    lb: int
    degree: int
    mag: int
    magnitude: float
    pairs: list

    @property
    def g_max(self) -> float:
        return max(p.G for p in self.pairs)


@dataclass(frozen=True)
class FLaunch:
    bra: int
    ket: int
    cls: tuple


@dataclass
class FPlan:
    tau: float
    n: int
    magnitudes: tuple
    eps: float
    g_max: float
    batches: list
    launches: list
    screened: list
    deleted: list
    n_prim_total: int

    def __iter__(self):
        return iter(self.launches)

    def __len__(self):
        return len(self.launches)

    @property
    def pairs(self):
        return [p for b in self.batches for p in b.pairs]


def canonical(basis: Basis, i: int, j: int) -> tuple[int, int]:
    """AO-order pair: la <= lb, lower shell index first on a tie."""
    li, lj = int(basis.shell_l[i]), int(basis.shell_l[j])
    return (i, j) if (li, i) <= (lj, j) else (j, i)


def unique_pairs(basis: Basis) -> list[tuple[int, int]]:
    n = basis.n_sh
    return [canonical(basis, i, j) for i in range(n) for j in range(i, n)]


def pair_tables(basis: Basis, D: np.ndarray | None = None):
    """Every canonical shell pair grouped by class: {(la, lb): (pairs, PrimPairs)}."""
    by_cls: dict = {}
    for a, b in unique_pairs(basis):
        by_cls.setdefault((int(basis.shell_l[a]), int(basis.shell_l[b])), []).append((a, b))
    return {c: (ps, primitive_pairs(basis, ps)) for c, ps in sorted(by_cls.items())}


def _block_absmax(basis: Basis, D: np.ndarray, a: int, b: int) -> float:
    oa, ob = int(basis.shell_ao[a]), int(basis.shell_ao[b])
    return float(np.abs(D[oa:oa + basis.shell_size[a], ob:ob + basis.shell_size[b]]).max())


def kept_pairs(basis: Basis, D: np.ndarray, tau: float, n: int, tables=None):
    """Bounds, pruning and the per-pair record data (shared by the upstream
    batch stream and the native builder). Returns (pairs, deleted, eps,
    g_max, n_prim_total, mags)."""
    tables = tables or pair_tables(basis)
    sums, bounds_at = {}, {}
    for cls, (ps, pp) in tables.items():
        # primitive rows are grouped by pair, pairs ascending
        at = np.searchsorted(pp.pair, np.arange(len(ps) + 1))
        bounds_at[cls] = at
        sums[cls] = np.array([math.fsum(pp.bound[at[k]:at[k + 1]]) for k in range(len(ps))])
    g_max = max((s.max() for s in sums.values() if len(s)), default=0.0)
    eps = tau / (2.0 * g_max) if tau > 0.0 and g_max > 0.0 else 0.0
    mags = magnitudes(tau, n) if tau > 0 else tuple([1.0] + [0.0] * (n - 1))
    pairs, deleted, n_prim_total = [], [], 0
    for cls, (ps, pp) in tables.items():
        kb_of = basis.shell_nprim
        at = bounds_at[cls]
        for k, (a, b) in enumerate(ps):
            rows = np.arange(at[k], at[k + 1])
            n_prim_total += len(rows)
            keep_local = prune(pp.bound[rows], eps)
            if len(keep_local) == 0:
                deleted.append((a, b))
                continue
            sel = rows[keep_local]
            prim_idx = pp.i[sel] * int(kb_of[b]) + pp.j[sel]
            rec = np.concatenate([pp.P[sel], pp.PA[sel], pp.zeta[sel, None], pp.U[sel, None]], axis=1)
            G = float(sums[cls][k])
            dmax = _block_absmax(basis, D, a, b)
            pairs.append(FPair(a, b, cls[0], cls[1], len(rows), prim_idx, G, dmax, G * dmax,
                               magnitude_index(G, mags) if tau > 0 else 0, np.ascontiguousarray(rec)))
    return pairs, deleted, eps, g_max, n_prim_total, mags


def f_launches(basis: Basis, D: np.ndarray, *, tau: float, n: int, tables=None) -> FPlan:
    """The upstream batch stream (launch_f.py:390-470)."""
    if n < 1:
        raise ValueError("n must be >= 1")
    if not np.array_equal(D, D.T):
        raise ValueError("the density must be exactly symmetric")
    pairs, deleted, eps, g_max, n_prim_total, mags = kept_pairs(basis, D, tau, n, tables)
    for p in pairs:
        if p.degree > MAX_PRIM_PAIRS:
            raise ValueError(f"pair {(p.a, p.b)} keeps {p.degree} primitive pairs, above "
                             f"MAX_PRIM_PAIRS = {MAX_PRIM_PAIRS}")
    bins: dict = {}
    for p in pairs:
        bins.setdefault((p.la, p.lb, p.degree, p.mag), []).append(p)
    batches = []
    for k in sorted(bins):
        ps = sorted(bins[k], key=lambda p: -p.key)  # stable descending
        batches.append(FBatch(k[0], k[1], k[2], k[3], mags[k[3]], ps))
    launches, screened = [], []
    for bi, bra in enumerate(batches):
        for ki in range(bi, len(batches)):
            ket = batches[ki]
            if tau > 0 and prelaunch_screened(bra.mag, ket.mag, n):
                if not bra.g_max * ket.g_max < tau:
                    raise ValueError(f"batch pair ({bi}, {ki}) holds G_ab G_cd >= tau: "
                                     "eq. 29 rounded past the bound")
                screened.append((bi, ki))
            else:
                launches.append(FLaunch(bi, ki, (bra.la, bra.lb, ket.la, ket.lb)))
    return FPlan(tau, n, mags, eps, g_max, batches, launches, screened, deleted, n_prim_total)


# ---------------------------------------------------------------------------
# staged records


@dataclass
class Packed:
    n: int
    degree: int
    prim: np.ndarray   # [n * degree * 8]
    sep: np.ndarray    # [n * 3]
    G: np.ndarray      # [n]
    sh: np.ndarray     # [n * 2] int32
    ao: np.ndarray     # [n * 2] int32


def sub_degree(k: int, chunk: int) -> int:
    """The sub-pair degree of a degree-k side: the largest divisor of k not
    above chunk (k itself when chunk is 0 or >= k)."""
    if chunk <= 0 or chunk >= k:
        return k
    return max(d for d in range(1, chunk + 1) if k % d == 0)


def pack(basis: Basis, pairs: list, prim_chunk: int = 0) -> Packed:
    """One side of a launch (upstream kernel_f.py:572-606). prim_chunk > 0
    packs each pair as sub-pairs of sub_degree(degree, prim_chunk) of its
    primitive records, each with the pair's G, separation, shells and AOs:
    the per-class kernels take one lane per quartet and its whole primitive
    loop, and the sub-quartets sum to the quartet (the HRR and digestion are
    linear in the contracted boundary, the screen reads the pair's G and
    shells, and on a diagonal launch the canonical half over sub-pairs with
    the k == bra halving gives the pair with itself exactly once)."""
    if not pairs:
        raise ValueError("an empty side has no pairs to stage")
    degrees = {p.degree for p in pairs}
    if len(degrees) != 1:
        raise ValueError(f"degrees {sorted(degrees)}: a batch is degree-uniform")
    (degree,) = degrees
    if not 1 <= degree <= MAX_PRIM_PAIRS:
        raise ValueError(f"degree {degree} outside 1..{MAX_PRIM_PAIRS}")
    d = sub_degree(degree, prim_chunk)
    m = degree // d
    prim = np.concatenate([p.rec.ravel() for p in pairs])
    sep = np.repeat(np.array([basis.shell_centre[p.a] - basis.shell_centre[p.b] for p in pairs]), m, axis=0)
    G = np.repeat(np.array([p.G for p in pairs]), m)
    sh = np.repeat(np.array([(p.a, p.b) for p in pairs], dtype=np.int32), m, axis=0).ravel()
    ao = np.repeat(np.array([(basis.shell_ao[p.a], basis.shell_ao[p.b]) for p in pairs], dtype=np.int32),
                   m, axis=0).ravel()
    return Packed(len(pairs) * m, d, np.ascontiguousarray(prim), np.ascontiguousarray(sep.ravel()), G, sh, ao)


def density_blocks(basis: Basis, D: np.ndarray) -> np.ndarray:
    """Dsh[s][t] = max |D| over shell block (s, t) (upstream kernel_f.py:608-618)."""
    n = basis.n_sh
    out = np.zeros((n, n))
    for s in range(n):
        for t in range(n):
            out[s, t] = _block_absmax(basis, D, s, t)
    return out


def close(F: np.ndarray) -> np.ndarray:
    """F + F^T (upstream kernel_f.py:621-628)."""
    return F + F.T


def launch_args(basis: Basis, plan: FPlan, L: FLaunch, D: np.ndarray, Dsh: np.ndarray,
                prim_chunk: int = 0) -> dict:
    """The per-class entry's arguments for launch L, by upstream name
    (upstream walk_f.py:385-398), host numpy; prim_chunk as in pack."""
    pb = pack(basis, plan.batches[L.bra].pairs, prim_chunk)
    pk = pack(basis, plan.batches[L.ket].pairs, prim_chunk)
    _, lb, _, ld = L.cls
    a = {"n_bra": pb.n, "n_ket": pk.n, "k_ab": pb.degree, "k_cd": pk.degree,
         "n_sh": basis.n_sh, "n_ao": basis.n_ao, "diag": int(L.bra == L.ket),
         "tau": plan.tau, "bra_prim": pb.prim, "G_ab": pb.G, "bra_sh": pb.sh,
         "bra_ao": pb.ao, "ket_prim": pk.prim, "G_cd": pk.G, "ket_sh": pk.sh,
         "ket_ao": pk.ao, "Dsh": np.ascontiguousarray(Dsh.ravel()),
         "D": np.ascontiguousarray(D.ravel())}
    if lb:
        a["AB"] = pb.sep
    if ld:
        a["CD"] = pk.sep
    return a


def n_eri(cls) -> int:
    out = 1
    for x in cls:
        out *= n_cart(x)
    return out
