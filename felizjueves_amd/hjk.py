"""The split HGP J / K build: the ordered pair stream (upstream
`eri/hjk/launch_hjk.py:1-342`) and its walker (upstream
`eri/hjk/walk_hjk.py:1-367`), on the library's split kernels fj_hj_<rep> /
fj_hk_<rep> (one per orbit of the 8-fold ERI symmetry, the canonical class
mode).

    hjk_launches(basis, D, *, tau, mode="canonical", screen="j") -> HPlan
    walk_hjk(plan, basis, D, *, run, threads=64) -> HJKWalk
    GpuHJKRunner()        run(kernel, grid, block, args, X) through the C-ABI

Every ORDERED shell pair (a, b) is a pair (no symmetry pruning, no
primitive pruning: degree K_a K_b, refused above 128), with G = the sum of
its primitive pairs' bounds. Pairs are binned by (la, lb, degree) and
sorted by G descending (stably); every ORDERED batch pair (I, J) is one
launch, routed to the representative of its class's orbit with the
orientation that carries it there. The J kernel deposits J_ab += D_cd
(ab|cd), the K kernel K_ac += D_bd (ab|cd); each skips a quartet when its
one-block key G_ab G_cd max|D_block| < tau (J: the (c, d) block, K: the
(b, d) block; a skip, never a break). Summed over every launch, J and K are
the every-ordered-quartet sweeps, restricted to the kept quartets.

Only the canonical class mode is built (the upstream's literal default):
the ordered mode's one kernel per ordered class is refused.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import pairs as P
from .basis import Basis
from .orbits import Orientation, to_representative

MODES = ("canonical",)
SCREENS = ("j", "k")
MAX_PRIM_PAIRS = 128
SPLIT_LMAX = 2          # the split kernels are built through (dd|dd)


def j_key(q_ab: float, q_cd: float, dmax_cd: float) -> float:
    """J's one-block key Q_ab Q_cd max|D_cd| (launch_hjk.py:148-150)."""
    return q_ab * q_cd * dmax_cd


def k_key(q_ab: float, q_cd: float, dmax_bd: float) -> float:
    """K's one-block key Q_ab Q_cd max|D_bd| (launch_hjk.py:153-155)."""
    return q_ab * q_cd * dmax_bd


@dataclass(frozen=True)
class HPair:
    a: int
    b: int
    la: int
    lb: int
    degree: int
    G: float


@dataclass(frozen=True)
class HBatch:
    la: int
    lb: int
    degree: int
    pairs: tuple


@dataclass(frozen=True)
class HLaunch:
    bra: int                       # batch indices
    ket: int
    ordered: tuple                 # (la_I, lb_I, la_J, lb_J)
    orientation: Orientation

    @property
    def cls(self) -> tuple:
        """The class of the kernel that runs this launch."""
        return self.orientation.rep


@dataclass
class HPlan:
    tau: float
    mode: str
    screen: str
    batches: tuple
    launches: tuple
    G: dict                        # (a, b) -> G of the ordered pair
    dmax: np.ndarray               # [n_sh, n_sh] max |D| per shell block
    prim_chunk: int = 0            # pack pairs as sub-pairs of <= this many primitive pairs (0: whole)

    def __iter__(self):
        return iter(self.launches)

    def __len__(self) -> int:
        return len(self.launches)

    @property
    def pairs(self) -> tuple:
        return tuple(p for b in self.batches for p in b.pairs)

    @property
    def classes(self) -> tuple:
        return tuple(sorted({L.cls for L in self.launches}))

    def visits(self):
        """Every ordered quartet (launch index, (a, b, c, d), key, skipped)."""
        for li, L in enumerate(self.launches):
            for p in self.batches[L.bra].pairs:
                for q in self.batches[L.ket].pairs:
                    a, b, c, d = p.a, p.b, q.a, q.b
                    key = (j_key(p.G, q.G, self.dmax[c, d]) if self.screen == "j"
                           else k_key(p.G, q.G, self.dmax[b, d]))
                    yield li, (a, b, c, d), key, key < self.tau

    def counts(self) -> dict:
        n_q = sum(len(self.batches[L.bra].pairs) * len(self.batches[L.ket].pairs)
                  for L in self.launches)
        return {"pairs": len(self.pairs), "batches": len(self.batches),
                "launches": len(self.launches), "classes": len(self.classes), "quartets": n_q}


def _density(basis: Basis, D) -> np.ndarray:
    D = np.asarray(D, dtype=np.float64)
    n = basis.n_ao
    if D.shape != (n, n):
        raise ValueError(f"D is not an {n} x {n} matrix")
    if not np.all(np.isfinite(D)):
        raise ValueError("D has a non-finite entry")
    if not np.array_equal(D, D.T):
        raise ValueError("D is not symmetric")
    return D


# default sub-pair size: at most 2 x 2 primitive quartets per lane. Measured
# on the GPU (J of the three-atom s,s,p,d system / the 10-atom cluster, ms):
# whole pairs 3,182 / 16,797; 16: 1,042 / 7,539; 9: 483 / 3,721;
# 4: 180 / 1,399; 3: 157 / 1,103; 2: 136 / 973; 1: 137 / 1,340 (every
# sub-quartet pays its own HRR).
PRIM_CHUNK = 2


def hjk_launches(basis: Basis, D, *, tau: float, mode: str = "canonical",
                 screen: str = "j", prim_chunk: int = PRIM_CHUNK) -> HPlan:
    """The ordered pair stream of `basis` and the symmetric density D.

    prim_chunk: the kernels take one lane per contracted quartet and its
    whole primitive loop, so a side's pairs are packed as sub-pairs of at
    most `prim_chunk` primitive pairs (the largest divisor of the batch's
    degree not above it; 0 packs whole pairs, the upstream's packing). A
    sub-pair carries its pair's G, shells and AOs, so every sub-quartet of a
    quartet passes or fails the screen together, and the HRR and digestion
    are linear in the contracted boundary: the sub-quartets' J or K sum to
    the quartet's (each through its own FP64 HRR and atomics)."""
    if mode not in MODES:
        raise ValueError(f"mode = {mode!r}: only the canonical class mode is built ({MODES})")
    if screen not in SCREENS:
        raise ValueError(f"screen = {screen!r} is not one of {SCREENS}")
    if not (isinstance(tau, (int, float)) and not isinstance(tau, bool) and math.isfinite(tau)
            and tau >= 0.0):
        raise ValueError(f"tau = {tau!r} is not a finite float >= 0")
    if basis.l_max > SPLIT_LMAX:
        raise ValueError(f"shells up to l = {basis.l_max}: the split kernels are built through "
                         f"l = {SPLIT_LMAX}")
    D = _density(basis, D)
    ns = basis.n_sh
    by_cls: dict = {}
    for a in range(ns):
        for b in range(ns):
            by_cls.setdefault((int(basis.shell_l[a]), int(basis.shell_l[b])), []).append((a, b))
    G: dict = {}
    pairs = []
    for (la, lb), sps in by_cls.items():
        pp = P.primitive_pairs(basis, sps)
        for k, (a, b) in enumerate(sps):
            g = pp.bound[pp.pair == k]
            deg = len(g)
            if deg > MAX_PRIM_PAIRS:
                raise ValueError(f"shell pair ({a}, {b}) has degree {deg} > MAX_PRIM_PAIRS = "
                                 f"{MAX_PRIM_PAIRS}, the kernels' staging cap")
            G[a, b] = math.fsum(g)
            pairs.append(HPair(a, b, la, lb, deg, G[a, b]))
    pairs.sort(key=lambda p: (p.a, p.b))            # pair order: a outer, b inner
    bins: dict = {}
    for p in pairs:
        bins.setdefault((p.la, p.lb, p.degree), []).append(p)
    batches = tuple(HBatch(*k, tuple(sorted(bins[k], key=lambda p: -p.G))) for k in sorted(bins))
    launches = []
    for bi, bra in enumerate(batches):
        for ki, ket in enumerate(batches):
            ordered = (bra.la, bra.lb, ket.la, ket.lb)
            launches.append(HLaunch(bi, ki, ordered, to_representative(ordered)))
    if not (isinstance(prim_chunk, int) and prim_chunk >= 0):
        raise ValueError(f"prim_chunk = {prim_chunk!r} is not an int >= 0")
    return HPlan(float(tau), mode, screen, batches, tuple(launches), G, P.density_blocks(basis, D),
                 prim_chunk)


sub_degree = P.sub_degree   # the sub-pair degree of a degree-k batch (pairs.sub_degree)


def kernel_sides(L: HLaunch):
    """The kernel's two sides: (batch index, swapped) for its bra and its
    ket. The ordered quartet q = (a, b, c, d) runs as x[j] = q[shells[j]];
    every group element keeps the pairs whole, so the kernel's bra is one
    batch's pairs (read as (b, a) when swapped) and its ket the other's."""
    s0, s2 = L.orientation.shells[0], L.orientation.shells[2]
    if s0 in (0, 1):
        return (L.bra, s0 == 1), (L.ket, s2 == 3)
    return (L.ket, s0 == 3), (L.bra, s2 == 1)


def side_records(basis: Basis, plan: HPlan, batch: HBatch, swapped: bool) -> dict:
    """One kernel side: per pair (x, y) (= (b, a) when swapped) its staged
    records built on x, (P, P - X, zeta, U), G of the ordered pair (x, y),
    sep = X - Y, shells and first AOs; degree-uniform. With the plan's
    prim_chunk, each pair as its sub-pairs (sub_degree records each, the
    pair's G, sep, shells and AOs repeated)."""
    xy = [(p.b, p.a) if swapped else (p.a, p.b) for p in batch.pairs]
    la, lb = (batch.lb, batch.la) if swapped else (batch.la, batch.lb)
    pp = P.primitive_pairs(basis, xy)
    k = batch.degree
    prim = np.empty((len(xy), k, 8))
    for i in range(len(xy)):
        sel = pp.pair == i
        prim[i, :, 0:3] = pp.P[sel]
        prim[i, :, 3:6] = pp.PA[sel]
        prim[i, :, 6] = pp.zeta[sel]
        prim[i, :, 7] = pp.U[sel]
    c = basis.shell_centre
    # sub-pairs: pair i's k records as m consecutive runs of d (d m = k)
    d = sub_degree(k, plan.prim_chunk)
    m = k // d
    rep = lambda a: np.repeat(a, m, axis=0)  # noqa: E731
    return {"prim": prim.reshape(-1, 8), "k": d, "lb": lb,
            "sep": rep(np.array([c[x] - c[y] for x, y in xy], dtype=np.float64).reshape(-1, 3)),
            "G": rep(np.array([plan.G[x, y] for x, y in xy])),
            "sh": rep(np.array(xy, dtype=np.int32).reshape(-1, 2)),
            "ao": rep(np.array([(basis.shell_ao[x], basis.shell_ao[y]) for x, y in xy],
                               dtype=np.int32).reshape(-1, 2))}


def launch_args(basis: Basis, plan: HPlan, L: HLaunch, D: np.ndarray, Dsh: np.ndarray) -> dict:
    """The split kernel's arguments for launch L, by upstream parameter name
    (kernel_hj.py:305-315), host numpy."""
    (bi, bs), (ki, ks) = kernel_sides(L)
    b = side_records(basis, plan, plan.batches[bi], bs)
    k = side_records(basis, plan, plan.batches[ki], ks)
    a = {"n_bra": len(b["G"]), "n_ket": len(k["G"]), "k_ab": b["k"], "k_cd": k["k"],
         "n_sh": basis.n_sh, "n_ao": basis.n_ao, "perm": L.orientation.perm, "tau": plan.tau,
         "bra_prim": b["prim"], "G_ab": b["G"], "bra_sh": b["sh"], "bra_ao": b["ao"],
         "ket_prim": k["prim"], "G_cd": k["G"], "ket_sh": k["sh"], "ket_ao": k["ao"],
         "Dsh": np.ascontiguousarray(Dsh.ravel()), "D": np.ascontiguousarray(D.ravel())}
    if b["lb"]:
        a["AB"] = b["sep"]
    if k["lb"]:
        a["CD"] = k["sep"]
    return a


@dataclass
class HJKWalk:
    X: np.ndarray          # J (screen "j") or K (screen "k")
    screen: str
    mode: str
    n_launches: int
    kernels: tuple


def walk_hjk(plan: HPlan, basis: Basis, D, *, run, threads: int = 64) -> HJKWalk:
    """Execute every launch of `plan` through `run(kernel, grid, block, args,
    X)` (kernel = ("hj" | "hk", representative class)) and return the
    accumulated J or K (nothing closed or mirrored)."""
    D = _density(basis, D)
    if plan.mode not in MODES or plan.screen not in SCREENS:
        raise ValueError("not a split plan of a built mode and screen")
    n_sh = basis.n_sh
    seen = sorted((p.a, p.b) for p in plan.pairs)
    if seen != [(a, b) for a in range(n_sh) for b in range(n_sh)]:
        raise ValueError("the plan's pairs are not every ordered shell pair exactly once")
    nb = len(plan.batches)
    if sorted((L.bra, L.ket) for L in plan.launches) != [(i, j) for i in range(nb) for j in range(nb)]:
        raise ValueError("the plan's launches are not every ordered batch pair exactly once")
    kind = "hj" if plan.screen == "j" else "hk"
    X = run.zeros(basis.n_ao)
    kernels = set()
    for L in plan.launches:
        bra, ket = plan.batches[L.bra], plan.batches[L.ket]
        if L.ordered != (bra.la, bra.lb, ket.la, ket.lb):
            raise ValueError(f"launch {L} names the wrong ordered class")
        o = to_representative(L.ordered)
        if (o.rep, o.shells) != (L.cls, L.orientation.shells):
            raise ValueError(f"launch {L} is not routed to its orbit representative")
        args = launch_args(basis, plan, L, D, plan.dmax)
        kernels.add((kind, L.cls))
        run((kind, L.cls), args["n_bra"], (threads, 1), run.put(args), X)
    return HJKWalk(run.host(X), plan.screen, plan.mode, len(plan.launches), tuple(sorted(kernels)))


class GpuHJKRunner:
    """walk_hjk's `run` on the library: fj_hj_<rep> / fj_hk_<rep>."""

    def __init__(self):
        from . import native
        self.native = native
        self.dev = native.device()

    def zeros(self, n):
        import torch
        return torch.zeros((n, n), dtype=torch.float64, device=self.dev)

    def put(self, args: dict) -> dict:
        return {k: (self.native.to_dev(v) if isinstance(v, np.ndarray) else v) for k, v in args.items()}

    def __call__(self, kernel, grid, block, args, X):
        import ctypes as C
        kind, cls = kernel
        la, lb, lc, ld = cls
        nat = self.native
        fn = getattr(nat.lib(), f"fj_{kind}_" + "".join(map(str, cls)))
        fn.restype = C.c_int
        a = args
        vals = [C.c_void_p(nat.stream_handle()), C.c_int(block[0])]
        vals += [C.c_int(a[k]) for k in ("n_bra", "n_ket", "k_ab", "k_cd", "n_sh", "n_ao", "perm")]
        vals += [C.c_double(a["tau"]), C.c_void_p(nat.ptr(a["bra_prim"]))]
        if lb:
            vals.append(C.c_void_p(nat.ptr(a["AB"])))
        vals += [C.c_void_p(nat.ptr(a[k])) for k in ("G_ab", "bra_sh", "bra_ao", "ket_prim")]
        if ld:
            vals.append(C.c_void_p(nat.ptr(a["CD"])))
        vals += [C.c_void_p(nat.ptr(a[k])) for k in ("G_cd", "ket_sh", "ket_ao", "Dsh", "D")]
        vals.append(C.c_void_p(nat.ptr(X)))
        nat.check(fn(*vals), fn.__name__)

    def host(self, X) -> np.ndarray:
        import torch
        torch.cuda.synchronize()
        return X.cpu().numpy()


STREAMS = 4   # streams of a staged build (hardware queues per process on the MI355X boxes)


class HJKBuild:
    """A split J or K build staged on the device for repeated use: the
    plan's launch arguments on the device (each (batch, orientation) side
    once, D and its block maxima once), `run()` issuing every launch into
    one n_ao x n_ao accumulator on one stream, and `capture()` / `replay()`
    the same build as one HIP graph (the accumulator cleared by torch's fill
    kernel, no memset node). The build is for the density it was staged
    with; the result is walk_hjk's (nothing closed or mirrored).

    `quartets_computed` and `flops` count the quartets that pass the
    in-kernel screen (G_ab G_cd times the density block, D_cd for J and D_bd
    for K) and their algorithmic FLOPs (hgp.cost of the kernel's class)."""

    def __init__(self, basis: Basis, D, *, tau: float = 1e-12, screen: str = "j",
                 threads: int = 64, plan: HPlan | None = None):
        from . import native
        from .hgp import cost
        D = _density(basis, D)
        self.plan = plan if plan is not None else hjk_launches(basis, D, tau=tau, screen=screen)
        if self.plan.screen != screen:
            raise ValueError("the plan is for the other screen")
        self.basis, self.screen, self.threads = basis, screen, threads
        self.kind = "hj" if screen == "j" else "hk"
        self.runner = GpuHJKRunner()
        plan = self.plan
        dD = native.to_dev(np.ascontiguousarray(D.ravel()))
        dDsh = native.to_dev(np.ascontiguousarray(plan.dmax.ravel()))
        sides: dict = {}

        def side(bi, sw):
            if (bi, sw) not in sides:
                sides[(bi, sw)] = {k: (native.to_dev(v) if isinstance(v, np.ndarray) else v)
                                   for k, v in side_records(basis, plan, plan.batches[bi], sw).items()}
            return sides[(bi, sw)]
        self.staged, n_q, fl = [], 0, 0.0
        for L in plan.launches:
            (bi, bs), (ki, ks) = kernel_sides(L)
            b, k = side(bi, bs), side(ki, ks)
            a = {"n_bra": len(b["G"]), "n_ket": len(k["G"]), "k_ab": b["k"], "k_cd": k["k"],
                 "n_sh": basis.n_sh, "n_ao": basis.n_ao, "perm": L.orientation.perm,
                 "tau": plan.tau, "bra_prim": b["prim"], "G_ab": b["G"], "bra_sh": b["sh"],
                 "bra_ao": b["ao"], "ket_prim": k["prim"], "G_cd": k["G"], "ket_sh": k["sh"],
                 "ket_ao": k["ao"], "Dsh": dDsh, "D": dD}
            if b["lb"]:
                a["AB"] = b["sep"]
            if k["lb"]:
                a["CD"] = k["sep"]
            self.staged.append(((self.kind, L.cls), a))
            bra, ket = plan.batches[L.bra], plan.batches[L.ket]
            gb = np.array([p.G for p in bra.pairs])
            gk = np.array([q.G for q in ket.pairs])
            if screen == "j":
                dm = np.array([plan.dmax[q.a, q.b] for q in ket.pairs])[None, :]
            else:
                dm = plan.dmax[np.array([p.b for p in bra.pairs])[:, None],
                               np.array([q.b for q in ket.pairs])[None, :]]
            live = int(np.count_nonzero(gb[:, None] * gk[None, :] * dm >= plan.tau))
            n_q += live
            fl += live * cost(L.cls).flops(bra.degree, ket.degree)
        self.quartets_computed, self.flops = n_q, fl
        self.X = self.runner.zeros(basis.n_ao)
        self.graph = None

    def run(self, n_streams: int = 1):
        """Every launch of the build into X (cleared first); returns X (on
        the device). The launches are small and independent (they add into X
        with atomics), so with n_streams > 1 they go round-robin to that many
        side streams, forked from and joined to the current stream (the
        library keeps one workspace per stream for the classes that need
        one, so launches on different streams may overlap)."""
        import torch
        self.X.zero_()
        if n_streams <= 1:
            for kern, a in self.staged:
                self.runner(kern, a["n_bra"], (self.threads, 1), a, self.X)
            return self.X
        main = torch.cuda.current_stream()
        if len(getattr(self, "_streams", ())) != n_streams:
            self._streams = [torch.cuda.Stream() for _ in range(n_streams)]
        for st in self._streams:
            st.wait_stream(main)
        for i, (kern, a) in enumerate(self.staged):
            with torch.cuda.stream(self._streams[i % n_streams]):
                self.runner(kern, a["n_bra"], (self.threads, 1), a, self.X)
        for st in self._streams:
            main.wait_stream(st)
        return self.X

    def capture(self, n_streams: int = 1):
        """Capture one build into a HIP graph on a side stream, after an
        eager build on the same streams: the library's workspace is per
        stream and allocated on first use, which a capture may not do."""
        import torch
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.run(n_streams)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.run(n_streams)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()
        return self.X
