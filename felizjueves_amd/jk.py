"""Host drivers of the MMD families: the Schwarz matrix, the J chain
(DQ -> [p|q] -> JAB) and the K exchange build, all on the device.

Records (upstream `ao/density.py:151`, `eri/schwarz.py:152-176`): per
primitive pair of an ordered shell pair (a, b): ``zeta = alpha + beta``,
``P = (alpha A + beta B)/zeta``, ``X_PA = P - A``, ``X_PB = P - B`` and the
dressing ``s = c_a c_b (pi/zeta)^{3/2}``; the Hermite coefficients (with
their ``exp(-alpha beta/zeta |AB|^2)``) are formed in the kernels.

J chain (upstream `eri/j_coulomb/j.py:1-60,254-326`, `drive.py:367-594`):
unique ket pairs with weight 2 off the diagonal, 1 on it, folded into D_q;
kets sorted descending by ``Iq Dmax`` (Dmax read back, as upstream's one
round trip); one [p|q] launch per (bra class, ket class) into its own J_p
buffer; one JAB launch per buffer, atomically added into J with the mirror
and diagonal rules.

K (upstream `eri/k_exchange/kernel_k.py:1-110`, `launch_k.py:308-479`):
``K_ab = sum_cd D_cd (ac|bd)``; bra records of ordered pairs (a, c), ket
records of ordered pairs (b, d) of ONE b shell per launch, sorted by Iq; the
kernel's per-bra-record blocks K[rec, a, b] are scattered into K by a small
kernel.

Bounds (upstream `ao/pairs.py:1-120`): ``I_p(zeta)`` closed form as in
`pairs.hermite_bound_sq`; a record's ``I = sum_p I_p`` over
``hermite_order(L)``.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import native
from .basis import Basis, n_cart
from .pairs import hermite_bound_sq, hermite_order


K_LMAX = 3   # the K kernels through f (straight-line through d, the f classes in loop form)
K_STRAIGHT_LMAX = 2   # the straight-line K kernels (and their FP32-tower twins fj_k32_*)


def _records(basis: Basis, a: int, b: int):
    """(P [n,3], zeta [n], s [n], XPA [n,3], XPB [n,3]) of ordered pair (a, b)."""
    A, B = basis.shell_centre[a], basis.shell_centre[b]
    ia = slice(basis.shell_pstart[a], basis.shell_pstart[a] + basis.shell_nprim[a])
    ib = slice(basis.shell_pstart[b], basis.shell_pstart[b] + basis.shell_nprim[b])
    al, be = basis.exps[ia], basis.exps[ib]
    ca, cb = basis.coefs[ia], basis.coefs[ib]
    z = (al[:, None] + be[None, :]).ravel()
    P = ((al[:, None, None] * A + be[None, :, None] * B).reshape(-1, 3)) / z[:, None]
    s = (ca[:, None] * cb[None, :]).ravel() * (math.pi / z) ** 1.5
    return P, z, s, P - A, P - B


def record_bound(zeta: np.ndarray, L: int) -> np.ndarray:
    return sum(np.sqrt(hermite_bound_sq(zeta, t)) for t in hermite_order(L))


def _c(x):
    return C.c_void_p(native.ptr(x))


def _marshal(args):
    """ctypes values for `args`: a device tensor becomes its pointer, a
    ctypes value passes through. The caller keeps `args` (and so every
    tensor) alive until the launch is enqueued: a tensor freed earlier
    would hand its block to the next allocation on the stream, whose copy
    lands before the kernel reads it."""
    out = []
    for a in args:
        if hasattr(a, "data_ptr"):
            out.append(C.c_void_p(a.data_ptr()))
        else:
            out.append(a)
    return out


def _call(name: str, *args) -> None:
    """Launch entry `name` on the current stream. Device arguments are passed
    as tensors and stay referenced by `args` through the call."""
    fn = getattr(native.lib(), name)
    fn.restype = C.c_int
    native.check(fn(C.c_void_p(native.stream_handle()), *_marshal(args)), name)


def _canon(basis: Basis, i: int, j: int):
    li, lj = int(basis.shell_l[i]), int(basis.shell_l[j])
    return (i, j) if (li, i) <= (lj, j) else (j, i)


def _pair_classes(basis: Basis):
    out: dict = {}
    for i in range(basis.n_sh):
        for j in range(i, basis.n_sh):
            a, b = _canon(basis, i, j)
            out.setdefault((int(basis.shell_l[a]), int(basis.shell_l[b])), []).append((a, b))
    return dict(sorted(out.items()))


# ---------------------------------------------------------------------------


def schwarz_tables(basis: Basis) -> dict:
    """Per pair class (la, lb): the records of its shell pairs (numpy), the
    record range of each pair and the AO offsets of its two shells (upstream
    eri/schwarz.py:152-176 builds the same records)."""
    out = {}
    for (la, lb), prs in _pair_classes(basis).items():
        recs = [_records(basis, a, b) for a, b in prs]
        n_rec = np.cumsum([0] + [len(r[1]) for r in recs])
        cat = lambda k: np.concatenate([r[k] for r in recs])  # noqa: E731
        out[(la, lb)] = dict(
            n_pair=len(prs), n_rec=int(n_rec[-1]), P=cat(0), z=cat(1), s=cat(2), XA=cat(3),
            XB=cat(4), rec0=n_rec[:-1].astype(np.int32), rec1=n_rec[1:].astype(np.int32),
            oa=np.array([basis.shell_ao[a] for a, _ in prs], dtype=np.int32),
            ob=np.array([basis.shell_ao[b] for _, b in prs], dtype=np.int32))
    return out


def schwarz_matrix(basis: Basis, threads: int = 64):
    """Q[m, n] = sqrt(max((mn|mn), 0)) on the device (torch float64); pair
    classes through (ff), the two past la + lb = 4 on the loop-form kernel."""
    import torch
    dev = native.device()
    Q = torch.zeros((basis.n_ao, basis.n_ao), dtype=torch.float64, device=dev)
    for (la, lb), t in schwarz_tables(basis).items():
        d = {k: native.to_dev(v) for k, v in t.items() if isinstance(v, np.ndarray)}
        _call(f"fj_schwarz_{la}{lb}", C.c_int(threads), C.c_int(t["n_pair"]), C.c_int(t["n_rec"]),
              C.c_int(basis.n_ao), d["z"], d["P"], d["XA"], d["XB"], d["s"], d["rec0"], d["rec1"],
              d["oa"], d["ob"], Q)
    return Q


# ---------------------------------------------------------------------------


def _block_index(basis: Basis, prs, l1: int, l2: int) -> np.ndarray:
    """Flat indices into D (row-major n_ao x n_ao) of the (l1, l2) blocks of
    the shell pairs prs, one row of n_cart(l1) * n_cart(l2) per pair."""
    n = basis.n_ao
    r = np.arange(n_cart(l1))[:, None] * n + np.arange(n_cart(l2))[None, :]
    return np.array([basis.shell_ao[c] * n + basis.shell_ao[d] + r.ravel() for c, d in prs],
                    dtype=np.int64).reshape(len(prs), -1)


def _dev_density(D):
    """D as a contiguous float64 device tensor (a host array is copied)."""
    import torch
    if isinstance(D, torch.Tensor):
        return D.to(device=native.device(), dtype=torch.float64).contiguous()
    return native.to_dev(np.ascontiguousarray(D, dtype=np.float64))


class CoulombChain:
    """The J chain for one basis; `run(D)` returns J on the device. The
    pair records and their device tables are built once here; `run` only
    gathers D's blocks on the device and launches (D a host array or a
    device tensor).

    mixed: the J mixed lever (upstream eri/j_coulomb/levers/mixed.py:82-191)
    -- every [p|q] launch evaluates the Hermite recurrence in FP32
    (fj_pq32_*), with Boys, D_q, the J_p accumulation and JAB in FP64."""

    def __init__(self, basis: Basis, tau: float = 0.0, mixed: bool = False):
        self.basis = basis
        self.tau = float(tau)
        self.mixed = bool(mixed)
        self.classes = _pair_classes(basis)
        b = basis
        self.side = {}
        for (l1, l2), prs in self.classes.items():
            recs = [_records(b, c, d) for c, d in prs]
            cat = lambda k: np.concatenate([r[k] for r in recs])  # noqa: E731
            P, z, s, X1, X2 = (cat(k) for k in range(5))
            pair = np.concatenate([np.full(len(r[1]), k) for k, r in enumerate(recs)]).astype(np.int32)
            w = np.concatenate([np.full(len(r[1]), 1.0 if c == d else 2.0)
                                for (c, d), r in zip(prs, recs)])
            oa = np.concatenate([np.full(len(r[1]), b.shell_ao[c]) for (c, _), r in zip(prs, recs)])
            ob = np.concatenate([np.full(len(r[1]), b.shell_ao[d]) for (_, d), r in zip(prs, recs)])
            dev = lambda x: native.to_dev(np.ascontiguousarray(x))  # noqa: E731
            self.side[(l1, l2)] = dict(
                n=len(z), n_pairs=len(prs), P=dev(P), z=dev(z), s=dev(s), X1=dev(X1), X2=dev(X2),
                pair=dev(pair), w=dev(w), Iq=dev(record_bound(z, l1 + l2)),
                Ip=dev(record_bound(z, l1 + l2)), oa=dev(oa.astype(np.int32)),
                ob=dev(ob.astype(np.int32)), rows=dev(np.arange(len(z), dtype=np.int32)),
                didx=dev(_block_index(b, prs, l1, l2)))

    def run(self, D):
        import torch
        b = self.basis
        dev = native.device()
        Dd = _dev_density(D).view(-1)
        J = torch.zeros((b.n_ao, b.n_ao), dtype=torch.float64, device=dev)
        # the ket side: DQ per class, then the kets sorted by bound x Dmax
        # (descending, stable) on the device
        kets = {}
        for (lc, ld), t in self.side.items():
            nq = len(hermite_order(lc + ld))
            Dq = torch.zeros((t["n"], nq), dtype=torch.float64, device=dev)
            Dm = torch.zeros(t["n"], dtype=torch.float64, device=dev)
            Dcd = Dd[t["didx"]].contiguous()
            _call(f"fj_dq_{lc}_{ld}", C.c_int(64), C.c_int(t["n"]), C.c_int(t["n_pairs"]),
                  t["z"], t["X1"], t["X2"], t["s"], t["w"], t["pair"], Dcd, Dq, Dm)
            order = torch.argsort(-(t["Iq"] * Dm), stable=True)
            kets[(lc, ld)] = dict(Q=t["P"][order].contiguous(), eta=t["z"][order].contiguous(),
                                  Iq=t["Iq"][order].contiguous(), Dmax=Dm[order].contiguous(),
                                  D=Dq[order].contiguous(), n=t["n"])
        # the bra side: [p|q] per ket class, then JAB per buffer
        for (la, lb), t in self.side.items():
            lp = la + lb
            npp = len(hermite_order(lp))
            for (lc, ld), k in kets.items():
                lq = lc + ld
                Jp = torch.zeros((t["n"], npp), dtype=torch.float64, device=dev)
                kn = "pq32" if self.mixed else "pq"
                _call(f"fj_{kn}_L{lp + lq}_p{lp}_q{lq}", C.c_int(256), C.c_int(t["n"]), C.c_int(k["n"]),
                      C.c_double(self.tau), _c(t["P"]), _c(t["z"]), _c(t["Ip"]), _c(k["Q"]),
                      _c(k["eta"]), _c(k["Iq"]), _c(k["Dmax"]), _c(k["D"]), _c(Jp))
                _call(f"fj_jab_{la}{lb}", C.c_int(64), C.c_int(t["n"]), C.c_int(b.n_ao),
                      C.c_int(t["n"]), _c(t["z"]), _c(t["X1"]), _c(t["X2"]), _c(t["s"]),
                      _c(t["oa"]), _c(t["ob"]), _c(t["rows"]), _c(Jp), _c(J))
        return J


def _exchange_side(basis: Basis, by_l: dict, l1: int, l2: int, fixed: int | None = None):
    """Records of every ordered pair (s1, s2), l(s1) = l1, l(s2) = l2 (s1 ==
    fixed if given), with each record's s1 index, s2 index and s2's position
    among the shells of momentum l2."""
    firsts = [fixed] if fixed is not None else by_l.get(l1, [])
    recs, s1, s2, pos2 = [], [], [], []
    for u in firsts:
        for j, v in enumerate(by_l.get(l2, [])):
            r = _records(basis, u, v)
            recs.append(r)
            n = len(r[1])
            s1.append(np.full(n, u))
            s2.append(np.full(n, v))
            pos2.append(np.full(n, j))
    if not recs:
        return None
    cat = lambda k: np.concatenate([r[k] for r in recs])  # noqa: E731
    return dict(P=cat(0), z=cat(1), s=cat(2), X1=cat(3), X2=cat(4),
                s1=np.concatenate(s1), s2=np.concatenate(s2), pos2=np.concatenate(pos2))


def exchange_tables(basis: Basis) -> dict:
    """The K kernel's host tables (numpy): bra records per (la, lc) with
    their D-block row per ld, ket records per (lb, ld, b shell) sorted by
    bound, and the D-block gather indices per (lc, ld) (upstream
    eri/k_exchange/launch_k.py:308-401 builds the same records per launch)."""
    b = basis
    by_l = {}
    for s in range(b.n_sh):
        by_l.setdefault(int(b.shell_l[s]), []).append(s)
    lmax = max(by_l)
    if lmax > K_LMAX:
        raise ValueError(f"shells up to l = {lmax}: the K kernels are built through l = {K_LMAX}")
    L = sorted(by_l)
    bras, kets, dblk = {}, {}, {}
    for la in L:
        for lc in L:
            bra = _exchange_side(b, by_l, la, lc)
            if bra is None:
                continue
            t = dict(n=len(bra["z"]), P=bra["P"], z=bra["z"], s=bra["s"], XA=bra["X1"],
                     XC=bra["X2"], Ip=record_bound(bra["z"], la + lc),
                     oa=b.shell_ao[bra["s1"]].astype(np.int32))
            for ld in L:
                t[("dc", ld)] = (np.searchsorted(by_l[lc], bra["s2"]) * len(by_l[ld])).astype(np.int32)
            bras[(la, lc)] = t
    for lc in L:
        for ld in L:
            prs = [(c, d) for c in by_l[lc] for d in by_l[ld]]
            dblk[(lc, ld)] = dict(n=len(prs), idx=_block_index(b, prs, lc, ld))
    for lb in L:
        for ld in L:
            for bs in by_l[lb]:
                ket = _exchange_side(b, by_l, lb, ld, fixed=bs)
                if ket is None:
                    continue
                Iq = record_bound(ket["z"], lb + ld)
                order = np.argsort(-Iq, kind="stable")
                kets[(lb, ld, bs)] = dict(
                    n=len(ket["z"]), Q=ket["P"][order], eta=ket["z"][order], s=ket["s"][order],
                    XB=ket["X1"][order], XD=ket["X2"][order], Iq=Iq[order],
                    pos=ket["pos2"][order].astype(np.int32))
    return dict(by_l=by_l, bras=bras, kets=kets, dblk=dblk)


def exchange_launches(tables: dict):
    """(la, lc, lb, ld, b shell) of every K launch, in launch order: each
    writes the K rows of the (la, lc) bra records against the kets of one b
    shell, scattered into K's (a, b) block."""
    by_l = tables["by_l"]
    L = sorted(by_l)
    for (la, lc) in tables["bras"]:
        for ld in L:
            for lb in L:
                for bs in by_l[lb]:
                    if (lb, ld, bs) in tables["kets"]:
                        yield la, lc, lb, ld, bs


class ExchangeBuild:
    """K_ab = sum_cd D_cd (ac|bd) through the K kernel; `run(D)` returns K
    on the device. The host tables (`exchange_tables`) are built and copied
    to the device once here; `run` gathers D's blocks on the device and
    launches. mixed: the K mixed lever (upstream k_exchange/levers/mixed.py,
    `kmixed`): the straight-line classes' Hermite towers in FP32
    (`fj_k32_*`), everything that accumulates in FP64."""

    def __init__(self, basis: Basis, tau: float = 0.0, mixed: bool = False):
        self.basis = basis
        self.tau = float(tau)
        self.mixed = bool(mixed)
        host = exchange_tables(basis)
        self.by_l = host["by_l"]
        self.lmax = max(self.by_l)
        self.order = list(exchange_launches(host))
        dev = lambda x: native.to_dev(np.ascontiguousarray(x))  # noqa: E731
        up = lambda t: {k: (dev(v) if isinstance(v, np.ndarray) else v) for k, v in t.items()}  # noqa: E731
        self.bras = {k: up(t) for k, t in host["bras"].items()}
        self.kets = {k: up(t) for k, t in host["kets"].items()}
        self.dblk = {k: up(t) for k, t in host["dblk"].items()}

    def run(self, D, n_streams: int = 1):
        """K on the device. Each launch's K kernel writes its bra rows into a
        scratch and the scatter adds them into K with atomics, so with
        n_streams > 1 the launches (kernel and scatter together) go
        round-robin to that many side streams, each with its own scratch,
        forked from and joined to the current stream."""
        import torch
        b = self.basis
        dev = native.device()
        Dd = _dev_density(D)
        K = torch.zeros((b.n_ao, b.n_ao), dtype=torch.float64, device=dev)
        maxD = float(Dd.abs().max().item())
        Dflat = Dd.view(-1)
        Dcd = {k: Dflat[t["idx"]].contiguous() for k, t in self.dblk.items()}
        # one scratch per stream for every launch's rows: the K kernel
        # writes each bra row whole, and the scatter reads it before the
        # next launch on the same stream writes it again
        n_rows = max(self.bras[(la, lc)]["n"] * n_cart(la) * n_cart(lb)
                     for la, lc, lb, _, _ in self.order)
        main = torch.cuda.current_stream()
        streams = [main] if n_streams <= 1 else [torch.cuda.Stream() for _ in range(n_streams)]
        scratch = [torch.empty(n_rows, dtype=torch.float64, device=dev) for _ in streams]
        if n_streams > 1:
            for st in streams:
                st.wait_stream(main)
        for i, (la, lc, lb, ld, bs) in enumerate(self.order):
            t, k = self.bras[(la, lc)], self.kets[(lb, ld, bs)]
            with torch.cuda.stream(streams[i % len(streams)]):
                self._launch(la, lc, lb, ld, bs, t, k, maxD, Dcd,
                             scratch[i % len(streams)][:t["n"] * n_cart(la) * n_cart(lb)], K)
        if n_streams > 1:
            # the current stream waits for every side stream, so the
            # tensors allocated on it are not reused before they are done
            for st in streams:
                main.wait_stream(st)
        return K

    def _launch(self, la, lc, lb, ld, bs, t, k, maxD, Dcd, Kr, K):
        """One launch (the K kernel into Kr, then its scatter into K) on the
        current stream."""
        b = self.basis
        # the K mixed lever: the FP32 tower on the straight-line classes;
        # the f classes' loop-form kernel stays FP64
        kn = "k32" if self.mixed and max(la, lc, lb, ld) <= K_STRAIGHT_LMAX else "k"
        _call(f"fj_{kn}_{la}{lc}{lb}{ld}", C.c_int(256), C.c_int(t["n"]), C.c_int(k["n"]),
              C.c_int(self.dblk[(lc, ld)]["n"]), C.c_double(self.tau),
              C.c_double(maxD), _c(t["P"]), _c(t["z"]), _c(t["XA"]), _c(t["XC"]),
              _c(t["s"]), _c(t["Ip"]), _c(t[("dc", ld)]), _c(k["Q"]), _c(k["eta"]),
              _c(k["XB"]), _c(k["XD"]), _c(k["s"]), _c(k["Iq"]), _c(k["pos"]),
              _c(Dcd[(lc, ld)]), _c(Kr))
        _call("fj_k_scatter", C.c_int(t["n"]), C.c_int(n_cart(la)),
              C.c_int(n_cart(lb)), C.c_int(b.n_ao), _c(t["oa"]),
              C.c_int(int(b.shell_ao[bs])), _c(Kr), _c(K))


def fock_jk(basis: Basis, D: np.ndarray, tau: float = 0.0, mixed: bool = False):
    """G = J - K/2 through the MMD families (device tensors J, K, G); mixed:
    both levers (the FP32 Hermite towers of [p|q] and K)."""
    J = CoulombChain(basis, tau, mixed=mixed).run(D)
    K = ExchangeBuild(basis, tau, mixed=mixed).run(D)
    return J, K, J - 0.5 * K
