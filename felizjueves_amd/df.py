"""Density fitting on the GPU: the two-center Coulomb metric (P|Q) of an
auxiliary basis (upstream `eri/density_fit/metric.py`, its kernel
`kernel_metric.py`) and the three-center integral (mn|P) (upstream
`eri/density_fit/eri3.py`, `kernel_eri3.py`), on the library's loop-form
kernel (`fj_metric_loop` / `fj_eri3_loop` and their per-class entries).

    AuxBasis.from_arrays(...)              an auxiliary basis (shells through i)
    metric_classes(aux) / metric_matrix(aux)          -> (P|Q) [n_aux, n_aux]
    eri3_classes(basis, aux) / eri3(basis, aux)       -> (mn|P) [n_ao, n_ao, n_aux]
    flops_metric(aux) / flops_eri3(basis, aux)        -> the loop form's arithmetic
    DfIntegrals(aux, basis).metric() / .eri3()         the tables staged once

An aux primitive is a Hermite "pair" of one Gaussian: record (A, alpha, s =
c (pi/alpha)^{3/2}) (upstream `metric.aux_dress`, the one place the dressing
goes); a bra record of (mn|P) is a primitive pair (zeta, P, X_PA, X_PB, s =
c_m c_n (pi/zeta)^{3/2}), its Gaussian factor exp(-alpha beta/zeta |AB|^2)
carried by the E seed. The staging follows the upstream's: every aux shell
pair P <= Q (by index) joins the class (lP, lQ) with lP <= lQ, staged
swapped when P has the higher l (the mirrored store writes the same two
elements); every AO shell pair m <= n joins (lm, ln) likewise, with every
aux shell. Blocks are written whole, mirrored; nothing accumulates, so the
outputs need no clearing.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import native
from . import pairs as P
from .basis import Basis, n_cart

AUX_LMAX = 6


class AuxBasis:
    """Auxiliary shells (Cartesian, contracted, l <= 6) as flat arrays, the
    AO basis's conventions (coefficients multiply unnormalised primitives)."""

    def __init__(self, shell_l, shell_centre, shell_nprim, exps, coefs):
        self.shell_l = np.asarray(shell_l, dtype=np.int32)
        if self.shell_l.size == 0:
            raise ValueError("an auxiliary basis needs at least one shell")
        if self.shell_l.min() < 0 or self.shell_l.max() > AUX_LMAX:
            raise ValueError(f"aux shells are built through l = {AUX_LMAX}")
        self.shell_centre = np.ascontiguousarray(shell_centre, dtype=np.float64).reshape(-1, 3)
        self.shell_nprim = np.asarray(shell_nprim, dtype=np.int32)
        self.exps = np.asarray(exps, dtype=np.float64)
        self.coefs = np.asarray(coefs, dtype=np.float64)
        if not np.all(self.exps > 0) or len(self.exps) != int(self.shell_nprim.sum()):
            raise ValueError("aux exponents must be > 0, shell_nprim of them per shell")
        self.n_sh = len(self.shell_l)
        self.shell_pstart = np.concatenate([[0], np.cumsum(self.shell_nprim)[:-1]]).astype(np.int32)
        self.shell_size = np.array([n_cart(int(l)) for l in self.shell_l], dtype=np.int32)  # noqa: E741
        self.shell_ao = np.concatenate([[0], np.cumsum(self.shell_size)[:-1]]).astype(np.int32)
        self.n_aux = int(self.shell_size.sum())

    @classmethod
    def from_npz(cls, z, prefix: str = "aux_") -> "AuxBasis":
        return cls(z[prefix + "shell_l"], z[prefix + "shell_centre"], z[prefix + "shell_nprim"],
                   z[prefix + "exps"], z[prefix + "coefs"])

    def records(self):
        """Per primitive: A [n, 3], alpha [n], s = c (pi/alpha)^{3/2} [n]
        (upstream metric.aux_dress), and per shell its record range."""
        A = np.repeat(self.shell_centre, self.shell_nprim, axis=0)
        s = self.coefs * (math.pi / self.exps) ** 1.5
        rec0 = self.shell_pstart
        return A, self.exps.copy(), s, rec0, rec0 + self.shell_nprim


@dataclass
class DfClass:
    cls: tuple
    n: int
    arrays: dict          # host int32 arrays of the class's pairs / cells


def metric_classes(aux: AuxBasis) -> dict:
    """{(lP, lQ): DfClass} of every aux shell pair P <= Q, staged with the
    lower l first (swapped where P has the higher l)."""
    _, _, _, rec0, rec1 = aux.records()
    by: dict = {}
    for p in range(aux.n_sh):
        for q in range(p, aux.n_sh):
            a, b = (p, q) if aux.shell_l[p] <= aux.shell_l[q] else (q, p)
            by.setdefault((int(aux.shell_l[a]), int(aux.shell_l[b])), []).append((a, b))
    out = {}
    for c, prs in sorted(by.items()):
        a = np.array([x for x, _ in prs])
        b = np.array([y for _, y in prs])
        out[c] = DfClass(c, len(prs), {
            "p_rec0": rec0[a].astype(np.int32), "p_rec1": rec1[a].astype(np.int32),
            "q_rec0": rec0[b].astype(np.int32), "q_rec1": rec1[b].astype(np.int32),
            "off_p": aux.shell_ao[a].astype(np.int32), "off_q": aux.shell_ao[b].astype(np.int32)})
    return out


def eri3_classes(basis: Basis, aux: AuxBasis) -> tuple[dict, dict]:
    """({(la, lb, lP): DfClass} of every cell (AO shell pair m <= n staged
    with the lower l first, aux shell P), bra record arrays): one bra record
    table over every staged AO pair (its primitive pairs contiguous)."""
    pairs = []
    for m in range(basis.n_sh):
        for n in range(m, basis.n_sh):
            a, b = (m, n) if basis.shell_l[m] <= basis.shell_l[n] else (n, m)
            pairs.append((a, b))
    by_pc: dict = {}
    for a, b in pairs:
        by_pc.setdefault((int(basis.shell_l[a]), int(basis.shell_l[b])), []).append((a, b))
    z, Pc, XA, XB, s, r0, r1, pa, pb = [], [], [], [], [], [], [], [], []
    n_rec = 0
    for c, prs in sorted(by_pc.items()):
        pp = P.primitive_pairs(basis, prs)
        ca = basis.coefs[basis.shell_pstart[np.array(prs)[pp.pair, 0]] + pp.i]
        cb = basis.coefs[basis.shell_pstart[np.array(prs)[pp.pair, 1]] + pp.j]
        at = np.searchsorted(pp.pair, np.arange(len(prs) + 1))
        z.append(pp.zeta)
        Pc.append(pp.P)
        XA.append(pp.PA)
        XB.append(pp.P - basis.shell_centre[np.array(prs)[pp.pair, 1]])
        s.append(ca * cb * (math.pi / pp.zeta) ** 1.5)
        for k, (a, b) in enumerate(prs):
            r0.append(n_rec + at[k])
            r1.append(n_rec + at[k + 1])
            pa.append(a)
            pb.append(b)
        n_rec += len(pp.zeta)
    bra = {"zeta_ab": np.concatenate(z), "P": np.concatenate(Pc), "XPA": np.concatenate(XA),
           "XPB": np.concatenate(XB), "s_ab": np.concatenate(s)}
    _, _, _, q0, q1 = aux.records()
    by: dict = {}
    for k, (a, b) in enumerate(zip(pa, pb)):
        for p in range(aux.n_sh):
            by.setdefault((int(basis.shell_l[a]), int(basis.shell_l[b]), int(aux.shell_l[p])),
                          []).append((k, p))
    out = {}
    for c, cells in sorted(by.items()):
        k = np.array([x for x, _ in cells])
        p = np.array([y for _, y in cells])
        out[c] = DfClass(c, len(cells), {
            "ab_rec0": np.array(r0, dtype=np.int32)[k], "ab_rec1": np.array(r1, dtype=np.int32)[k],
            "p_rec0": q0[p].astype(np.int32), "p_rec1": q1[p].astype(np.int32),
            "off_a": basis.shell_ao[np.array(pa)[k]].astype(np.int32),
            "off_b": basis.shell_ao[np.array(pb)[k]].astype(np.int32),
            "off_p": aux.shell_ao[p].astype(np.int32)})
    return out, bra


def _cell_flops(la: int, lb: int, lc: int) -> int:
    """The loop-form kernel's arithmetic per (bra record, aux record) pair of
    a class (la, lb | lc): the tower (3 per entry per level, |t| <= L), and
    per output element ((a, b), c) for each Hermite index p of the bra box
    (p_d <= a_d + b_d) the aux parity box (q_d = c_d, c_d - 2, ...): 4 per q
    (the E product, the signed multiply-add with [p + q]), 4 per p (the bra E
    product and its multiply-add), 2 per element (the accumulation)."""
    from .basis import cartesians
    L = la + lb + lc
    herm = lambda n: n * (n + 1) * (n + 2) // 6  # noqa: E731
    f = sum(3 * herm(L - m + 1) for m in range(L))
    for a in cartesians(la):
        for b in cartesians(lb):
            nbox = (a[0] + b[0] + 1) * (a[1] + b[1] + 1) * (a[2] + b[2] + 1)
            for c in cartesians(lc):
                nq = (c[0] // 2 + 1) * (c[1] // 2 + 1) * (c[2] // 2 + 1)
                f += nbox * (4 * nq + 4) + 2
    return f


def flops_metric(aux: AuxBasis) -> int:
    """The arithmetic of one metric_matrix call (_cell_flops over every
    record pair of every aux shell pair)."""
    tot = 0
    for (lp, lq), dc in metric_classes(aux).items():
        a = dc.arrays
        n_rp = (a["p_rec1"] - a["p_rec0"]).astype(np.int64) * (a["q_rec1"] - a["q_rec0"])
        tot += int(n_rp.sum()) * _cell_flops(lp, 0, lq)
    return tot


def flops_eri3(basis: Basis, aux: AuxBasis) -> int:
    """The arithmetic of one eri3 call (_cell_flops over every record pair
    of every cell)."""
    classes, _ = eri3_classes(basis, aux)
    tot = 0
    for (la, lb, lp), dc in classes.items():
        a = dc.arrays
        n_rp = (a["ab_rec1"] - a["ab_rec0"]).astype(np.int64) * (a["p_rec1"] - a["p_rec0"])
        tot += int(n_rp.sum()) * _cell_flops(la, lb, lp)
    return tot


def _p(t):
    return C.c_void_p(native.ptr(t))


THREADS = 128   # lanes per workgroup of the DF kernel


class DfIntegrals:
    """The density-fitting tables of one (basis, aux) staged on the device
    once; `metric(out)` and `eri3(out)` launch every class into device
    arrays (the host work -- class tables, record arrays, uploads -- is paid
    here, not per call)."""

    def __init__(self, aux: AuxBasis, basis: Basis | None = None, threads: int = THREADS):
        self.aux, self.basis, self.threads = aux, basis, threads
        A, alpha, s, _, _ = aux.records()
        self.n_rec = len(alpha)
        self.d = {k: native.to_dev(v) for k, v in (("A", A), ("alpha", alpha), ("s", s))}
        lib = native.lib()
        self.fm = lib.fj_metric_loop_rp
        self.fm.restype = C.c_int
        self.fe = lib.fj_eri3_loop_rp
        self.fe.restype = C.c_int
        self.mcls = []
        for (lp, lq), dc in metric_classes(aux).items():
            a = dc.arrays
            max_rp = int(((a["p_rec1"] - a["p_rec0"]) * (a["q_rec1"] - a["q_rec0"])).max())
            self.mcls.append(((lp, lq), dc.n, max_rp, {k: native.to_dev(v) for k, v in a.items()}))
        self.ecls = []
        if basis is not None:
            classes, bra = eri3_classes(basis, aux)
            self.n_bra = len(bra["zeta_ab"])
            self.b = {k: native.to_dev(v) for k, v in bra.items()}
            for cls, dc in classes.items():
                a = dc.arrays
                max_rp = int(((a["ab_rec1"] - a["ab_rec0"]) * (a["p_rec1"] - a["p_rec0"])).max())
                self.ecls.append((cls, dc.n, max_rp, {k: native.to_dev(v) for k, v in a.items()}))

    def metric(self, out=None, stream=None):
        """(P|Q) [n_aux, n_aux] into `out` (a new device array if None)."""
        import torch
        aux, d = self.aux, self.d
        if out is None:
            out = torch.empty((aux.n_aux, aux.n_aux), dtype=torch.float64, device=native.device())
        h = C.c_void_p(native.stream_handle(stream))
        for (lp, lq), n, max_rp, ar in self.mcls:
            native.check(self.fm(h, C.c_int(lp), C.c_int(lq), C.c_int(self.threads), C.c_int(n),
                                 C.c_int(self.n_rec), C.c_int(aux.n_aux), _p(d["A"]), _p(d["alpha"]),
                                 _p(d["s"]), *[_p(ar[k]) for k in ("p_rec0", "p_rec1", "q_rec0", "q_rec1",
                                                                   "off_p", "off_q")],
                                 _p(out), C.c_int(max_rp)), "fj_metric_loop_rp")
        return out

    def eri3(self, out=None, stream=None, n_streams: int = 1):
        """(mn|P) [n_ao, n_ao, n_aux] into `out` (a new device array if None).
        Each class writes its own cells, so with n_streams > 1 the classes
        go round-robin (costliest first) to that many side streams, forked
        from and joined to `stream` (the current stream if None)."""
        import torch
        if self.basis is None:
            raise ValueError("no AO basis staged")
        aux, d, b = self.aux, self.d, self.b
        n_ao = self.basis.n_ao
        if out is None:
            out = torch.empty((n_ao, n_ao, aux.n_aux), dtype=torch.float64, device=native.device())
        main = stream if stream is not None else torch.cuda.current_stream()
        if n_streams > 1:
            if len(getattr(self, "_streams", ())) != n_streams:
                self._streams = [torch.cuda.Stream() for _ in range(n_streams)]
            for st in self._streams:
                st.wait_stream(main)
            order = sorted(self.ecls, key=lambda c: -c[1] * c[2] * (c[0][0] + c[0][1] + c[0][2] + 1) ** 4)
        else:
            order = self.ecls
        for i, ((la, lb, lp), n, max_rp, ar) in enumerate(order):
            h = C.c_void_p(native.stream_handle(self._streams[i % n_streams] if n_streams > 1 else main))
            native.check(self.fe(h, C.c_int(la), C.c_int(lb), C.c_int(lp), C.c_int(self.threads),
                                 C.c_int(n), C.c_int(self.n_bra), C.c_int(self.n_rec), C.c_int(n_ao),
                                 C.c_int(aux.n_aux),
                                 *[_p(b[k]) for k in ("zeta_ab", "P", "XPA", "XPB", "s_ab")],
                                 _p(d["A"]), _p(d["alpha"]), _p(d["s"]),
                                 *[_p(ar[k]) for k in ("ab_rec0", "ab_rec1", "p_rec0", "p_rec1", "off_a",
                                                       "off_b", "off_p")], _p(out), C.c_int(max_rp)),
                         "fj_eri3_loop_rp")
        if n_streams > 1:
            for st in self._streams:
                main.wait_stream(st)
        return out


def metric_matrix(aux: AuxBasis, stream=None, threads: int = THREADS):
    """(P|Q) of every aux function on the device, [n_aux, n_aux] (float64)."""
    import torch
    out = DfIntegrals(aux, threads=threads).metric(stream=stream)
    torch.cuda.synchronize()
    return out


def eri3(basis: Basis, aux: AuxBasis, stream=None, threads: int = THREADS):
    """(mn|P) on the device, [n_ao, n_ao, n_aux] (float64, aux fastest)."""
    import torch
    out = DfIntegrals(aux, basis, threads=threads).eri3(stream=stream)
    torch.cuda.synchronize()
    return out
