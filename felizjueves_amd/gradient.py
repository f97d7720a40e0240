"""The fused two-electron nuclear gradient on the GPU: dE_2e/dR for
E_2e = 1/2 tr(D (J - K/2)), RHF's Coulomb and exchange gradient terms in one
kernel per F class (upstream `gradient/kernel_gf.py`, its stream
`gradient/interpreted.py:293-351`, `gf_launches` / `gf_run` /
`gf_gradient`), on the library's loop-form kernel (`fj_gf_loop` and its
per-class entries `fj_gf_<la><lb><lc><ld>`).

    gf_launches(basis, shell_atom)            -> GfStream    the stream
    gf_gradient(basis, shell_atom, n_atom, D) -> G [n_atom, 3]
    GpuGfRunner                                the device side, reusable
                                               (per class one merged launch,
                                               or the per-launch entries)

THE STREAM (`interpreted.py:293-310`): every unordered shell pair written
canonically (la <= lb, the lower index first on a tie), batched by (la, lb,
contraction degree), each batch pair I <= J one launch of class
(la_I lb_I | la_J lb_J) with diag = 1 exactly on I == J, so every unique
quartet is met once; nothing is screened or pruned (the upstream's gf_ has
no screen). A side's record per primitive pair (i outer) is the fused F
record and the two exponents (`kernel_gf.py:168`, `GREC`): P, X = P - A,
zeta, U = c_a c_b (pi/zeta)^{3/2} exp(-alpha beta/zeta |AB|^2), alpha, beta.

The density must be symmetric (the derived weights assume D_ab = D_ba,
`interpreted.py:345-347`). Shells through d (the upstream emits gf_ through
l = 2).
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import native
from .basis import Basis

GF_LMAX = 2
GREC = 10
ITEMS = 2048   # work items one merged class launch aims for


@dataclass
class GfSide:
    """One packed side: prim [n, degree, 10], sep [n, 3], sh / ao / at [n, 2]."""

    la: int
    lb: int
    degree: int
    prim: np.ndarray
    sep: np.ndarray
    sh: np.ndarray
    ao: np.ndarray
    at: np.ndarray

    @property
    def n(self) -> int:
        return len(self.sh)


@dataclass
class GfLaunch:
    cls: tuple
    diag: int
    bra: int      # index into the stream's sides
    ket: int


@dataclass
class GfStream:
    sides: list
    launches: list


def _records(basis: Basis, a: int, b: int) -> np.ndarray:
    """The (a, b) pair's GREC rows, every primitive pair, a's primitive outer."""
    A, B = basis.shell_centre[a], basis.shell_centre[b]
    pa, pb = int(basis.shell_pstart[a]), int(basis.shell_pstart[b])
    na, nb = int(basis.shell_nprim[a]), int(basis.shell_nprim[b])
    al = np.repeat(basis.exps[pa:pa + na], nb)
    be = np.tile(basis.exps[pb:pb + nb], na)
    ca = np.repeat(basis.coefs[pa:pa + na], nb)
    cb = np.tile(basis.coefs[pb:pb + nb], na)
    z = al + be
    P = (al[:, None] * A + be[:, None] * B) / z[:, None]
    ab2 = float(((A - B) ** 2).sum())
    U = ca * cb * (math.pi / z) ** 1.5 * np.exp(-al * be / z * ab2)
    return np.concatenate([P, P - A, z[:, None], U[:, None], al[:, None], be[:, None]], axis=1)


def gf_launches(basis: Basis, shell_atom) -> GfStream:
    """The gradient's stream over `basis` (module docstring)."""
    shell_atom = np.asarray(shell_atom, dtype=np.int32)
    if len(shell_atom) != basis.n_sh:
        raise ValueError(f"shell_atom has {len(shell_atom)} entries for {basis.n_sh} shells")
    if int(basis.shell_l.max()) > GF_LMAX:
        raise ValueError(f"the gradient is built through l = {GF_LMAX}")
    L = basis.shell_l
    batches: dict = {}
    for i in range(basis.n_sh):
        for j in range(i, basis.n_sh):
            a, b = (i, j) if (L[i], i) <= (L[j], j) else (j, i)
            key = (int(L[a]), int(L[b]), int(basis.shell_nprim[a] * basis.shell_nprim[b]))
            batches.setdefault(key, []).append((a, b))
    keys = sorted(batches)
    sides = []
    for la, lb, deg in keys:
        prs = batches[(la, lb, deg)]
        sides.append(GfSide(
            la, lb, deg,
            np.ascontiguousarray(np.stack([_records(basis, a, b) for a, b in prs])),
            np.ascontiguousarray(np.array([basis.shell_centre[a] - basis.shell_centre[b] for a, b in prs])),
            np.array(prs, dtype=np.int32),
            np.array([(basis.shell_ao[a], basis.shell_ao[b]) for a, b in prs], dtype=np.int32),
            np.array([(shell_atom[a], shell_atom[b]) for a, b in prs], dtype=np.int32)))
    launches = [GfLaunch((sides[i].la, sides[i].lb, sides[j].la, sides[j].lb), int(i == j), i, j)
                for i in range(len(sides)) for j in range(i, len(sides))]
    return GfStream(sides, launches)


def default_threads(cls) -> int:
    """Workgroup size of a class: 256 lanes (four wavefronts) for every
    class. Measured per class on the three-atom s,s,p,d system and the
    10-atom cluster at 64, 128 and 256 lanes: 256 was fastest or within 15%
    of the fastest everywhere (the register-tiled stages have few tiles per
    primitive quartet, so more lanes take more primitive quartets per
    batch)."""
    return 256


def quartets(stream: GfStream) -> int:
    """Unique shell quartets the stream visits."""
    n = 0
    for L in stream.launches:
        nb, nk = stream.sides[L.bra].n, stream.sides[L.ket].n
        n += nb * (nb + 1) // 2 if L.diag else nb * nk
    return n


def _box(l1: int, l2: int) -> tuple[int, int]:
    """(valid (ab, t) elements of the derivative loop, Hermite t with
    |t| <= l1 + l2): t inside the box t_k <= a_k + b_k, +1 on at most one
    axis, over every Cartesian pair (a, b)."""
    from .basis import cartesians
    herm = [(x, y, n - x - y) for n in range(l1 + l2 + 2) for x in range(n, -1, -1)
            for y in range(n - x, -1, -1)]
    n = 0
    for a in cartesians(l1):
        for b in cartesians(l2):
            for t in herm:
                o = [t[k] - a[k] - b[k] for k in range(3)]
                if max(o) <= 1 and sum(1 for v in o if v == 1) <= 1:
                    n += 1
    return n, (l1 + l2 + 1) * (l1 + l2 + 2) * (l1 + l2 + 3) // 6


def flops_per_prim_quartet(cls) -> int:
    """The arithmetic of one primitive quartet in the loop form's
    formulation: the Hermite-space weights 2 nab ncd (nu + nt) (the E
    products counted once per (cd, u) / (ab, t)), the contraction with the
    tower 2 nu per valid bra element and 2 nt per valid ket element, 13 per
    valid element for the differentiated E and the six sums, and the tower
    (3 per entry per level). No FMA credit beyond 2."""
    la, lb, lc, ld = cls
    nab = (la + 1) * (la + 2) // 2 * ((lb + 1) * (lb + 2) // 2)
    ncd = (lc + 1) * (lc + 2) // 2 * ((ld + 1) * (ld + 2) // 2)
    vb, nt = _box(la, lb)
    vk, nu = _box(lc, ld)
    Lt = sum(cls) + 1
    herm = lambda n: n * (n + 1) * (n + 2) // 6  # noqa: E731
    tower = sum(3 * herm(Lt - m + 1) for m in range(Lt))
    return (2 * nab * ncd * (nu + nt) + 2 * ncd * nu * 2 + 2 * nab * nt * 2
            + vb * (2 * nu + 13) + vk * (2 * nt + 13) + tower)


def flops(stream: GfStream) -> int:
    """The stream's arithmetic (flops_per_prim_quartet over every primitive
    quartet of every launch)."""
    tot = 0
    cache: dict = {}
    for L in stream.launches:
        b, k = stream.sides[L.bra], stream.sides[L.ket]
        nq = b.n * (b.n + 1) // 2 if L.diag else b.n * k.n
        if L.cls not in cache:
            cache[L.cls] = flops_per_prim_quartet(L.cls)
        tot += nq * b.degree * k.degree * cache[L.cls]
    return tot


def prim_quartets(stream: GfStream) -> int:
    n = 0
    for L in stream.launches:
        b, k = stream.sides[L.bra], stream.sides[L.ket]
        n += (b.n * (b.n + 1) // 2 if L.diag else b.n * k.n) * b.degree * k.degree
    return n


def _p(t):
    return C.c_void_p(native.ptr(t))


JOB_DTYPE = np.dtype([("bra_row", "<i4"), ("ket_row", "<i4"), ("n_bra", "<i4"), ("n_ket", "<i4"),
                      ("k_ab", "<i4"), ("k_cd", "<i4"), ("diag", "<i4"), ("pad", "<i4"),
                      ("bra_prim", "<i8"), ("ket_prim", "<i8"), ("q0", "<i8")])   # FjGfJob


class GpuGfRunner:
    """The stream's sides on the device, reused across densities; `run`
    launches every class into a zeroed G on `stream`.

    mode "jobs" (default): every launch of a class in one kernel launch
    (fj_gf_jobs over a job table: the stream's launches are many and small,
    one per pair of contraction degrees, and one of them alone leaves the
    chip idle); mode "launches": one fj_gf_<cls> call per launch of the
    stream, the upstream gf_ parameter list."""

    def __init__(self, stream_: GfStream, n_atom: int, n_ao: int, threads: dict | None = None,
                 mode: str = "jobs"):
        import torch
        if mode not in ("jobs", "launches"):
            raise ValueError(f"mode {mode!r} is not 'jobs' or 'launches'")
        self.s, self.n_atom, self.n_ao, self.mode = stream_, int(n_atom), int(n_ao), mode
        self.threads = threads or {}
        self.G = torch.zeros((self.n_atom, 3), dtype=torch.float64, device=native.device())
        lib = native.lib()
        self.fns = {}
        if mode == "launches":
            self.dev = [{k: native.to_dev(getattr(sd, k)) for k in ("prim", "sep", "sh", "ao", "at")}
                        for sd in stream_.sides]
            for L in stream_.launches:
                if L.cls not in self.fns:
                    fn = getattr(lib, "fj_gf_" + "".join(map(str, L.cls)))
                    fn.restype = C.c_int
                    self.fns[L.cls] = fn
            return
        # the sides concatenated: rows and record offsets per side
        rows = np.concatenate([[0], np.cumsum([sd.n for sd in stream_.sides])]).astype(np.int64)
        recs = np.concatenate([[0], np.cumsum([sd.prim.size for sd in stream_.sides])]).astype(np.int64)
        cat = lambda k: np.concatenate([getattr(sd, k).reshape(sd.n, -1) for sd in stream_.sides])  # noqa: E731
        self.cat = {"prim": native.to_dev(np.concatenate([sd.prim.ravel() for sd in stream_.sides])),
                    **{k: native.to_dev(np.ascontiguousarray(cat(k))) for k in ("sep", "sh", "ao", "at")}}
        by: dict = {}
        for L in stream_.launches:
            by.setdefault(L.cls, []).append(L)
        self.classes = []
        lib.fj_gf_batch.restype = C.c_int
        for cls, Ls in sorted(by.items()):
            jobs = np.zeros(len(Ls), dtype=JOB_DTYPE)
            nq = [stream_.sides[L.bra].n * (stream_.sides[L.bra].n + 1) // 2 if L.diag
                  else stream_.sides[L.bra].n * stream_.sides[L.ket].n for L in Ls]
            max_pq = max(stream_.sides[L.bra].degree * stream_.sides[L.ket].degree for L in Ls)
            th = self.threads.get(cls, default_threads(cls))
            B = int(lib.fj_gf_batch(*[C.c_int(x) for x in cls], C.c_int(max_pq), C.c_int(th)))
            # a class of few quartets splits each quartet's primitive quartets
            # into pieces (multiples of B) so that about ITEMS items fill the chip
            pieces = max(1, -(-ITEMS // sum(nq)))
            q0 = 0
            for j, L in enumerate(Ls):
                b, k = stream_.sides[L.bra], stream_.sides[L.ket]
                n_pq = b.degree * k.degree
                chunk = n_pq if pieces == 1 else min(n_pq, -(-(-(-n_pq // pieces)) // B) * B)
                jobs[j] = (rows[L.bra], rows[L.ket], b.n, k.n, b.degree, k.degree, L.diag, chunk,
                           recs[L.bra], recs[L.ket], q0)
                q0 += nq[j] * -(-n_pq // chunk)
            jd = native.to_dev(np.frombuffer(jobs.tobytes(), dtype=np.uint8))
            self.classes.append((cls, jd, len(Ls), q0, max_pq))
        self.fj_gf_jobs = lib.fj_gf_jobs
        self.fj_gf_jobs.restype = C.c_int

    def run(self, D_dev, stream=None, n_streams: int = 1):
        """G [n_atom, 3] on the device (float64), zeroed then accumulated.
        The job-table launches add into G with atomics and are independent,
        so with n_streams > 1 (job mode) they go round-robin, costliest
        class first, to that many side streams forked from and joined to
        `stream` (the current stream if None)."""
        import torch
        self.G.zero_()
        main = stream if stream is not None else torch.cuda.current_stream()
        h = C.c_void_p(native.stream_handle(main))
        if self.mode == "jobs":
            c = self.cat
            order = self.classes
            if n_streams > 1:
                if len(getattr(self, "_streams", ())) != n_streams:
                    self._streams = [torch.cuda.Stream() for _ in range(n_streams)]
                for st in self._streams:
                    st.wait_stream(main)
                order = sorted(self.classes, key=lambda r: -r[3] * flops_per_prim_quartet(r[0]))
            for i, (cls, jd, n_jobs, n_q, max_pq) in enumerate(order):
                th = self.threads.get(cls, default_threads(cls))
                hs = C.c_void_p(native.stream_handle(self._streams[i % n_streams])) if n_streams > 1 else h
                native.check(self.fj_gf_jobs(
                    hs, *[C.c_int(x) for x in cls], C.c_int(th), C.c_int(n_jobs), _p(jd),
                    C.c_longlong(n_q), C.c_int(max_pq), C.c_int(self.n_atom), C.c_int(self.n_ao),
                    _p(c["prim"]), _p(c["sep"]), _p(c["sh"]), _p(c["ao"]), _p(c["at"]), _p(D_dev),
                    _p(self.G)), "fj_gf_jobs")
            if n_streams > 1:
                for st in self._streams:
                    main.wait_stream(st)
            return self.G
        for L in self.s.launches:
            b, k = self.s.sides[L.bra], self.s.sides[L.ket]
            db, dk = self.dev[L.bra], self.dev[L.ket]
            th = self.threads.get(L.cls, default_threads(L.cls))
            native.check(self.fns[L.cls](
                h, C.c_int(th), C.c_int(b.n), C.c_int(k.n), C.c_int(b.degree), C.c_int(k.degree),
                C.c_int(self.n_atom), C.c_int(self.n_ao), C.c_int(L.diag),
                _p(db["prim"]), _p(db["sep"]), _p(db["sh"]), _p(db["ao"]), _p(db["at"]),
                _p(dk["prim"]), _p(dk["sep"]), _p(dk["sh"]), _p(dk["ao"]), _p(dk["at"]),
                _p(D_dev), _p(self.G)), "fj_gf_" + "".join(map(str, L.cls)))
        return self.G


def gf_gradient(basis: Basis, shell_atom, n_atom: int, D, *, runner: GpuGfRunner | None = None,
                stream=None) -> np.ndarray:
    """dE_2e/dR [n_atom, 3] from every launch of the stream into one zeroed G."""
    import torch
    Dx = np.asarray(D.cpu().numpy() if isinstance(D, torch.Tensor) else D, dtype=np.float64)
    if Dx.shape != (basis.n_ao, basis.n_ao) or not np.array_equal(Dx, Dx.T):
        raise ValueError(f"D is not a symmetric {basis.n_ao} x {basis.n_ao} matrix "
                         "(the derived weights assume D_ab = D_ba)")
    if runner is None:
        runner = GpuGfRunner(gf_launches(basis, shell_atom), n_atom, basis.n_ao)
    D_dev = D if isinstance(D, torch.Tensor) else native.to_dev(np.ascontiguousarray(Dx))
    G = runner.run(D_dev, stream)
    torch.cuda.synchronize()
    return G.cpu().numpy()


def shell_atoms(basis: Basis, coords) -> np.ndarray:
    """The atom each shell sits on: the atom whose coordinates equal the
    shell's centre (refuses a shell on no atom)."""
    coords = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    out = np.empty(basis.n_sh, dtype=np.int32)
    for s in range(basis.n_sh):
        d = np.abs(coords - basis.shell_centre[s]).max(axis=1)
        k = int(np.argmin(d))
        if d[k] > 1e-12:
            raise ValueError(f"shell {s} sits on no atom")
        out[s] = k
    return out
