"""The fused two-electron nuclear gradient on the GPU: dE_2e/dR for
E_2e = 1/2 tr(D (J - K/2)), RHF's Coulomb and exchange gradient terms in one
kernel per F class (upstream `gradient/kernel_gf.py`, its stream
`gradient/interpreted.py:293-351`, `gf_launches` / `gf_run` /
`gf_gradient`), on the library's loop-form kernel (`fj_gf_loop` and its
per-class entries `fj_gf_<la><lb><lc><ld>`).

    gf_launches(basis, shell_atom)            -> [GfLaunch]  the stream
    gf_gradient(basis, shell_atom, n_atom, D) -> G [n_atom, 3]
    GpuGfRunner                                the device side, reusable

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
    """Workgroup size of a class: one wavefront for the small classes, more
    lanes where the (ab, t) and (cd, u) loops are long."""
    la, lb, lc, ld = cls
    herm = lambda n: n * (n + 1) * (n + 2) // 6  # noqa: E731
    nab = (la + 1) * (la + 2) // 2 * ((lb + 1) * (lb + 2) // 2)
    ncd = (lc + 1) * (lc + 2) // 2 * ((ld + 1) * (ld + 2) // 2)
    n = max(nab * herm(la + lb + 2), ncd * herm(lc + ld + 2))
    return 64 if n <= 256 else (128 if n <= 1024 else 256)


def quartets(stream: GfStream) -> int:
    """Unique shell quartets the stream visits."""
    n = 0
    for L in stream.launches:
        nb, nk = stream.sides[L.bra].n, stream.sides[L.ket].n
        n += nb * (nb + 1) // 2 if L.diag else nb * nk
    return n


def _p(t):
    return C.c_void_p(native.ptr(t))


class GpuGfRunner:
    """The stream's sides on the device, reused across densities; `run`
    launches every class into a zeroed G on `stream`."""

    def __init__(self, stream_: GfStream, n_atom: int, n_ao: int, threads: dict | None = None):
        import torch
        self.s, self.n_atom, self.n_ao = stream_, int(n_atom), int(n_ao)
        self.threads = threads or {}
        self.dev = [{k: native.to_dev(getattr(sd, k)) for k in ("prim", "sep", "sh", "ao", "at")}
                    for sd in stream_.sides]
        self.G = torch.zeros((self.n_atom, 3), dtype=torch.float64, device=native.device())
        lib = native.lib()
        self.fns = {}
        for L in stream_.launches:
            if L.cls not in self.fns:
                fn = getattr(lib, "fj_gf_" + "".join(map(str, L.cls)))
                fn.restype = C.c_int
                self.fns[L.cls] = fn

    def run(self, D_dev, stream=None):
        """G [n_atom, 3] on the device (float64), zeroed then accumulated."""
        self.G.zero_()
        h = C.c_void_p(native.stream_handle(stream))
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
