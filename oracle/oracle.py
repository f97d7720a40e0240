"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package never
imports this module.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build() -> Path:
    """Compile liboracle.so from fj_oracle.c (make)."""
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.oracle_boys.argtypes = [C.c_int, C.c_double, f64p]
        basis_args = [i32p, f64p, i32p, i32p, f64p, f64p]
        for fn in ("oracle_eri_mmd", "oracle_eri_hgp"):
            getattr(L, fn).argtypes = [C.c_int] * 4 + basis_args + [f64p]
        L.oracle_fock_ordered.argtypes = ([C.c_int] + basis_args +
                                          [i32p, C.c_int, f64p, f64p, f64p])
        L.oracle_schwarz.argtypes = [C.c_int] + basis_args + [i32p, C.c_int, f64p]
        L.oracle_grad_ordered.argtypes = ([C.c_int] + basis_args +
                                          [i32p, i32p, C.c_int, f64p, C.c_int, f64p])
        vp = C.c_void_p
        L.oracle_f_launch.argtypes = ([C.c_int] * 11 + [C.c_double] +
                                      [vp] * 13)
        L.oracle_fx_launch.argtypes = ([C.c_int] * 8 + [vp] * 2 + [C.c_int] * 3 +
                                       [C.c_double, C.c_int] + [vp] * 13)
        L.oracle_fx_launch_mmd.argtypes = [C.c_int] * 9 + [C.c_double, C.c_int] + [vp] * 15
        _lib = L
    return _lib


def _b(basis):
    return (basis.shell_l, np.ascontiguousarray(basis.shell_centre),
            basis.shell_nprim, basis.shell_pstart, basis.exps, basis.coefs)


def boys(M: int, T: float) -> np.ndarray:
    out = np.zeros(M + 1)
    lib().oracle_boys(M, float(T), out)
    return out


def eri(basis, q, method: str = "hgp") -> np.ndarray:
    shape = tuple(basis.shells[s].n_cart for s in q)
    out = np.zeros(int(np.prod(shape)))
    fn = lib().oracle_eri_hgp if method == "hgp" else lib().oracle_eri_mmd
    fn(*[int(s) for s in q], *_b(basis), out)
    return out.reshape(shape)


def fock_ordered(basis, D) -> tuple[np.ndarray, np.ndarray]:
    n = basis.n_ao
    J, K = np.zeros((n, n)), np.zeros((n, n))
    lib().oracle_fock_ordered(basis.n_sh, *_b(basis), basis.shell_ao, n,
                              np.ascontiguousarray(D, dtype=np.float64), J, K)
    return J, K


def grad_ordered(basis, shell_atom, n_atom: int, D) -> np.ndarray:
    """dE_2e/dR [n_atom, 3] by every ordered shell quartet, the first
    centre differentiated (fj_oracle.c's oracle_grad_ordered)."""
    G = np.zeros((n_atom, 3))
    lib().oracle_grad_ordered(basis.n_sh, *_b(basis), basis.shell_ao,
                              np.ascontiguousarray(shell_atom, dtype=np.int32), basis.n_ao,
                              np.ascontiguousarray(D, dtype=np.float64), int(n_atom), G)
    return G


def energy_2e(basis, D) -> float:
    """E_2e = 1/2 tr(D J) - 1/4 tr(D K) by the ordered sweep."""
    J, K = fock_ordered(basis, D)
    return float(0.5 * np.sum(D * J) - 0.25 * np.sum(D * K))


def threads() -> int:
    """The OpenMP threads of the oracle's parallel loops."""
    return int(lib().oracle_threads())


class _Flat:
    """Shell arrays in the oracle's layout (any l the MMD path takes)."""

    def __init__(self, parts):
        self.shell_l = np.concatenate([p[0] for p in parts]).astype(np.int32)
        self.shell_centre = np.ascontiguousarray(np.concatenate([p[1] for p in parts]), dtype=np.float64)
        self.shell_nprim = np.concatenate([p[2] for p in parts]).astype(np.int32)
        self.exps = np.ascontiguousarray(np.concatenate([p[3] for p in parts]), dtype=np.float64)
        self.coefs = np.ascontiguousarray(np.concatenate([p[4] for p in parts]), dtype=np.float64)
        self.shell_pstart = np.concatenate([[0], np.cumsum(self.shell_nprim)[:-1]]).astype(np.int32)
        self.shells = [type("S", (), {"n_cart": (int(l) + 1) * (int(l) + 2) // 2})()
                       for l in self.shell_l]  # noqa: E741


def _parts(b):
    return (b.shell_l, b.shell_centre, b.shell_nprim, b.exps, b.coefs)


# the unit s "shell": one primitive of exponent 0 and coefficient 1. Paired
# with an aux shell P it is P's single-Gaussian Hermite pair: zeta = alpha,
# the product centre A, no Gaussian factor (exp(0) = 1), dressing
# c (pi/alpha)^{3/2} -- the upstream's aux record (metric.py:1-90) -- so the
# oracle's four-center MMD integral (P unit|Q unit) is (P|Q) and (mn|P unit)
# is (mn|P). Pinned to libcint's int2c2e / int3c2e in the tests.
_UNIT = (np.array([0]), np.zeros((1, 3)), np.array([1]), np.array([0.0]), np.array([1.0]))


def df_metric(aux) -> np.ndarray:
    """(P|Q) of every aux function, [n_aux, n_aux]."""
    f = _Flat([_parts(aux), _UNIT])
    u = len(f.shell_l) - 1
    off = np.concatenate([[0], np.cumsum([s.n_cart for s in f.shells[:u]])])
    M = np.zeros((off[-1], off[-1]))
    for p in range(u):
        for q in range(p, u):
            x = eri(f, (p, u, q, u), "mmd")[:, 0, :, 0]
            M[off[p]:off[p + 1], off[q]:off[q + 1]] = x
            M[off[q]:off[q + 1], off[p]:off[p + 1]] = x.T
    return M


def df_eri3(basis, aux) -> np.ndarray:
    """(mn|P), [n_ao, n_ao, n_aux]."""
    f = _Flat([_parts(basis), _parts(aux), _UNIT])
    nb, u = len(basis.shell_l), len(f.shell_l) - 1
    sz = [s.n_cart for s in f.shells]
    ao = np.concatenate([[0], np.cumsum(sz[:nb])])
    xo = np.concatenate([[0], np.cumsum(sz[nb:u])])
    out = np.zeros((ao[-1], ao[-1], xo[-1]))
    for m in range(nb):
        for n in range(m, nb):
            for p in range(u - nb):
                x = eri(f, (m, n, nb + p, u), "mmd")[:, :, :, 0]
                out[ao[m]:ao[m + 1], ao[n]:ao[n + 1], xo[p]:xo[p + 1]] = x
                out[ao[n]:ao[n + 1], ao[m]:ao[m + 1], xo[p]:xo[p + 1]] = x.transpose(1, 0, 2)
    return out


def schwarz(basis) -> np.ndarray:
    n = basis.n_ao
    Q = np.zeros((n, n))
    lib().oracle_schwarz(basis.n_sh, *_b(basis), basis.shell_ao, n, Q)
    return Q


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def f_launch(cls, args: dict, F: np.ndarray) -> None:
    """Run one fused-F launch on the CPU, adding into F (n_ao x n_ao).
    `args` holds the device entry's arguments by name (numpy arrays)."""
    la, lb, lc, ld = cls
    a = args
    lib().oracle_f_launch(
        la, lb, lc, ld, a["n_bra"], a["n_ket"], a["k_ab"], a["k_cd"], a["n_sh"],
        a["n_ao"], a["diag"], float(a["tau"]), _ptr(a["bra_prim"]), _ptr(a.get("AB")),
        _ptr(a["G_ab"]), _ptr(a["bra_sh"]), _ptr(a["bra_ao"]), _ptr(a["ket_prim"]),
        _ptr(a.get("CD")), _ptr(a["G_cd"]), _ptr(a["ket_sh"]), _ptr(a["ket_ao"]),
        _ptr(a["Dsh"]), _ptr(a["D"]), _ptr(F))


def fx_launch(cls, bra: dict, ket: dict, diag: int, n_sh: int, n_ao: int, tau: float,
              Dsh, D, F, pre: bool = False) -> None:
    """Ragged-degree launch (pair i's records prim[off[i]:off[i+1]]), adding
    into F. Each side: prim [n_rec, 8], off [n+1], sep [n, 3], G, sh, ao.
    pre: also skip the quartets with G_ab G_cd < tau (the density-free
    pre-launch screen, per quartet)."""
    la, lb, lc, ld = cls
    c = lambda a: np.ascontiguousarray(a)  # noqa: E731
    b, k = {x: c(v) for x, v in bra.items()}, {x: c(v) for x, v in ket.items()}
    lib().oracle_fx_launch(
        la, lb, lc, ld, len(b["G"]), len(k["G"]), 0, 0, _ptr(b["off"]), _ptr(k["off"]),
        n_sh, n_ao, diag, float(tau), int(bool(pre)), _ptr(b["prim"]), _ptr(b["sep"]), _ptr(b["G"]),
        _ptr(b["sh"]), _ptr(b["ao"]), _ptr(k["prim"]), _ptr(k["sep"]), _ptr(k["G"]),
        _ptr(k["sh"]), _ptr(k["ao"]), _ptr(c(Dsh)), _ptr(c(D)), _ptr(F))


def fx_launch_mmd(cls, bra: dict, ket: dict, diag: int, basis, tau: float, Dsh, D, F,
                  pre: bool = False) -> None:
    """fx_launch's screen, weights and digestion with every quartet's
    integrals by MMD over the shells' full primitive lists (no HRR: full
    FP64 accuracy for tight-diffuse pairs, where the HGP transfer cancels).
    Each side needs G, sh, ao."""
    la, lb, lc, ld = cls
    c = lambda a: np.ascontiguousarray(a)  # noqa: E731
    b, k = {x: c(v) for x, v in bra.items()}, {x: c(v) for x, v in ket.items()}
    Dsh, D = c(np.asarray(Dsh, dtype=np.float64)), c(np.asarray(D, dtype=np.float64))
    arrs = [c(a) for a in _b(basis)]
    lib().oracle_fx_launch_mmd(
        la, lb, lc, ld, len(b["G"]), len(k["G"]), basis.n_sh, basis.n_ao, diag, tau,
        int(bool(pre)), _ptr(b["G"]), _ptr(b["sh"]), _ptr(b["ao"]), _ptr(k["G"]), _ptr(k["sh"]), _ptr(k["ao"]),
        _ptr(Dsh), _ptr(D), *[_ptr(a) for a in arrs], _ptr(F))
