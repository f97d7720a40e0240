"""The loop-form MMD kernels (csrc/fj_kloop.h) run unchanged on the CPU
(CPU only): compiled with g++ against the emulation of the device header
(one thread per lane, barriers for __syncthreads) under AddressSanitizer.

K takes every K launch of a two-atom s,p,d,f basis -- the same host tables
and launch order as `jk.ExchangeBuild` -- and, after the scatter, its K is
compared with the oracle's K (every ordered shell quartet by
McMurchie-Davidson in C). Schwarz takes every pair class of the same basis
(`jk.schwarz_tables`) and its Q is compared with the oracle's.

Every class through (ff|ff) and (ff) runs through the loop forms here, the
lower classes included, so each kernel is checked over all the classes the
library launches it for or could."""

import os
import shutil
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

from felizjueves_amd import jk
from felizjueves_amd.basis import Basis, Shell, n_cart, primitive_norm
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent
EMU = ROOT / "tests" / "emu"
CSRC = ROOT / "felizjueves_amd" / "csrc"


def _shell(l, centre, exps, raw):  # noqa: E741
    c = [w * primitive_norm(l, a) for w, a in zip(raw, exps)]
    return Shell(l, tuple(centre), tuple(exps), tuple(c))


def _system(contracted: bool = False):
    """Two atoms 1.9 bohr apart with s, p, d, f each (the first atom's s
    contracted over two primitives); D a random symmetric matrix. With
    `contracted`, every s shell has three primitives and every p shell two,
    so that a quartet has up to 81 primitive quartets."""
    A, B = (0.0, 0.0, 0.0), (0.7, -1.1, 1.3)

    def sh(l, X, e):  # noqa: E741
        if not contracted:
            return _shell(l, X, (e,), (1.0,))
        if l == 0:
            return _shell(0, X, (4.0 * e, e, 0.3 * e), (0.3, 0.5, 0.4))
        if l == 1:
            return _shell(1, X, (2.5 * e, e), (0.4, 0.7))
        return _shell(l, X, (e,), (1.0,))
    shells = [sh(0, A, 0.45) if contracted else _shell(0, A, (3.1, 0.45), (0.6, 0.5))]
    for X, sc in ((A, 1.0), (B, 0.8)):
        if X is B:
            shells.append(sh(0, X, 0.5))
        shells += [sh(1, X, 1.3 * sc), sh(2, X, 0.9 * sc), sh(3, X, 0.7 * sc)]
    basis = Basis(shells)
    rng = np.random.default_rng(5)
    M = rng.normal(size=(basis.n_ao, basis.n_ao))
    return basis, 0.1 * (M + M.T)


def _dump(path: Path, basis, D, tables, launches, tau, maxD) -> None:
    def w(f, x, dt):
        x = np.ascontiguousarray(x, dtype=dt).ravel()
        np.array([x.size], dtype=np.int64).tofile(f)
        x.tofile(f)
    with open(path, "wb") as f:
        np.array([len(launches)], dtype=np.int32).tofile(f)
        for la, lc, lb, ld, bs in launches:
            t, k = tables["bras"][(la, lc)], tables["kets"][(lb, ld, bs)]
            blk = tables["dblk"][(lc, ld)]
            np.array([la, lc, lb, ld, t["n"], k["n"], blk["n"]], dtype=np.int32).tofile(f)
            np.array([tau, maxD], dtype=np.float64).tofile(f)
            for x in (t["P"], t["z"], t["XA"], t["XC"], t["s"], t["Ip"]):
                w(f, x, np.float64)
            w(f, t[("dc", ld)], np.int32)
            for x in (k["Q"], k["eta"], k["XB"], k["XD"], k["s"], k["Iq"]):
                w(f, x, np.float64)
            w(f, k["pos"], np.int32)
            w(f, D.ravel()[blk["idx"]], np.float64)


def _build(td: Path, driver: str) -> Path:
    (td / "gen").mkdir(exist_ok=True)
    shutil.copy(EMU / "fj_device_emu.h", td / "fj_device.h")
    shutil.copy(CSRC / "gen" / "fj_boys_table.h", td / "gen" / "fj_boys_table.h")
    exe = td / Path(driver).stem
    cmd = ["g++", "-std=c++20", "-O1", "-g0", "-fsanitize=address", f"-I{ROOT / 'include'}",
           f"-I{td}", f'-DKLOOP_HEADER="{CSRC / "fj_kloop.h"}"', str(EMU / driver),
           "-o", str(exe), "-lpthread"]
    r = subprocess.run(cmd, cwd=td, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _run(exe: Path, td: Path, out: str, *extra: str) -> np.ndarray:
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([str(exe), str(td / "case.bin"), str(td / out), *extra], cwd=td, env=env,
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    return np.fromfile(td / out)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_loop_schwarz_kernel_on_cpu_vs_oracle():
    basis, _ = _system()
    tables = jk.schwarz_tables(basis)
    assert (3, 3) in tables and (2, 3) in tables
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        exe = _build(td, "schwarz_loop_emu.cpp")
        with open(td / "case.bin", "wb") as f:
            np.array([len(tables), basis.n_ao], dtype=np.int32).tofile(f)
            for (la, lb), t in tables.items():
                np.array([la, lb, t["n_pair"], t["n_rec"]], dtype=np.int32).tofile(f)
                for k, dt in (("z", np.float64), ("P", np.float64), ("XA", np.float64),
                              ("XB", np.float64), ("s", np.float64), ("rec0", np.int32),
                              ("rec1", np.int32), ("oa", np.int32), ("ob", np.int32)):
                    x = np.ascontiguousarray(t[k], dtype=dt).ravel()
                    np.array([x.size], dtype=np.int64).tofile(f)
                    x.tofile(f)
        Q = _run(exe, td, "Q.bin").reshape(basis.n_ao, basis.n_ao)
    Qo = O.schwarz(basis)
    err = np.abs(Q - Qo).max() / Qo.max()
    print("loop-form Schwarz vs oracle, max rel", err, flush=True)
    assert err < 1e-12
    assert np.array_equal(Q, Q.T)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_loop_k_kernel_on_cpu_vs_oracle():
    basis, D = _system()
    tables = jk.exchange_tables(basis)
    launches = list(jk.exchange_launches(tables))
    maxD = float(np.abs(D).max())
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        exe = _build(td, "kloop_emu.cpp")
        _dump(td / "case.bin", basis, D, tables, launches, 0.0, maxD)
        # two wavefronts, as the library launches it; the kets cut into three
        # slices per bra, adding into the cleared K
        rows = _run(exe, td, "K.bin", "128", "3")
    K = np.zeros((basis.n_ao, basis.n_ao))
    o = 0
    for la, lc, lb, ld, bs in launches:
        t = tables["bras"][(la, lc)]
        na, nb = n_cart(la), n_cart(lb)
        Kr = rows[o:o + t["n"] * na * nb].reshape(t["n"], na, nb)
        o += Kr.size
        ob = int(basis.shell_ao[bs])
        for r_, oa in enumerate(t["oa"]):
            K[oa:oa + na, ob:ob + nb] += Kr[r_]
    assert o == rows.size
    _, Ko = O.fock_ordered(basis, D)
    err = np.abs(K - Ko).max()
    print("loop-form K vs oracle", err, "scale", np.abs(Ko).max(), flush=True)
    assert err <= 1e-12 * max(1.0, np.abs(Ko).max())


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
@pytest.mark.parametrize("mode,contracted,threads", [("one", False, "64"), ("batched", True, "64")])
def test_loop_f_kernel_on_cpu_vs_oracle(mode, contracted, threads):
    """The loop-form fused F kernels on every walker launch of the s,p,d,f
    basis (every canonical class through (ff|ff)), F against the oracle's
    launch-by-launch witness (the same screen and digestion, integrals by
    McMurchie-Davidson). "one": every class on the one-at-a-time kernel;
    "batched": the classes the launcher gives the batched kernel (a batch
    of two or more primitive quartets fits its LDS), on the contracted basis
    with batches of at most 7 (quartets with an f shell of up to 27
    primitive quartets: full batches, remainders and several batches per
    ket), F
    against the oracle's launches of the same classes. Both run one
    wavefront per workgroup (the kernels' strides and wavefront loops take
    the width as an argument; the library launches four): each emulated
    matrix-core step is an exchange at the wavefront's barrier, and one
    wavefront keeps them short."""
    from felizjueves_amd import pairs as P
    basis, D = _system(contracted)
    plan = P.f_launches(basis, D, tau=0.0, n=1)
    Dsh = P.density_blocks(basis, D)
    launches = plan.launches
    assert max(max(L.cls) for L in launches) == 3
    Fo = np.zeros((basis.n_ao, basis.n_ao))
    parts = []   # each launch's F from the oracle
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        (td / "gen").mkdir(exist_ok=True)
        shutil.copy(EMU / "fj_device_emu.h", td / "fj_device.h")
        shutil.copy(CSRC / "gen" / "fj_boys_table.h", td / "gen" / "fj_boys_table.h")
        exe = td / "floop_emu"
        cmd = ["g++", "-std=c++20", "-O1", "-g0", "-fsanitize=address", f"-I{ROOT / 'include'}",
               f"-I{td}", f'-DKLOOP_HEADER="{CSRC / "fj_kloop.h"}"',
               f'-DFLOOP_HEADER="{CSRC / "fj_floop.h"}"', str(EMU / "floop_emu.cpp"),
               "-o", str(exe), "-lpthread"]
        r = subprocess.run(cmd, cwd=td, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        with open(td / "case.bin", "wb") as f:
            np.array([len(launches), basis.n_ao], dtype=np.int32).tofile(f)

            def w(x, dt):
                x = np.ascontiguousarray(x, dtype=dt).ravel()
                np.array([x.size], dtype=np.int64).tofile(f)
                x.tofile(f)
            for L in launches:
                a = P.launch_args(basis, plan, L, D, Dsh)
                Fl = np.zeros_like(Fo)
                O.f_launch(L.cls, a, Fl)
                parts.append(Fl)
                np.array([*L.cls, a["n_bra"], a["n_ket"], a["k_ab"], a["k_cd"], a["n_sh"],
                          a["n_ao"], a["diag"]], dtype=np.int32).tofile(f)
                np.array([a["tau"]], dtype=np.float64).tofile(f)
                w(a["bra_prim"], np.float64)
                if L.cls[1]:
                    w(a["AB"], np.float64)
                w(a["G_ab"], np.float64)
                w(a["bra_sh"], np.int32)
                w(a["bra_ao"], np.int32)
                w(a["ket_prim"], np.float64)
                if L.cls[3]:
                    w(a["CD"], np.float64)
                w(a["G_cd"], np.float64)
                w(a["ket_sh"], np.int32)
                w(a["ket_ao"], np.int32)
                w(a["Dsh"], np.float64)
                w(a["D"], np.float64)
        # batches of at most 7 on the contracted basis: full batches,
        # remainders and several batches per ket in every batched class
        extra = ("7",) if mode == "batched" else ()
        F = _run(exe, td, "F.bin", threads, mode, *extra).reshape(basis.n_ao, basis.n_ao)
        ran = np.fromfile(td / "F.bin.ran", dtype=np.int32)
    assert len(ran) == len(launches)
    if mode == "batched":   # both kernels' classes are present; only the batched ones ran
        assert 0 < ran.sum() < len(ran)
    else:
        assert ran.all()
    Fo = sum(Fl for Fl, r in zip(parts, ran) if r)
    err = np.abs(F - Fo).max()
    print("loop-form F vs oracle", err, "scale", np.abs(Fo).max(), "launches", len(launches), flush=True)
    assert err <= 1e-12 * max(1.0, np.abs(Fo).max())
