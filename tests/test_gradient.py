"""The fused two-electron gradient on the CPU: the oracle pinned, the
stream's structure, the host refusals, and the device kernel run unchanged
under emulation (CPU only).

THE ORACLE (oracle_grad_ordered: every ordered shell quartet, the first
centre differentiated with the symmetrised Gam, derivative integrals from
raised and lowered shells by McMurchie-Davidson) is held to

- PySCF's two-electron gradient terms grad:j + grad:k on the five upstream
  gradient fixtures (H2O STO-3G, STO-3G plus a d on O, 6-31G, cc-pVDZ,
  NH2F 6-31G), within 1e-10 absolute (the upstream's bound,
  test/test_kernel_gf.py ABS_TOL);
- central finite differences of its own energy E_2e = 1/2 tr(D J) -
  1/4 tr(D K) from the ordered J, K sweep (an independent route: no
  derivative integral), within the O(h^2) error of the difference;
- no net force.

A mistake in the weights (a dropped exchange term, a wrong factor), in the
raised shell's 2 alpha or in the lowered term's a_k fails the first two.

THE KERNEL (csrc/fj_gfloop.h) runs on every launch of a small s,p,d basis
where shells coincide on an atom (so a = b, c = d and the diagonal launch's
P = Q all occur) under AddressSanitizer, with two wavefronts per workgroup,
against the oracle."""

import os
import shutil
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

from felizjueves_amd import gradient as GR
from felizjueves_amd.basis import Basis, Shell, load_fixture, primitive_norm
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent
EMU = ROOT / "tests" / "emu"
CSRC = ROOT / "felizjueves_amd" / "csrc"
FIXTURES = ("grad_h2o_sto3g", "grad_h2o_sto3g_d", "grad_h2o_631g", "grad_h2o_ccpvdz",
            "grad_nh2f_631g")
ABS_TOL = 1e-10


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_is_pyscfs_two_electron_gradient(name):
    fx = load_fixture(name)
    a = fx.arrays
    G = O.grad_ordered(fx.basis, a["shell_atom"], len(a["atom_coord"]), fx.density)
    want = a["grad_j"] + a["grad_k"]
    err = np.abs(G - want).max()
    print(name, "max |G - PySCF|", err, flush=True)
    assert err <= ABS_TOL
    assert np.abs(G.sum(axis=0)).max() <= 1e-12 * np.abs(G).max()


def test_oracle_is_the_derivative_of_its_energy():
    fx = load_fixture("grad_h2o_sto3g_d")
    a = fx.arrays
    G = O.grad_ordered(fx.basis, a["shell_atom"], 3, fx.density)
    h = 1e-4
    for at in range(3):
        for x in range(3):
            e = []
            for s in (1, -1):
                C = a["shell_centre"].copy()
                C[a["shell_atom"] == at, x] += s * h
                b = Basis.from_arrays(a["shell_l"], C, a["shell_nprim"], a["exps"], a["coefs"])
                e.append(O.energy_2e(b, fx.density))
            fd = (e[0] - e[1]) / (2 * h)
            assert abs(G[at, x] - fd) <= 1e-7 * max(1.0, abs(fd)), (at, x, G[at, x], fd)


def _shell(l, centre, exps, raw):  # noqa: E741
    return Shell(l, tuple(centre), tuple(exps), tuple(w * primitive_norm(l, e) for w, e in zip(raw, exps)))


def small_system():
    """Two atoms: s (two primitives), p, d on the first; s, d on the second;
    D a random symmetric matrix."""
    A, B = (0.0, 0.0, 0.0), (0.6, -0.9, 1.2)
    shells = [_shell(0, A, (2.9, 0.4), (0.6, 0.5)), _shell(1, A, (1.2,), (1.0,)),
              _shell(2, A, (0.9,), (1.0,)), _shell(0, B, (1.1,), (1.0,)), _shell(2, B, (1.4,), (1.0,))]
    basis = Basis(shells)
    M = np.random.default_rng(11).normal(size=(basis.n_ao, basis.n_ao))
    return basis, np.array([0, 0, 0, 1, 1], dtype=np.int32), 0.1 * (M + M.T)


def test_the_stream_meets_every_unique_quartet_once():
    basis, at, _ = small_system()
    st = GR.gf_launches(basis, at)
    n_pair = basis.n_sh * (basis.n_sh + 1) // 2
    assert GR.quartets(st) == n_pair * (n_pair + 1) // 2
    seen = set()
    for L in st.launches:
        b, k = st.sides[L.bra], st.sides[L.ket]
        assert L.cls == (b.la, b.lb, k.la, k.lb) and b.la <= b.lb and k.la <= k.lb
        assert L.cls[:2] <= L.cls[2:]
        for i in range(b.n):
            for j in range(i if L.diag else 0, k.n):
                key = frozenset((tuple(b.sh[i]), tuple(k.sh[j])))
                assert key not in seen
                seen.add(key)
    assert len(seen) == n_pair * (n_pair + 1) // 2


def test_records_are_the_fused_record_and_the_exponents():
    basis, at, _ = small_system()
    st = GR.gf_launches(basis, at)
    side = next(s for s in st.sides if s.degree == 2)
    a, b = side.sh[0]
    rec = side.prim[0]
    e1 = basis.exps[basis.shell_pstart[a]:basis.shell_pstart[a] + 2]
    e2 = basis.exps[basis.shell_pstart[b]]
    assert np.allclose(rec[:, 8], e1) and np.allclose(rec[:, 9], e2)
    assert np.allclose(rec[:, 6], e1 + e2)
    A, B = basis.shell_centre[a], basis.shell_centre[b]
    assert np.allclose(rec[:, 3:6], rec[:, 0:3] - A)
    assert np.allclose(side.sep[0], A - B)


def test_refusals():
    basis, at, D = small_system()
    with pytest.raises(ValueError, match="symmetric"):
        GR.gf_gradient(basis, at, 2, D + np.triu(np.ones_like(D), 1))
    with pytest.raises(ValueError, match="entries"):
        GR.gf_launches(basis, at[:-1])
    f = Basis([_shell(3, (0.0, 0.0, 0.0), (1.0,), (1.0,))])
    with pytest.raises(ValueError, match="through l = 2"):
        GR.gf_launches(f, [0])


def _emu(td: Path, driver: str) -> Path:
    (td / "gen").mkdir(exist_ok=True)
    shutil.copy(EMU / "fj_device_emu.h", td / "fj_device.h")
    shutil.copy(CSRC / "gen" / "fj_boys_table.h", td / "gen" / "fj_boys_table.h")
    exe = td / Path(driver).stem
    r = subprocess.run(["g++", "-std=c++20", "-O1", "-g0", "-fsanitize=address", f"-I{ROOT / 'include'}",
                        f"-I{td}", f'-DGFLOOP_HEADER="{CSRC / "fj_gfloop.h"}"', str(EMU / driver),
                        "-o", str(exe), "-lpthread"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _w(f, x, dt):
    x = np.ascontiguousarray(x, dtype=dt).ravel()
    np.array([x.size], dtype=np.int64).tofile(f)
    x.tofile(f)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_gradient_kernel_on_cpu_vs_oracle():
    """Every launch of the stream through the single-launch kernel (the
    per-class entries' path), each quartet split into work items of half its
    primitive quartets, one wavefront per workgroup."""
    basis, at, D = small_system()
    st = GR.gf_launches(basis, at)
    assert {L.cls for L in st.launches} >= {(0, 0, 0, 0), (2, 2, 2, 2), (0, 2, 1, 2), (1, 2, 2, 2)}
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        exe = _emu(td, "gfloop_emu.cpp")
        w = _w
        with open(td / "case.bin", "wb") as f:
            np.array([len(st.launches), 2, basis.n_ao], dtype=np.int32).tofile(f)
            w(f, D, np.float64)
            for L in st.launches:
                b, k = st.sides[L.bra], st.sides[L.ket]
                # items of half a quartet's primitive quartets (uneven pieces)
                chunk = max(1, b.degree * k.degree // 2)
                np.array([*L.cls, b.n, k.n, b.degree, k.degree, 2, basis.n_ao, L.diag, chunk],
                         dtype=np.int32).tofile(f)
                for s in (b, k):
                    w(f, s.prim, np.float64)
                    w(f, s.sep, np.float64)
                    w(f, s.sh, np.int32)
                    w(f, s.ao, np.int32)
                    w(f, s.at, np.int32)
        r = subprocess.run([str(exe), str(td / "case.bin"), str(td / "G.bin"), "64"], cwd=td,
                           env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"), capture_output=True,
                           text=True, timeout=900)
        assert r.returncode == 0, r.stderr[-3000:]
        G = np.fromfile(td / "G.bin").reshape(2, 3)
    Go = O.grad_ordered(basis, at, 2, D)
    err = np.abs(G - Go).max()
    print("emulated gradient kernel vs oracle", err, "scale", np.abs(Go).max(), flush=True)
    assert err <= 1e-12 * np.abs(Go).max()
    assert np.abs(G.sum(axis=0)).max() <= 1e-12 * np.abs(G).max()


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_gradient_merged_launches_on_cpu_vs_oracle():
    """Every class as one merged launch over its job table (fj_gfjobs_k, the
    runner's default path): the job lookup and the concatenated sides' offsets
    under AddressSanitizer, a grid of seven workgroups across the jobs, each
    quartet split into items of half its primitive quartets (the GPU tests
    run it at two and four wavefronts per workgroup)."""
    basis, at, D = small_system()
    st = GR.gf_launches(basis, at)
    rows = np.concatenate([[0], np.cumsum([sd.n for sd in st.sides])])
    recs = np.concatenate([[0], np.cumsum([sd.prim.size for sd in st.sides])])
    by: dict = {}
    for L in st.launches:
        by.setdefault(L.cls, []).append(L)
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        exe = _emu(td, "gfjobs_emu.cpp")
        with open(td / "case.bin", "wb") as f:
            np.array([len(by), 2, basis.n_ao], dtype=np.int32).tofile(f)
            _w(f, D, np.float64)
            _w(f, np.concatenate([sd.prim.ravel() for sd in st.sides]), np.float64)
            for k in ("sep", "sh", "ao", "at"):
                _w(f, np.concatenate([getattr(sd, k) for sd in st.sides]),
                   np.float64 if k == "sep" else np.int32)
            for cls, Ls in sorted(by.items()):
                jobs = np.zeros(len(Ls), dtype=GR.JOB_DTYPE)
                q0 = 0
                for j, L in enumerate(Ls):
                    b, k = st.sides[L.bra], st.sides[L.ket]
                    n_pq = b.degree * k.degree
                    chunk = max(1, n_pq // 2)     # items of half a quartet's primitive quartets
                    jobs[j] = (rows[L.bra], rows[L.ket], b.n, k.n, b.degree, k.degree, L.diag, chunk,
                               recs[L.bra], recs[L.ket], q0)
                    q0 += (b.n * (b.n + 1) // 2 if L.diag else b.n * k.n) * -(-n_pq // chunk)
                np.array([*cls, len(Ls), int((jobs["k_ab"] * jobs["k_cd"]).max())], dtype=np.int32).tofile(f)
                np.array([q0], dtype=np.int64).tofile(f)
                np.array([len(Ls)], dtype=np.int64).tofile(f)
                f.write(jobs.tobytes())
        r = subprocess.run([str(exe), str(td / "case.bin"), str(td / "G.bin"), "64"], cwd=td,
                           env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"), capture_output=True,
                           text=True, timeout=900)
        assert r.returncode == 0, r.stderr[-3000:]
        G = np.fromfile(td / "G.bin").reshape(2, 3)
    Go = O.grad_ordered(basis, at, 2, D)
    err = np.abs(G - Go).max()
    print("emulated merged launches vs oracle", err, "scale", np.abs(Go).max(), flush=True)
    assert err <= 1e-12 * np.abs(Go).max()
