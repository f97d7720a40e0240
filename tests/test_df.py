"""Density fitting on the CPU: the oracle's metric (P|Q) and three-center
(mn|P) against libcint's int2c2e_cart / int3c2e_cart (the upstream's
tracked fixtures, aux shells through i), and the staging of the device
kernels (every element written by exactly one cell, mirrored)."""

import numpy as np
import pytest

from felizjueves_amd import df
from felizjueves_amd.basis import Basis
from oracle import oracle as O
from pathlib import Path

GOLDEN = Path(__file__).resolve().parent / "golden"


def _load(name):
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    aux = df.AuxBasis.from_npz(z)
    basis = Basis.from_arrays(z["shell_l"], z["shell_centre"], z["shell_nprim"], z["exps"],
                              z["coefs"]) if "shell_l" in z else None
    return z, aux, basis


@pytest.mark.parametrize("name", ["df_prim_spd_f_auxi_metric", "df_h2o_ccpvtz_jkfit_metric"])
def test_oracle_metric_vs_libcint(name):
    z, aux, _ = _load(name)
    M = O.df_metric(aux)
    ref = z["int2c2e"]
    assert M.shape == ref.shape and max(aux.shell_l) >= 4
    assert np.abs(M - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["df_prim_spd_f_auxi_3c", "df_h2o_sto3g_jkfit_3c"])
def test_oracle_eri3_vs_libcint(name):
    z, aux, basis = _load(name)
    X = O.df_eri3(basis, aux)
    ref = z["int3c2e"]
    assert X.shape == ref.shape
    assert np.abs(X - ref).max() < 1e-12 * np.abs(ref).max()


def test_staging_writes_every_element_once():
    """Metric pairs cover every (P, Q) block exactly once with their
    mirrors; eri3 cells every (m, n, P) block likewise; a class never
    stages the higher l first."""
    z, aux, basis = _load("df_prim_spd_f_auxi_3c")
    cover = np.zeros((aux.n_aux, aux.n_aux), dtype=int)
    for (lp, lq), dc in df.metric_classes(aux).items():
        assert lp <= lq
        for op, oq in zip(dc.arrays["off_p"], dc.arrays["off_q"]):
            sp = aux.shell_size[list(aux.shell_ao).index(op)]
            sq = aux.shell_size[list(aux.shell_ao).index(oq)]
            cover[op:op + sp, oq:oq + sq] += 1
            if op != oq:
                cover[oq:oq + sq, op:op + sp] += 1
    assert (cover == 1).all()
    classes, bra = df.eri3_classes(basis, aux)
    cov = np.zeros((basis.n_ao, basis.n_ao, aux.n_aux), dtype=int)
    ao = list(basis.shell_ao)
    for (la, lb, lp), dc in classes.items():
        assert la <= lb
        a = dc.arrays
        for oa, ob, op, r0, r1 in zip(a["off_a"], a["off_b"], a["off_p"], a["ab_rec0"], a["ab_rec1"]):
            sa, sb = basis.shell_size[ao.index(oa)], basis.shell_size[ao.index(ob)]
            sp = aux.shell_size[list(aux.shell_ao).index(op)]
            assert r1 - r0 == basis.shell_nprim[ao.index(oa)] * basis.shell_nprim[ao.index(ob)]
            cov[oa:oa + sa, ob:ob + sb, op:op + sp] += 1
            if oa != ob:
                cov[ob:ob + sb, oa:oa + sa, op:op + sp] += 1
    assert (cov == 1).all()


def test_aux_basis_refusals():
    with pytest.raises(ValueError):
        df.AuxBasis([7], [[0, 0, 0]], [1], [1.0], [1.0])
    with pytest.raises(ValueError):
        df.AuxBasis([0], [[0, 0, 0]], [1], [0.0], [1.0])


ROOT = Path(__file__).resolve().parent.parent
EMU = ROOT / "tests" / "emu"
CSRC = ROOT / "felizjueves_amd" / "csrc"


def _emulate(kind: str, aux, basis, threads: int) -> np.ndarray:
    """Every class of one metric_matrix / eri3 call through the device
    kernel compiled for the CPU (tests/emu: a thread per lane) under
    AddressSanitizer, the host tables exactly as df.py stages them."""
    import os
    import shutil
    import subprocess
    import tempfile
    A, alpha, s_aux, _, _ = aux.records()
    if kind == "metric":
        classes = df.metric_classes(aux)
        bra = {"zeta_ab": alpha, "P": A, "XPA": np.zeros((len(alpha), 3)), "XPB": np.zeros((len(alpha), 3)),
               "s_ab": s_aux}
        keys = ("p_rec0", "p_rec1", "q_rec0", "q_rec1", "off_p", "off_p", "off_q")
        n_ao = 0
    else:
        classes, bra = df.eri3_classes(basis, aux)
        keys = ("ab_rec0", "ab_rec1", "p_rec0", "p_rec1", "off_a", "off_b", "off_p")
        n_ao = basis.n_ao
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        (td / "gen").mkdir()
        shutil.copy(EMU / "fj_device_emu.h", td / "fj_device.h")
        shutil.copy(CSRC / "gen" / "fj_boys_table.h", td / "gen" / "fj_boys_table.h")
        exe = td / "dfloop_emu"
        r = subprocess.run(["g++", "-std=c++20", "-O1", "-g0", "-fsanitize=address", f"-I{ROOT / 'include'}",
                            f"-I{td}", f'-DDFLOOP_HEADER="{CSRC / "fj_dfloop.h"}"', str(EMU / "dfloop_emu.cpp"),
                            "-o", str(exe), "-lpthread"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]

        def w(f, x, dt):
            x = np.ascontiguousarray(x, dtype=dt).ravel()
            np.array([x.size], dtype=np.int64).tofile(f)
            x.tofile(f)
        with open(td / "case.bin", "wb") as f:
            np.array([1 if kind == "metric" else 0, n_ao, aux.n_aux, len(classes)], dtype=np.int32).tofile(f)
            for k in ("zeta_ab", "P", "XPA", "XPB", "s_ab"):
                w(f, bra[k], np.float64)
            for x in (A, alpha, s_aux):
                w(f, x, np.float64)
            for cls, dc in sorted(classes.items()):
                a = dc.arrays
                r0, r1, q0, q1 = (a[k] for k in keys[:4])
                max_rp = int(((r1 - r0) * (q1 - q0)).max())
                l3 = (cls[0], 0, cls[1]) if kind == "metric" else cls
                np.array([*l3, dc.n, max_rp], dtype=np.int32).tofile(f)
                for k in keys:
                    w(f, a[k], np.int32)
        r = subprocess.run([str(exe), str(td / "case.bin"), str(td / "out.bin"), str(threads)], cwd=td,
                           env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"), capture_output=True,
                           text=True, timeout=900)
        assert r.returncode == 0, r.stderr[-3000:]
        out = np.fromfile(td / "out.bin")
    return out.reshape((aux.n_aux, aux.n_aux) if kind == "metric" else (n_ao, n_ao, aux.n_aux))


def test_df_kernel_on_cpu_vs_oracle():
    """The device kernel run unchanged on the CPU: the metric of the s..i aux
    set (every metric class through (i|i), record pairs batched) and the
    three-center integrals of H2O/STO-3G with the JKFIT aux set (every cell
    with its bra primitive pairs batched), two wavefronts per workgroup,
    against the oracle."""
    z, aux, _ = _load("df_prim_spd_f_auxi_metric")
    M = _emulate("metric", aux, None, 128)
    Mo = O.df_metric(aux)
    err = np.abs(M - Mo).max() / np.abs(Mo).max()
    print("emulated metric vs oracle", err, flush=True)
    assert err < 1e-13
    z, aux, basis = _load("df_h2o_sto3g_jkfit_3c")
    X = _emulate("eri3", aux, basis, 64)
    Xo = O.df_eri3(basis, aux)
    err = np.abs(X - Xo).max() / np.abs(Xo).max()
    print("emulated eri3 vs oracle", err, flush=True)
    assert err < 1e-13
