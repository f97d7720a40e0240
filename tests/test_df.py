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
