"""Density fitting on the GPU: the metric (P|Q) and the three-center (mn|P)
from the loop-form kernel against libcint's int2c2e_cart / int3c2e_cart
(aux shells through i, AO shells through f: every class up to (ff|i) and
(i|i)) and against the oracle."""

import numpy as np
import pytest

from felizjueves_amd import df
from oracle import oracle as O
from test_df import _load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["df_prim_spd_f_auxi_metric", "df_h2o_ccpvtz_jkfit_metric"])
def test_metric_vs_libcint(name):
    z, aux, _ = _load(name)
    M = df.metric_matrix(aux).cpu().numpy()
    ref = z["int2c2e"]
    err = np.abs(M - ref).max() / np.abs(ref).max()
    print(name, "n_aux", aux.n_aux, "max l", int(aux.shell_l.max()), "rel err", err, flush=True)
    assert err < 1e-12
    assert np.array_equal(M, M.T)


@pytest.mark.parametrize("name", ["df_prim_spd_f_auxi_3c", "df_h2o_sto3g_jkfit_3c"])
def test_eri3_vs_libcint(name):
    z, aux, basis = _load(name)
    X = df.eri3(basis, aux).cpu().numpy()
    ref = z["int3c2e"]
    err = np.abs(X - ref).max() / np.abs(ref).max()
    eo = np.abs(X - O.df_eri3(basis, aux)).max() / np.abs(ref).max()
    print(name, "shape", X.shape, "rel err vs libcint", err, "vs oracle", eo, flush=True)
    assert err < 1e-12 and eo < 1e-12
    assert np.array_equal(X, X.transpose(1, 0, 2))


def test_eri3_streams_match_one_stream():
    """The staged (mn|P) with its classes over four streams equals the
    one-stream launch bit for bit (every class writes its own cells)."""
    import torch
    z, aux, basis = _load("df_prim_spd_f_auxi_3c")
    st = df.DfIntegrals(aux, basis)
    one = st.eri3().cpu().numpy()
    four = st.eri3(n_streams=4)
    torch.cuda.synchronize()
    assert np.array_equal(one, four.cpu().numpy())
