"""The fused two-electron gradient on the GPU, every launch of the stream
into one zeroed G: each class as one merged launch over its job table
(fj_gf_jobs, the runner's default) and, separately, one fj_gf_<cls> call per
launch (the upstream gf_ parameter list):

- PySCF's grad:j + grad:k on the five upstream gradient fixtures within
  1e-10 absolute (the upstream's bound, test/test_kernel_gf.py ABS_TOL),
  and the oracle (oracle_grad_ordered) within 1e-12 of max|G|;
- the three-atom s,s,p,d system (contractions 1-6, shells coinciding on
  every atom, so a = b, c = d and the diagonal launch's P = Q occur)
  against the oracle within 1e-11 of max|G|, at every workgroup size (its
  exponents reach 1e4, so the raised integrals carry 2 alpha ~ 2e4 and
  cancel over the quartets: both sides' FP64 rounding is ~1e-12 of max|G|,
  two orders under the 1e-10 absolute bar of the fused F build);
- no net force (all four centres are differentiated, translational
  invariance is not used, so this is a check);
- the entry's refusals (a class past d, a diagonal launch of unequal
  sides, a missing shell-index array where la == lb)."""

import ctypes as C

import numpy as np
import pytest

from felizjueves_amd import gradient as GR
from felizjueves_amd import native
from felizjueves_amd.basis import load_fixture
from felizjueves_amd.systems import config1
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIXTURES = ("grad_h2o_sto3g", "grad_h2o_sto3g_d", "grad_h2o_631g", "grad_h2o_ccpvdz",
            "grad_nh2f_631g")


@pytest.mark.parametrize("name", FIXTURES)
def test_gradient_vs_pyscf_and_oracle(name):
    fx = load_fixture(name)
    a = fx.arrays
    n_atom = len(a["atom_coord"])
    G = GR.gf_gradient(fx.basis, a["shell_atom"], n_atom, fx.density)
    want = a["grad_j"] + a["grad_k"]
    Go = O.grad_ordered(fx.basis, a["shell_atom"], n_atom, fx.density)
    e_py, e_or = np.abs(G - want).max(), np.abs(G - Go).max()
    print(name, "max |G - PySCF|", e_py, "max |G - oracle|", e_or, flush=True)
    assert e_py <= 1e-10
    assert e_or <= 1e-12 * np.abs(Go).max()
    assert np.abs(G.sum(axis=0)).max() <= 1e-12 * np.abs(G).max()


@pytest.mark.parametrize("threads", [64, 128, 256])
def test_gradient_three_atom_spd_vs_oracle(threads):
    s = config1()
    at = GR.shell_atoms(s.basis, s.coords)
    st = GR.gf_launches(s.basis, at)
    assert len({L.cls for L in st.launches}) == 21   # every class through (dd|dd)
    run = GR.GpuGfRunner(st, len(s.coords), s.basis.n_ao,
                         threads={L.cls: threads for L in st.launches})
    G = GR.gf_gradient(s.basis, at, len(s.coords), s.D, runner=run)
    Go = O.grad_ordered(s.basis, at, len(s.coords), s.D)
    err = np.abs(G - Go).max()
    print("threads", threads, "max |G - oracle|", err, "scale", np.abs(Go).max(), "quartets",
          GR.quartets(st), flush=True)
    assert err <= 1e-11 * np.abs(Go).max()
    assert np.abs(G.sum(axis=0)).max() <= 1e-11 * np.abs(G).max()


def test_gradient_per_launch_entries_vs_merged():
    """The per-class entries (one call per launch of the stream, the
    upstream ABI) give the merged launches' G, both the oracle's."""
    s = config1()
    at = GR.shell_atoms(s.basis, s.coords)
    st = GR.gf_launches(s.basis, at)
    Gl = GR.gf_gradient(s.basis, at, len(s.coords), s.D,
                        runner=GR.GpuGfRunner(st, len(s.coords), s.basis.n_ao, mode="launches"))
    Gj = GR.gf_gradient(s.basis, at, len(s.coords), s.D,
                        runner=GR.GpuGfRunner(st, len(s.coords), s.basis.n_ao))
    Go = O.grad_ordered(s.basis, at, len(s.coords), s.D)
    print("launches vs merged", np.abs(Gl - Gj).max(), "launches vs oracle", np.abs(Gl - Go).max(),
          "calls", len(st.launches), flush=True)
    assert np.abs(Gl - Gj).max() <= 1e-12 * np.abs(Go).max()
    assert np.abs(Gl - Go).max() <= 1e-11 * np.abs(Go).max()


def test_gradient_entry_refusals():
    lib = native.lib()
    fn = lib.fj_gf_loop
    fn.restype = C.c_int
    D = native.to_dev(np.eye(4))
    G = native.to_dev(np.zeros((1, 3)))
    # one primitive pair at the origin: zeta = 1 from exponents 1/2, 1/2, U = 1
    rec = native.to_dev(np.array([[[0, 0, 0, 0, 0, 0, 1.0, 1.0, 0.5, 0.5]]]))
    sep = native.to_dev(np.zeros((1, 3)))
    i2 = native.to_dev(np.zeros((1, 2), dtype=np.int32))
    p = lambda t: C.c_void_p(native.ptr(t))  # noqa: E731

    def call(la, lb, lc, ld, n_bra=1, n_ket=1, diag=0, sh=True):
        return fn(None, la, lb, lc, ld, 64, n_bra, n_ket, 1, 1, 1, 4, diag, p(rec), p(sep),
                  p(i2) if sh else None, p(i2), p(i2), p(rec), p(sep), p(i2) if sh else None, p(i2),
                  p(i2), p(D), p(G))
    assert call(3, 3, 0, 0) == 9004                      # FJ_ERR_CLASS
    assert call(0, 0, 0, 0, n_bra=1, n_ket=2, diag=1) == 9003   # FJ_ERR_SHAPE
    assert call(0, 0, 0, 1, diag=1) == 9003
    assert call(0, 0, 0, 0, sh=False) == 9005            # FJ_ERR_ARG
    assert call(0, 1, 0, 1, sh=False) == 0               # no shell indices needed there
    import torch
    torch.cuda.synchronize()
    assert np.isfinite(G.cpu().numpy()).all()


def test_gradient_streams_match_one_stream():
    """The merged class launches over four streams give the one-stream
    gradient (atomic sums in another order: within 1e-13 of its scale)."""
    import torch
    from felizjueves_amd import native
    s = config1()
    at = GR.shell_atoms(s.basis, s.coords)
    st = GR.gf_launches(s.basis, at)
    run = GR.GpuGfRunner(st, len(s.coords), s.basis.n_ao)
    Dd = native.to_dev(s.D)
    one = run.run(Dd).cpu().numpy()
    four = run.run(Dd, n_streams=4)
    torch.cuda.synchronize()
    four = four.cpu().numpy()
    d = np.abs(four - one).max()
    print("four streams vs one", d, flush=True)
    assert d <= 1e-13 * np.abs(one).max()
