"""Edge cases and refusals on the host (CPU only): what the upstream
refuses before any launch (eri/fock/walk_f.py:213-304, kernel_f.py:572-600)
and what the builder refuses or plans for degenerate inputs; the C entries'
argument checks that return before any device call."""

import ctypes as C

import numpy as np
import pytest

from felizjueves_amd import native
from felizjueves_amd import pairs as P
from felizjueves_amd.basis import Basis, Shell
from felizjueves_amd.fock import FockBuilder
from felizjueves_amd.systems import config1

ERR_THREADS, ERR_ARG = 9001, 9005


def _one_s_shell():
    return Basis([Shell(0, (0.0, 0.0, 0.0), (1.3, 0.4), (0.6, 0.5))])


def test_walker_stream_refuses_asymmetric_density():
    s = config1()
    D = s.D.copy()
    D[0, 1] += 1e-9
    with pytest.raises(ValueError, match="symmetric"):
        P.f_launches(s.basis, D, tau=1e-12, n=1)


def test_walker_stream_refuses_no_magnitude_bins():
    s = config1()
    with pytest.raises(ValueError):
        P.f_launches(s.basis, s.D, tau=1e-12, n=0)


def test_builder_plans_f_classes_onto_built_entries():
    """An f shell is planned (no refusal), and every class a launch names has
    its fj_fx_ and fj_fxm_ entries in the header (the f classes on the
    loop-form kernel)."""
    from pathlib import Path
    b = Basis([Shell(3, (0.0, 0.0, 0.0), (1.0,), (1.0,)), Shell(0, (0.0, 0.0, 1.0), (1.0,), (1.0,))])
    D = np.eye(b.n_ao) * 0.1
    fb = FockBuilder(b, D, tau=1e-14)
    classes = {tuple(L.cls) for L in fb.launches}
    assert any(3 in c for c in classes)
    hdr = (Path(__file__).resolve().parent.parent / "include" / "fj_mi355x.h").read_text()
    for c in classes:
        n = "".join(map(str, c))
        assert f"int fj_fx_{n}(" in hdr and f"int fj_fxm_{n}(" in hdr, c


def test_builder_refuses_more_primitive_pairs_than_the_staging_holds():
    # 12 x 12 = 144 primitive pairs > 128 on the diagonal pair
    ex = tuple(float(x) for x in np.geomspace(0.1, 1e3, 12))
    b = Basis([Shell(0, (0.0, 0.0, 0.0), ex, tuple([0.3] * 12))])
    with pytest.raises(ValueError, match="primitive"):
        FockBuilder(b, np.eye(1))


def test_batch_of_no_systems_is_refused():
    with pytest.raises(ValueError, match="no systems"):
        FockBuilder.batch([])


def test_everything_screened_plans_nothing():
    s = config1()
    fb = FockBuilder(s.basis, s.D, tau=1e30)
    assert fb.launches == [] and fb.n_quartets == 0 and fb.shell_quartets() == 0


def test_single_s_shell_plans_one_quartet():
    b = _one_s_shell()
    fb = FockBuilder(b, np.array([[0.7]]), tau=1e-12)
    assert fb.n_quartets == 1 and len(fb.launches) == 1
    L = fb.launches[0]
    assert L.cls == (0, 0, 0, 0) and L.diag == 1


def test_pack_refuses_empty_or_mixed_degree_sides():
    s = config1()
    kept = P.kept_pairs(s.basis, s.D, 1e-12, 1)[0]
    with pytest.raises(ValueError, match="empty"):
        P.pack(s.basis, [])
    by_deg = {}
    for p in kept:
        by_deg.setdefault(p.degree, p)
    assert len(by_deg) >= 2
    a, b = list(by_deg.values())[:2]
    with pytest.raises(ValueError, match="degree-uniform"):
        P.pack(s.basis, [a, b])


@pytest.mark.skipif(not native.LIB_PATH.exists(), reason="library not built")
def test_entries_refuse_bad_arguments_before_any_launch():
    lib = C.CDLL(str(native.LIB_PATH))
    # null tables
    f = lib.fj_fxm_0000
    f.restype = C.c_int
    rc = f(None, C.c_int(64), None, None, None, C.c_int(0), C.c_int(0), None, None, C.c_int(0),
           C.c_double(0.0), None, C.c_longlong(0))
    assert rc == ERR_ARG
    # a quantile outside (0, 1)
    g = lib.fj_mlp_backward_quantile
    g.restype = C.c_int
    rc = g(None, C.c_int(10), C.c_int(14), None, None, None, None, None, None, C.c_float(1.5))
    assert rc == ERR_ARG
    # threads not a multiple of 64 (checked before the launch)
    side = native.FjSide(0, 0, 1, 0, 0, 0, 0, 1)
    work = (C.c_int * 4)(0, 0, 0, 1)
    sysd = (C.c_int * 16)()
    h = lib.fj_fx_0000
    h.restype = C.c_int
    rc = h(None, C.c_int(96), C.c_int(0), C.byref(side), C.byref(side), work, C.c_int(1), None,
           sysd, C.c_int(0), C.c_double(0.0), None, C.c_longlong(0))
    assert rc == ERR_THREADS
    # the loop-form K kernel: a class past f, then threads not a multiple of 64
    k = lib.fj_k_loop
    k.restype = C.c_int
    nulls = [None] * 16
    rc = k(None, C.c_int(4), C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(64), C.c_int(1),
           C.c_int(1), C.c_int(1), C.c_double(0.0), C.c_double(1.0), *nulls)
    assert rc == ERR_ARG
    rc = k(None, C.c_int(3), C.c_int(3), C.c_int(3), C.c_int(3), C.c_int(100), C.c_int(1),
           C.c_int(1), C.c_int(1), C.c_double(0.0), C.c_double(1.0), *nulls)
    assert rc == ERR_THREADS
    # the loop-form Schwarz kernel: a class past f, then a bad thread count
    q = lib.fj_schwarz_loop
    q.restype = C.c_int
    nulls = [None] * 10
    rc = q(None, C.c_int(4), C.c_int(0), C.c_int(64), C.c_int(1), C.c_int(1), C.c_int(10), *nulls)
    assert rc == ERR_ARG
    rc = q(None, C.c_int(3), C.c_int(3), C.c_int(320), C.c_int(1), C.c_int(1), C.c_int(10), *nulls)
    assert rc == ERR_THREADS


def test_k_tables_cover_f_and_follow_the_launch_order():
    """The K host tables reach f (CO/cc-pVTZ) and every launch names a bra
    table, a ket table and a D-block table that exist."""
    from felizjueves_amd import jk
    from felizjueves_amd.basis import load_fixture
    t = jk.exchange_tables(load_fixture("co_ccpvtz").basis)
    L = list(jk.exchange_launches(t))
    assert max(la for la, *_ in L) == 3 and max(ld for _, _, _, ld, _ in L) == 3
    for la, lc, lb, ld, bs in L:
        assert (la, lc) in t["bras"] and (lb, ld, bs) in t["kets"] and (lc, ld) in t["dblk"]
        # every D block a ket can name lies inside the gathered table
        dc = t["bras"][(la, lc)][("dc", ld)]
        assert dc.max() + t["kets"][(lb, ld, bs)]["pos"].max() < t["dblk"][(lc, ld)]["n"]


def test_schwarz_tables_cover_f_pairs():
    """The Schwarz host tables reach (ff) on CO/cc-pVTZ; each pair's record
    range lies inside its class's records."""
    from felizjueves_amd import jk
    from felizjueves_amd.basis import load_fixture
    t = jk.schwarz_tables(load_fixture("co_ccpvtz").basis)
    assert (2, 3) in t and (3, 3) in t
    for c in t.values():
        assert c["rec0"][0] == 0 and c["rec1"][-1] == c["n_rec"] == len(c["z"])
        assert (c["rec1"] > c["rec0"]).all() and len(c["oa"]) == c["n_pair"]
