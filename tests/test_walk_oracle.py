"""The pair stream + walker with the oracle as the launcher, against the
libcint fixtures and the every-ordered-quartet sweep (CPU)."""

import numpy as np
import pytest

from felizjueves_amd import pairs as P
from felizjueves_amd.basis import FOCK_FIXTURES, load_fixture
from felizjueves_amd.walk import walk_f
from helpers import OracleRunner


@pytest.mark.parametrize("name", FOCK_FIXTURES)
@pytest.mark.parametrize("tau,n", [(0.0, 1), (1e-12, 3)])
def test_walk_vs_libcint(name, tau, n):
    fx = load_fixture(name)
    D = fx.density
    plan = P.f_launches(fx.basis, D, tau=tau, n=n)
    w = walk_f(plan, fx.basis, D, run=OracleRunner())
    G_ref = fx.arrays["vj"] - 0.5 * fx.arrays["vk"]
    err = np.abs(w.F - G_ref).max()
    # libcint vs the oracle at 1e-12 relative to the G scale; screening at
    # tau moves no ERI by tau or more
    assert err < 1e-11 * max(1.0, np.abs(G_ref).max()) + 50 * tau * np.abs(D).sum()
    assert np.allclose(w.F, w.F.T, atol=1e-13, rtol=0)


def test_walk_vs_ordered_sweep(oracle):
    fx = load_fixture("h2o_631g")
    D = fx.density
    plan = P.f_launches(fx.basis, D, tau=0.0, n=1)
    w = walk_f(plan, fx.basis, D, run=OracleRunner())
    J, K = oracle.fock_ordered(fx.basis, D)
    assert np.abs(w.F - (J - 0.5 * K)).max() < 1e-12


def test_sub_pair_packing_matches_whole_pairs():
    """The walker's default sub-pair packing (pairs.pack with prim_chunk)
    against the upstream's whole-pair packing, both through the oracle's
    launch on the three-atom s,s,p,d system (contractions up to 6 x 6, so
    diagonal launches of split pairs occur): the same G to rounding, and the
    sub-pairs repeat each pair's G, shells and AOs over its primitive runs."""
    from felizjueves_amd.systems import config1
    from felizjueves_amd.walk import PRIM_CHUNK
    s = config1()
    plan = P.f_launches(s.basis, s.D, tau=1e-12, n=3)
    w0 = walk_f(plan, s.basis, s.D, run=OracleRunner(), prim_chunk=0)
    w2 = walk_f(plan, s.basis, s.D, run=OracleRunner(), prim_chunk=PRIM_CHUNK)
    assert w0.n_quartets_launched == w2.n_quartets_launched
    assert np.abs(w0.F - w2.F).max() <= 1e-12 * np.abs(w0.F).max()
    b = max(plan.batches, key=lambda x: x.degree)
    whole, sub = P.pack(s.basis, b.pairs), P.pack(s.basis, b.pairs, PRIM_CHUNK)
    m = whole.degree // sub.degree
    assert m > 1 and sub.n == whole.n * m and np.array_equal(whole.prim, sub.prim)
    assert np.array_equal(np.repeat(whole.G, m), sub.G)
    assert np.array_equal(np.repeat(whole.sh.reshape(-1, 2), m, axis=0).ravel(), sub.sh)
