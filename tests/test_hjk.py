"""The split J / K build on the CPU: the symmetry orbits and their
representatives (pinned to the upstream's counts and choices), the
orientation's block map against the oracle's ERIs, and the ordered pair
stream (every ordered quartet exactly once, the one-block screens), walked
with the oracle's ERIs per ordered quartet against the oracle's
every-ordered-quartet sweep and libcint."""

import numpy as np
import pytest

from felizjueves_amd import hjk
from felizjueves_amd import orbits as OB
from felizjueves_amd.basis import load_fixture
from felizjueves_amd.systems import config1
from oracle import oracle as O


def test_orbit_counts_and_representatives():
    """The upstream's counts (hgp/orbits.py docstring): 50 ordered classes
    and 13 orbits at L <= 4 with shells through d, 70 and 17 uncapped; its
    tie-breaks: (ss|ps) over (ps|ss), (ds|ps), (fs|ps), (dp|ps), (ds|pp)
    by VRR nodes against the lexicographic order."""
    assert len(OB.ordered_classes(4)) == 50 and len(OB.canonical_classes(4)) == 13
    assert len(OB.ordered_classes(4, None)) == 70 and len(OB.canonical_classes(4, None)) == 17
    assert OB.representative((1, 0, 0, 0)) == (0, 0, 1, 0)
    for rep, other in [((2, 0, 1, 0), (1, 0, 2, 0)), ((3, 0, 1, 0), (1, 0, 3, 0)),
                       ((2, 1, 1, 0), (1, 0, 2, 1)), ((2, 0, 1, 1), (1, 1, 2, 0))]:
        assert OB.representative(other) == rep
        assert OB.rank_key(rep)[:2] < OB.rank_key(other)[:2]
    # the HRR statement counts that rank (dp|ss) over (pd|ss), (ps|dp) over (sp|pd)
    assert OB.rank_key((2, 1, 0, 0))[0] == 34 and OB.rank_key((1, 2, 0, 0))[0] == 60
    assert OB.rank_key((1, 0, 2, 1))[0] == 102 and OB.rank_key((0, 1, 1, 2))[0] == 256
    assert len(OB.canonical_classes(8)) == 21


def test_orientation_maps_blocks():
    """For every ordered class of a quartet of h2o/6-31G's shells: the
    representative's block on the permuted shells, transposed by `axes`
    (and gathered by `gather`), is the ordered quartet's block."""
    fx = load_fixture("h2o_631g")
    b = fx.basis
    ls = [int(x) for x in b.shell_l]
    sh = {l: ls.index(l) for l in set(ls)}
    for cls in OB.ordered_classes(3):
        if any(x not in sh for x in cls):
            continue
        q = tuple(sh[x] for x in cls)
        o = OB.to_representative(cls)
        assert OB.apply(o.shells, cls) == o.rep
        x = tuple(q[j] for j in o.shells)
        want = O.eri(b, q, "mmd")
        rep = O.eri(b, x, "mmd")
        assert np.array_equal(rep.transpose(o.axes), want) or np.allclose(rep.transpose(o.axes), want,
                                                                          rtol=0, atol=1e-15)
        assert np.allclose(rep.ravel()[list(o.gather)], want.ravel(), rtol=0, atol=1e-15)
        assert o.perm in OB.orientations(o.rep)


def _walk_oracle(plan, basis, D):
    """The plan walked with the oracle's ERIs, one ordered quartet at a time
    in its own order, deposited per the screen (no orientation involved)."""
    n = basis.n_ao
    X = np.zeros((n, n))
    ao = basis.shell_ao
    sz = basis.shell_size
    for _, (a, b, c, d), key, skipped in plan.visits():
        if skipped:
            continue
        x = O.eri(basis, (a, b, c, d), "mmd")
        if plan.screen == "j":
            Dcd = D[ao[c]:ao[c] + sz[c], ao[d]:ao[d] + sz[d]]
            X[ao[a]:ao[a] + sz[a], ao[b]:ao[b] + sz[b]] += np.einsum("ijkl,kl->ij", x, Dcd)
        else:
            Dbd = D[ao[b]:ao[b] + sz[b], ao[d]:ao[d] + sz[d]]
            X[ao[a]:ao[a] + sz[a], ao[c]:ao[c] + sz[c]] += np.einsum("ijkl,jl->ik", x, Dbd)
    return X


@pytest.mark.parametrize("screen", ["j", "k"])
def test_plan_covers_every_ordered_quartet_once(screen):
    fx = load_fixture("h2o_631g")
    b = fx.basis
    plan = hjk.hjk_launches(b, fx.density, tau=1e-12, screen=screen)
    c = plan.counts()
    assert c["pairs"] == b.n_sh ** 2 and c["launches"] == c["batches"] ** 2
    seen = [q for _, q, _, _ in plan.visits()]
    assert len(seen) == len(set(seen)) == b.n_sh ** 4
    for L in plan.launches:
        (bi, bs), (ki, ks) = hjk.kernel_sides(L)
        # the kernel's sides, each pair read in the built order, make the
        # launch's ordered quartets exactly once
        made = set()
        for p in plan.batches[bi].pairs:
            for q in plan.batches[ki].pairs:
                x = ((p.b, p.a) if bs else (p.a, p.b)) + ((q.b, q.a) if ks else (q.a, q.b))
                made.add(tuple(x[j] for j in L.orientation.axes))
        want = {(p.a, p.b, q.a, q.b) for p in plan.batches[L.bra].pairs
                for q in plan.batches[L.ket].pairs}
        assert made == want


@pytest.mark.parametrize("screen", ["j", "k"])
def test_walk_with_oracle_eris_vs_ordered_sweep(screen):
    """tau = 0: the stream's quartets summed = the oracle's every-ordered-
    quartet sweep (C1, s,s,p,d); h2o/6-31G at tau = 1e-12 within libcint."""
    s = config1()
    plan = hjk.hjk_launches(s.basis, s.D, tau=0.0, screen=screen)
    X = _walk_oracle(plan, s.basis, s.D)
    J, K = O.fock_ordered(s.basis, s.D)
    ref = J if screen == "j" else K
    assert np.abs(X - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
    fx = load_fixture("h2o_631g")
    plan = hjk.hjk_launches(fx.basis, fx.density, tau=1e-12, screen=screen)
    X = _walk_oracle(plan, fx.basis, fx.density)
    ref = fx.arrays["vj" if screen == "j" else "vk"]
    assert np.abs(X - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


def test_sub_pairs_partition_each_pair():
    """The sub-pair packing (prim_chunk) splits each pair's primitive records
    into runs of sub_degree that concatenate to the whole pair's, and repeats
    its G, separation, shells and AOs per run."""
    s = config1()
    whole = hjk.hjk_launches(s.basis, s.D, tau=1e-12, screen="j", prim_chunk=0)
    sub = hjk.hjk_launches(s.basis, s.D, tau=1e-12, screen="j")
    assert sub.prim_chunk == hjk.PRIM_CHUNK and whole.prim_chunk == 0
    assert [hjk.sub_degree(k, 9) for k in (1, 5, 6, 7, 12, 18, 25, 36)] == [1, 5, 6, 7, 6, 9, 5, 9]
    checked = 0
    for bi, bt in enumerate(whole.batches):
        for sw in (False, True):
            a = hjk.side_records(s.basis, whole, bt, sw)
            b = hjk.side_records(s.basis, sub, sub.batches[bi], sw)
            m = a["k"] // b["k"]
            assert a["k"] == bt.degree and b["k"] == hjk.sub_degree(bt.degree, hjk.PRIM_CHUNK)
            assert np.array_equal(a["prim"], b["prim"])       # the same records, regrouped
            for key in ("G", "sep", "sh", "ao"):
                assert np.array_equal(np.repeat(a[key], m, axis=0), b[key])
            checked += m > 1
    assert checked > 0


def test_refusals():
    fx = load_fixture("h2o_631g")
    with pytest.raises(ValueError):
        hjk.hjk_launches(fx.basis, fx.density, tau=1e-12, mode="ordered")
    with pytest.raises(ValueError):
        hjk.hjk_launches(fx.basis, fx.density, tau=-1.0)
    with pytest.raises(ValueError):
        hjk.hjk_launches(fx.basis, fx.density + np.triu(np.ones_like(fx.density), 1), tau=0.0)
    co = load_fixture("co_ccpvtz")
    with pytest.raises(ValueError):
        hjk.hjk_launches(co.basis, co.density, tau=1e-12)
    with pytest.raises(ValueError):
        hjk.hjk_launches(fx.basis, fx.density, tau=1e-12, prim_chunk=-1)
