"""The builder's work lists (CPU): every quartet the density-free screen
keeps (G_ab G_cd >= tau) has its primitive loop covered exactly once by the
(work item bra chunk) x (ket-list lane chunk) products, for whole tiles,
for the automatic split and for a forced fine split, in both tiers; the
ones the planning density's screen drops sit in the cold items."""

import numpy as np
import pytest

from felizjueves_amd.fock import FP32, MIXED, FockBuilder
from felizjueves_amd.systems import config1, config2


def _decode_cover(fb):
    """{(class, bra, ket): [kab x kcd] coverage counts} from the work lists,
    decoding the fields as the kernel front end does."""
    cover = {}
    for L in fb.launches:
        bra, ket = L.bra, L.ket
        for b, k0, z, k1 in L.work:
            assert z & 0xFFFF == bra.mol[b]
            kab = int(bra.deg[b])
            kb_lo, kb_n = (z >> 16) & 0xFF, (z >> 24) & 0xFF
            assert kb_lo < kab
            b1 = kab if kb_n == 0 else min(kab, kb_lo + kb_n)
            assert k1 - k0 <= L.threads
            for v in L.ket_list[k0:k1]:
                k, lo, n = v & 0x1FFFF, (v >> 17) & 0x7F, (v >> 24) & 0x7F
                assert ket.mol[k] == bra.mol[b]
                kcd = int(ket.deg[k])
                c1 = kcd if n == 0 else min(kcd, lo + n)
                key = (L.cls, int(b), int(k))
                m = cover.setdefault(key, np.zeros((kab, kcd), dtype=np.int32))
                m[kb_lo:b1, lo:c1] += 1
    return cover


def _expected(fb, density: bool = False):
    """{(class, bra, ket)} the plan must cover: G_ab G_cd >= tau over the
    kets of the bra's own system (density: also the planning density's
    six-block screen)."""
    out = set()
    classes = sorted(fb.tables)
    for i, X in enumerate(classes):
        for Y in classes[i:]:
            bra, ket = fb.tables[X], fb.tables[Y]
            for b in range(bra.n):
                lo, hi = ket.mol_range.get(int(bra.mol[b]), (0, 0))
                live = np.zeros(ket.n, dtype=bool)
                live[lo:hi] = bra.G[b] * ket.G[lo:hi] >= fb.tau
                if density:
                    live &= fb._bounds(bra, ket, b) >= fb.tau
                if X == Y:
                    live[:b] = False
                for k in np.nonzero(live)[0]:
                    out.add(((*X, *Y), b, int(k)))
    return out


def _cold(fb):
    """{(class, bra, ket)} of the cold items (the last n_cold FP64 items)."""
    out = set()
    for L in fb.launches:
        end = L.n_fp64 if L.precision == MIXED else L.work.shape[0]
        for b, k0, z, k1 in L.work[end - L.n_cold:end]:
            assert z >> 16 == 0                        # the whole bra
            for v in L.ket_list[k0:k1]:
                assert v >> 17 == 0                    # whole kets
                out.add((L.cls, int(b), int(v)))
    return out


@pytest.mark.parametrize("mixed", [False, True])
def test_cold_items_hold_exactly_the_density_screened_quartets(mixed):
    """Cold items: the quartets G_ab G_cd >= tau that the planning density's
    screen drops, not counted as work (the launch counts sum over the other
    items)."""
    s = config2(seed=11, n_atoms=10)
    fb = FockBuilder(s.basis, s.D, tau=1e-12, mixed=mixed)
    cold = _cold(fb)
    assert cold and cold == _expected(fb) - _expected(fb, density=True)
    for L in fb.launches:
        assert FockBuilder.item_quartets(L).sum() == L.n_quartets, L.cls
    assert sum(L.n_quartets for L in fb.launches) == len(_expected(fb, density=True))


@pytest.mark.parametrize("kw", [dict(), dict(lane_flops=3e4), dict(lane_flops=1e30),
                                dict(mixed=True, selector=lambda X: np.arange(len(X)) % 2 == 0)])
def test_work_lists_cover_each_primitive_quartet_once(kw):
    s = config1()
    fb = FockBuilder(s.basis, s.D, tau=1e-12, **kw)
    cover = _decode_cover(fb)
    assert set(cover) == _expected(fb)
    for key, m in cover.items():
        assert (m == 1).all(), key
    if "selector" in kw:
        # both tiers of a class pair run as one launch, each item in its tier
        assert any(L.precision == MIXED for L in fb.launches)
        split = FockBuilder(s.basis, s.D, tau=1e-12, merge_tiers=False, **kw)
        assert any(L.precision == FP32 for L in split.launches)
        assert fb.fp32_share() == split.fp32_share() > 0.0
        assert _decode_cover(split).keys() == cover.keys()


def test_small_system_is_split_finer_than_a_large_one():
    small = FockBuilder(config1().basis, config1().D, tau=1e-12)
    s = config2(seed=11, n_atoms=10)
    large = FockBuilder(s.basis, s.D, tau=1e-12)
    assert small.lane_flops_used < large.lane_flops_used
    n_lanes = sum(len(L.ket_list) for L in small.launches)
    assert n_lanes > 5 * small.n_quartets


def test_batch_work_lists_match_the_single_system_plans():
    """A batch of three systems: each system's quartets covered once, kets
    only from the bra's own system, and the same quartets as each system's
    own plan."""
    from felizjueves_amd.systems import config1, config2
    systems = [config1(seed=1), config2(seed=5, n_atoms=4), config1(seed=3)]
    fb = FockBuilder.batch([(s.basis, s.D) for s in systems], tau=1e-12)
    cover = _decode_cover(fb)
    for key, m in cover.items():
        assert (m == 1).all(), key
    assert set(cover) == _expected(fb)
    for i, s in enumerate(systems):
        one = FockBuilder(s.basis, s.D, tau=1e-12)
        mine = {(c, fb.tables[c[:2]].pairs[b].a, fb.tables[c[:2]].pairs[b].b,
                 fb.tables[c[2:]].pairs[k].a, fb.tables[c[2:]].pairs[k].b)
                for c, b, k in cover if fb.tables[c[:2]].mol[b] == i}
        theirs = {(c, one.tables[c[:2]].pairs[b].a, one.tables[c[:2]].pairs[b].b,
                   one.tables[c[2:]].pairs[k].a, one.tables[c[2:]].pairs[k].b)
                  for c, b, k in _decode_cover(one)}
        assert mine == theirs, i


def _same_plan(a, b):
    assert len(a.launches) == len(b.launches)
    for La, Lb in zip(a.launches, b.launches):
        assert (La.cls, La.precision, La.threads) == (Lb.cls, Lb.precision, Lb.threads)
        assert np.array_equal(La.work, Lb.work), La.cls
        assert np.array_equal(La.ket_list, Lb.ket_list), La.cls
        assert (La.n_quartets, La.n_prim) == (Lb.n_quartets, Lb.n_prim)
        assert abs(La.flops - Lb.flops) <= 1e-9 * La.flops


@pytest.mark.parametrize("kw", [dict(), dict(lane_flops=3e4), dict(mixed=True),
                                dict(mixed=True, fp32_bound=1e-6), dict(threads=128, lane_flops=1e5)])
def test_native_planner_equals_python_planner(kw):
    s = config1()
    _same_plan(FockBuilder(s.basis, s.D, tau=1e-12, planner="python", **kw),
               FockBuilder(s.basis, s.D, tau=1e-12, planner="native", **kw))


def test_native_planner_equals_python_planner_batch():
    systems = [config1(seed=1), config2(seed=5, n_atoms=4), config2(seed=11, n_atoms=10)]
    sy = [(s.basis, s.D) for s in systems]
    _same_plan(FockBuilder.batch(sy, tau=1e-12, planner="python"),
               FockBuilder.batch(sy, tau=1e-12, planner="native"))


def test_orient_pairs_partition_each_pair():
    """Oriented tables: each shell pair's primitive pairs appear exactly
    once over its parts, every part's records are (P, P - first centre,
    zeta, U) of its own orientation, and only pairs of separate centres are
    swapped or split."""
    from felizjueves_amd import pairs as PP
    from felizjueves_amd.fock import ORIENT_AMP_MAX, orient_pairs
    s = config2(seed=11, n_atoms=10)
    kept = PP.kept_pairs(s.basis, s.D, 1e-12, 1)[0]
    parts = orient_pairs(s.basis, kept)
    seen = {}
    for p in parts:
        key = (min(p.a, p.b), max(p.a, p.b)) if p.a != p.b else (p.a, p.b)
        first = s.basis.shell_centre[p.a]
        assert np.allclose(p.rec[:, 3:6], p.rec[:, 0:3] - first, rtol=0, atol=1e-12)
        assert (p.la, p.lb) == (int(s.basis.shell_l[p.a]), int(s.basis.shell_l[p.b]))
        seen.setdefault(key, []).append(p)
    swapped = split = 0
    for p in kept:
        key = (min(p.a, p.b), max(p.a, p.b)) if p.a != p.b else (p.a, p.b)
        got = seen[key]
        zs = np.sort(np.concatenate([q.rec[:, 6] for q in got]))
        us = np.sort(np.concatenate([q.rec[:, 7] for q in got]))
        assert np.array_equal(zs, np.sort(p.rec[:, 6])) and np.array_equal(us, np.sort(p.rec[:, 7]))
        same_centre = not np.any(s.basis.shell_centre[p.a] != s.basis.shell_centre[p.b])
        if p.a == p.b or same_centre:
            assert len(got) == 1 and (got[0].a, got[0].b) == (p.a, p.b)
        swapped += any((q.a, q.b) == (p.b, p.a) for q in got) and len(got) == 1
        split += len(got) == 2
    assert swapped > 0 and ORIENT_AMP_MAX > 0


def test_orientation_keeps_the_fock_matrix():
    """The oriented plan computes the same G as the unoriented one (C1, the
    CPU oracle's HGP launches over each plan's tables)."""
    from test_builder_oracle import oracle_build
    s = config1()
    a = oracle_build(FockBuilder(s.basis, s.D, tau=1e-12, orient=False), s.D)
    b = oracle_build(FockBuilder(s.basis, s.D, tau=1e-12), s.D)
    assert np.abs(a - b).max() < 1e-11
