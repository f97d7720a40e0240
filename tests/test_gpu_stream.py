"""The fused-F pair stream on the device (fj_stream_plan / fj_stream_launches,
stream.f_launches_device) against the host stream (pairs.f_launches, the
upstream's f_launches restated): the same deleted pairs, the same kept
primitive pairs per pair, the same bins with the same pairs (in the same
order wherever the keys differ), the same launches and pre-launch screens,
the bounds within rounding; and the walk of the device plan on the GPU
against libcint."""

import numpy as np
import pytest

from felizjueves_amd import pairs as P
from felizjueves_amd.basis import load_fixture
from felizjueves_amd.stream import f_launches_device
from felizjueves_amd.systems import config1, config2
from felizjueves_amd.walk import GpuRunner, walk_f

pytestmark = pytest.mark.gpu


def _systems():
    out = [(n, load_fixture(n).basis, load_fixture(n).density)
           for n in ("h2o_sto3g", "h2o_631g", "h2o_ccpvdz", "h10_chain_ccpvdz", "co_ccpvtz")]
    for s in (config1(), config2(seed=11, n_atoms=10)):
        out.append((s.name, s.basis, s.D))
    return out


@pytest.mark.parametrize("case", _systems(), ids=lambda c: c[0])
def test_device_stream_equals_host(case):
    name, basis, D = case
    host = P.f_launches(basis, D, tau=1e-12, n=3)
    dev, t = f_launches_device(basis, D, tau=1e-12, n=3)
    assert sorted(dev.deleted) == sorted(host.deleted)
    assert abs(dev.g_max - host.g_max) <= 1e-14 * host.g_max
    assert abs(dev.eps - host.eps) <= 1e-14 * host.eps
    assert len(dev.batches) == len(host.batches)
    worst_g = 0.0
    for bd, bh in zip(dev.batches, host.batches):
        assert (bd.la, bd.lb, bd.degree, bd.mag) == (bh.la, bh.lb, bh.degree, bh.mag)
        assert {(p.a, p.b) for p in bd.pairs} == {(p.a, p.b) for p in bh.pairs}
        for pd, ph in zip(bd.pairs, bh.pairs):
            if (pd.a, pd.b) != (ph.a, ph.b):   # only where the keys tie to rounding
                assert abs(pd.key - ph.key) <= 1e-13 * max(pd.key, ph.key)
        hp = {(p.a, p.b): p for p in bh.pairs}
        for p in bd.pairs:
            q = hp[(p.a, p.b)]
            assert np.array_equal(p.prims, q.prims), (p.a, p.b)
            worst_g = max(worst_g, abs(p.G - q.G) / q.G)
            assert np.allclose(p.rec, q.rec, rtol=0, atol=0)
    assert [(L.bra, L.ket) for L in dev.launches] == [(L.bra, L.ket) for L in host.launches]
    assert sorted(dev.screened) == sorted(host.screened)
    print(name, "pairs", sum(len(b.pairs) for b in dev.batches), "batches", len(dev.batches),
          "launches", len(dev.launches), "worst rel dG", worst_g, "device plan s", t["device_s"],
          flush=True)
    assert worst_g <= 1e-14


@pytest.mark.parametrize("name", ["h2o_631g", "h2o_ccpvdz", "h10_chain_ccpvdz"])
def test_walk_of_device_plan_vs_libcint(name):
    fx = load_fixture(name)
    plan, _ = f_launches_device(fx.basis, fx.density, tau=1e-12, n=3)
    w = walk_f(plan, fx.basis, fx.density, run=GpuRunner())
    ref = fx.arrays["vj"] - 0.5 * fx.arrays["vk"]
    err = np.abs(w.F - ref).max()
    print(name, "walk of the device plan: max |G - libcint|", err, flush=True)
    assert err < 1e-11 * max(1.0, np.abs(ref).max())
