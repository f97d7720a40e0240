"""A plan made with one density, built with another (the SCF use): the plan
covers every quartet the pair bounds keep (G_ab G_cd >= tau), the kernels
screen each quartet with the density of the build, and a mixed build
chooses its tiers per build on the device. Planned with a density 1e-4
times the built one, many quartets are cold (dropped by the planning
density's screen) and many tiles look FP32-safe; built with the full
density they must run, in FP64 where the error rule says so."""

import numpy as np
import pytest

from felizjueves_amd import native
from felizjueves_amd.fock import MIXED, FockBuilder, energy
from felizjueves_amd.systems import config2
from test_builder_oracle import oracle_build

pytestmark = pytest.mark.gpu

SCALE = 1e-4


def _print(*a):
    print(*a, flush=True)


def _system():
    return config2(seed=11, n_atoms=10)


def _reference(s, D):
    return oracle_build(FockBuilder(s.basis, D, tau=1e-12), D)


def test_plan_with_one_density_build_with_another_fp64():
    import torch
    s = _system()
    D1, D2 = SCALE * s.D, s.D
    fb = FockBuilder(s.basis, D1, tau=1e-12)
    full = FockBuilder(s.basis, D2, tau=1e-12)
    # the planning density drops quartets the built one needs
    assert full.n_quartets > fb.n_quartets
    assert sum(L.n_cold for L in fb.launches) > 0
    Dd = native.to_dev(D1)
    fb.capture(Dd)
    Dd.copy_(native.to_dev(D2))
    Gr = fb.replay().clone()
    G = fb.build(Dd).clone()
    torch.cuda.synchronize()
    Go = _reference(s, D2)
    for name, Gg in (("eager", G), ("graph", Gr)):
        Gg = Gg.cpu().numpy()
        dG = np.abs(Gg - Go).max()
        dE = abs(energy(D2, Gg) - energy(D2, Go))
        _print(f"planned with {SCALE} D, built with D ({name}): max|dG| {dG:.2e} |dE| {dE:.2e}; "
               f"quartets planned {fb.n_quartets}, needed {full.n_quartets}")
        assert dG <= 1e-10 and dE <= 1e-9


@pytest.mark.parametrize("rule", ["amplified", "literal"])
def test_plan_with_one_density_build_with_another_mixed(rule):
    """The tiers planned for the small density would put most work in FP32;
    the device re-derives them from the built density, so the mixed bars
    hold against the oracle on it."""
    import torch
    s = _system()
    D1, D2 = SCALE * s.D, s.D
    kw = dict(fp32_bound=1e-6) if rule == "literal" else {}
    fb = FockBuilder(s.basis, D1, tau=1e-12, mixed=True, **kw)
    assert all(L.precision == MIXED for L in fb.launches)
    planned = fb.fp32_share()
    Dd = native.to_dev(D2)
    G = fb.build(Dd)
    torch.cuda.synchronize()
    built = fb.fp32_share()
    Gg = G.cpu().numpy()
    Go = _reference(s, D2)
    dG = np.abs(Gg - Go).max()
    dE = abs(energy(D2, Gg) - energy(D2, Go))
    # the same rule planned with the built density gives the same tiers
    same = FockBuilder(s.basis, D2, tau=1e-12, mixed=True, **kw)
    _print(f"{rule}: fp32 share planned {planned:.3f}, built {built:.3f}, "
           f"planned with D {same.fp32_share():.3f}; max|dG| {dG:.2e} |dE| {dE:.2e}")
    assert built < planned
    assert dG <= 1e-8 and dE <= 1e-6


@pytest.mark.parametrize("rule", ["amplified", "literal"])
def test_device_tiers_equal_the_plan_under_the_planning_density(rule):
    """Built with its planning density, the device's per-item tiers are the
    plan's (the bound rule per ket at plan time, per item on the device:
    the plan never mixes tiers within an item)."""
    import torch
    s = _system()
    kw = dict(fp32_bound=1e-7) if rule == "literal" else {}
    fb = FockBuilder(s.basis, s.D, tau=1e-12, mixed=True, **kw)
    planned = fb.fp32_share()
    fb.build(native.to_dev(s.D))
    torch.cuda.synchronize()
    n_diff = 0
    for L in fb.launches:
        t = L.tier_dev.cpu().numpy()
        want = (np.arange(len(t)) >= L.n_fp64).astype(np.int32)
        hot = np.ones(len(t), dtype=bool)
        hot[L.n_fp64 - L.n_cold:L.n_fp64] = False      # cold items compute nothing
        n_diff += int((t[hot] != want[hot]).sum())
    _print(f"{rule}: fp32 share {planned:.4f}, items whose device tier differs {n_diff}")
    assert planned > 0.0 and n_diff == 0
    assert fb.fp32_share() == planned
