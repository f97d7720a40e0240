"""The batched workload on the GPU: C4 molecules (seeds 4 + i, s,s,p,d
shells) in ONE batch builder, graph-captured -- each molecule's G against
its own single build and the smallest against the oracle, FP64 bars --
and the learned selector on the same batch (the network picks tiers per
build on the device) against the bound rule, with the FP64 batch build as
the reference for its error."""

import numpy as np
import pytest

from felizjueves_amd import native
from felizjueves_amd.fock import FockBuilder, energy
from felizjueves_amd.systems import batch_molecule
from test_builder_oracle import oracle_build

pytestmark = pytest.mark.gpu

MOLECULES = (30, 49, 25, 7, 21, 8)   # 8, 8, 9, 11, 12 and 14 atoms


def _print(*a):
    print(*a, flush=True)


def _batch():
    return [batch_molecule(i, lmax=2) for i in MOLECULES]


def test_batch_fp64_graph_vs_single_builds_and_oracle():
    import torch
    systems = _batch()
    fb = FockBuilder.batch([(s.basis, s.D) for s in systems], tau=1e-12)
    Ds = [native.to_dev(s.D) for s in systems]
    fb.capture(Ds, n_streams=4)
    Gs = [G.cpu().numpy() for G in fb.replay()]
    torch.cuda.synchronize()
    worst = 0.0
    for s, G, Dd in zip(systems, Gs, Ds):
        one = FockBuilder(s.basis, s.D, tau=1e-12)
        G1 = one.build(Dd).cpu().numpy()
        worst = max(worst, float(np.abs(G - G1).max()))
        assert np.abs(G - G.T).max() == 0.0
    Go = oracle_build(FockBuilder(systems[0].basis, systems[0].D, tau=1e-12), systems[0].D)
    dG = float(np.abs(Gs[0] - Go).max())
    dE = abs(energy(systems[0].D, Gs[0]) - energy(systems[0].D, Go))
    _print(f"batch of {len(systems)} C4 molecules: max|dG| vs single builds {worst:.2e}; "
           f"{systems[0].name} vs oracle {dG:.2e}, |dE| {dE:.2e}")
    assert worst <= 1e-10
    assert dG <= 1e-10 and dE <= 1e-9


def test_batch_learned_selector_vs_bound_rule():
    import torch
    from pathlib import Path
    from felizjueves_amd.selector import LearnedSelector
    sel = LearnedSelector.load(Path(__file__).resolve().parent.parent / "felizjueves_amd" / "tuned"
                               / "selector")
    systems = _batch()
    sy = [(s.basis, s.D) for s in systems]
    Ds = [native.to_dev(s.D) for s in systems]
    G64 = [G.cpu().numpy() for G in FockBuilder.batch(sy, tau=1e-12).build(Ds)]
    out = {}
    for name, kw in (("bound_rule", {}), ("learned", {"selector": sel})):
        fb = FockBuilder.batch(sy, tau=1e-12, mixed=True, planner="native", **kw)
        Gs = [G.cpu().numpy() for G in fb.build(Ds)]
        torch.cuda.synchronize()
        dG = max(float(np.abs(G - R).max()) for G, R in zip(Gs, G64))
        dE = max(abs(energy(s.D, G) - energy(s.D, R)) for s, G, R in zip(systems, Gs, G64))
        out[name] = (fb.fp32_share(), dG, dE)
        _print(f"{name}: fp32 share {fb.fp32_share():.4f}, max|dG| {dG:.2e}, max|dE| {dE:.2e}")
        assert dG <= 1e-8 and dE <= 1e-6
    assert out["learned"][0] >= out["bound_rule"][0]
