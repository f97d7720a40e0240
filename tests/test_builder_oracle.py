"""The native builder's plan (ragged pair tables, tile work lists), evaluated
by the CPU oracle, against libcint and the upstream-stream walk (CPU)."""

import numpy as np
import pytest

from felizjueves_amd import pairs as P
from felizjueves_amd.basis import load_fixture
from felizjueves_amd.fock import FockBuilder
from felizjueves_amd.systems import config1
from felizjueves_amd.walk import walk_f
from helpers import OracleRunner
from oracle import oracle as O


def builder_tables_host(fb: FockBuilder):
    b = fb.basis
    out = {}
    for c, t in fb.tables.items():
        out[c] = {"prim": np.concatenate([p.rec for p in t.pairs]),
                  "off": np.concatenate([[0], np.cumsum(t.deg)]).astype(np.int32),
                  "sep": np.array([b.shell_centre[p.a] - b.shell_centre[p.b] for p in t.pairs]),
                  "G": t.G.copy(), "sh": t.sh.astype(np.int32),
                  "ao": np.array([(b.shell_ao[p.a], b.shell_ao[p.b]) for p in t.pairs], dtype=np.int32)}
    return out


def oracle_build(fb: FockBuilder, D: np.ndarray) -> np.ndarray:
    """What the builder's launches compute, on the CPU: every class pair of
    its tables through the oracle's ragged launch, then F + F^T. The
    quartets are the builder's: the density-free pre-launch screen
    (G_ab G_cd < tau, launch_f.py:448-466) and the six-block density screen
    with D (kernel_f.py:150-182)."""
    b = fb.basis
    host = builder_tables_host(fb)
    Dsh = P.density_blocks(b, D)
    F = np.zeros((b.n_ao, b.n_ao))
    classes = sorted(host)
    for i, X in enumerate(classes):
        for Y in classes[i:]:
            O.fx_launch((*X, *Y), host[X], host[Y], int(X == Y), b.n_sh, b.n_ao, fb.tau, Dsh, D, F,
                        pre=True)
    return F + F.T


def oracle_build_mmd(fb: FockBuilder, D: np.ndarray) -> np.ndarray:
    """oracle_build with every quartet's integrals by MMD over the shells'
    full primitive lists (no HRR): the accurate FP64 reference where HGP's
    transfer cancels (very tight primitives paired with distant diffuse
    shells). The integrals are per shell quartet, so it takes the
    unoriented tables (one entry per shell pair, whatever fb's orientation)."""
    fb = FockBuilder(fb.basis, D, tau=fb.tau, orient=False, merge_tiers=False)
    b = fb.basis
    host = builder_tables_host(fb)
    Dsh = P.density_blocks(b, D)
    F = np.zeros((b.n_ao, b.n_ao))
    classes = sorted(host)
    for i, X in enumerate(classes):
        for Y in classes[i:]:
            O.fx_launch_mmd((*X, *Y), host[X], host[Y], int(X == Y), b, fb.tau, Dsh, D, F, pre=True)
    return F + F.T


@pytest.mark.parametrize("name", ["h2o_sto3g", "h2o_ccpvdz", "h10_chain_ccpvdz"])
def test_builder_plan_mmd_vs_libcint(name):
    fx = load_fixture(name)
    fb = FockBuilder(fx.basis, fx.density, tau=1e-12)
    G = oracle_build_mmd(fb, fx.density)
    ref = fx.arrays["vj"] - 0.5 * fx.arrays["vk"]
    assert np.abs(G - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


def test_builder_plan_mmd_vs_hgp_config1():
    """On the three-atom system the two integral paths agree closely (the
    cancellation that separates them needs exponent ratios and distances
    of the working-size cluster)."""
    s = config1()
    fb = FockBuilder(s.basis, s.D, tau=1e-12)
    assert np.abs(oracle_build_mmd(fb, s.D) - oracle_build(fb, s.D)).max() < 1e-11


@pytest.mark.parametrize("name", ["h2o_sto3g", "h2o_ccpvdz", "h10_chain_ccpvdz"])
def test_builder_plan_vs_libcint(name):
    fx = load_fixture(name)
    fb = FockBuilder(fx.basis, fx.density, tau=1e-12)
    G = oracle_build(fb, fx.density)
    ref = fx.arrays["vj"] - 0.5 * fx.arrays["vk"]
    assert np.abs(G - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


def test_builder_plan_vs_walk_config1():
    s = config1()
    fb = FockBuilder(s.basis, s.D, tau=1e-12)
    G = oracle_build(fb, s.D)
    plan = P.f_launches(s.basis, s.D, tau=1e-12, n=3)
    w = walk_f(plan, s.basis, s.D, run=OracleRunner())
    assert np.abs(G - w.F).max() < 1e-11
    # the plan's counts: every computed quartet is a unique quartet
    n_pairs = s.basis.n_sh * (s.basis.n_sh + 1) // 2
    assert 0 < fb.shell_quartets() <= n_pairs * (n_pairs + 1) // 2
