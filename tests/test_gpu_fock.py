"""GPU parity: the gfx950 library against the CPU oracle and libcint.

Every test here launches the HIP library on cuda:0 and is marked `gpu`.
Bars: FP64 max|dG| <= 1e-10 and |dE| <= 1e-9 Eh against the oracle; the
FP32 tier |dE| <= 1e-6 Eh and max|dG| <= 1e-8."""

import numpy as np
import pytest

from felizjueves_amd import native
from felizjueves_amd import pairs as P
from felizjueves_amd.basis import load_fixture
from felizjueves_amd.fock import FockBuilder, energy
from felizjueves_amd.systems import config1, config2
from felizjueves_amd.walk import GpuRunner, walk_f
from helpers import OracleRunner
from oracle import oracle as O
from test_builder_oracle import oracle_build, oracle_build_mmd

pytestmark = pytest.mark.gpu

F_OFF = ("the HGP f-class kernels are switched off on the GPU: the first launch of the (ss|ff) "
         "workspace kernel (upstream-ABI walker, two-atom s,s,p,d,f system) ended in a memory aperture "
         "violation; the same generated code emulated on the CPU under AddressSanitizer "
         "(tests/test_kernel_emulation.py) stays in bounds and matches the oracle on those inputs, so "
         "the cause is not shown and the library carries HGP kernels through (dd|dd) only. The f "
         "classes now run on the loop-form kernel instead, tested by the *_loop_f tests")


def _print(*a):
    print(*a, flush=True)


def test_library_loads():
    assert native.lib().fj_version() >= 1


def test_boys_device_vs_oracle():
    rng = np.random.default_rng(7)
    T = np.concatenate([
        [0.0, 1e-14, 1e-8, 1e-4, 0.0625, 0.1, 1.0, 39.9375, 39.99, 40.0, 40.0001, 79.9, 80.0, 1e3, 1e6],
        rng.uniform(0, 45, 2000), np.exp(rng.uniform(np.log(1e-6), np.log(1e6), 2000)),
        np.arange(0, 40, 0.125) + 1.0 / 16 - 1e-12])
    for M in (0, 1, 4, 8, 12):
        got = native.boys(M, T)
        ref = np.array([O.boys(M, t) for t in T])
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
        _print("boys M", M, "max rel", rel.max())
        assert rel.max() < 2e-14, (M, T[np.argmax(rel.max(axis=1))])


def _launch_errors(system_basis, D, tau, n, mmd_for_f: bool = False):
    """Every launch of the upstream stream, device vs oracle, per class. With
    mmd_for_f the classes with an f shell (the loop-form kernel, integrals by
    McMurchie-Davidson) are checked against the oracle's MMD witness of the
    same launch (same screen and digestion): the HGP witness's contracted
    HRR loses digits on tight-diffuse f pairs."""
    plan = P.f_launches(system_basis, D, tau=tau, n=n)
    Dsh = P.density_blocks(system_basis, D)
    gpu = GpuRunner()
    worst = {}
    for L in plan.launches:
        args = P.launch_args(system_basis, plan, L, D, Dsh)
        Fo = np.zeros((system_basis.n_ao,) * 2)
        if mmd_for_f and 3 in L.cls:
            side = lambda g, sh, ao: {"G": args[g], "sh": args[sh], "ao": args[ao]}  # noqa: E731
            O.fx_launch_mmd(L.cls, side("G_ab", "bra_sh", "bra_ao"), side("G_cd", "ket_sh", "ket_ao"),
                            args["diag"], system_basis, args["tau"], Dsh, D, Fo)
        else:
            O.f_launch(L.cls, args, Fo)
        Fg = gpu.zeros(system_basis.n_ao)
        gpu(L.cls, args["n_bra"], (64, 1), gpu.put(args), Fg)
        err = np.abs(gpu.host(Fg) - Fo).max()
        worst[L.cls] = max(worst.get(L.cls, 0.0), err)
    return worst


def test_every_class_config1_launches():
    s = config1()
    worst = _launch_errors(s.basis, s.D, 0.0, 1)
    _print("per-class max |dF|:", {k: f"{v:.1e}" for k, v in sorted(worst.items())})
    assert len(worst) == 21
    assert max(worst.values()) < 1e-11


@pytest.mark.parametrize("name", ["h2o_sto3g", "h2o_631g", "h2o_ccpvdz", "h10_chain_ccpvdz"])
def test_walk_gpu_vs_libcint(name):
    fx = load_fixture(name)
    plan = P.f_launches(fx.basis, fx.density, tau=1e-12, n=3)
    w = walk_f(plan, fx.basis, fx.density, run=GpuRunner())
    ref = fx.arrays["vj"] - 0.5 * fx.arrays["vk"]
    wo = walk_f(plan, fx.basis, fx.density, run=OracleRunner())
    _print(name, "gpu-libcint", np.abs(w.F - ref).max(), "gpu-oracle", np.abs(w.F - wo.F).max())
    assert np.abs(w.F - wo.F).max() < 1e-12 * max(1.0, np.abs(ref).max())
    assert np.abs(w.F - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


def _builder_vs_oracle(s, mixed: bool):
    import torch
    fb = FockBuilder(s.basis, s.D, tau=1e-12, mixed=mixed)
    Dd = native.to_dev(s.D)
    G = fb.build(Dd)
    torch.cuda.synchronize()
    Gg = G.cpu().numpy()
    Go = oracle_build(fb if not mixed else FockBuilder(s.basis, s.D, tau=1e-12), s.D)
    dG = np.abs(Gg - Go).max()
    dE = abs(energy(s.D, Gg) - energy(s.D, Go))
    return fb, dG, dE, Gg


def test_builder_config1_fp64():
    s = config1()
    fb, dG, dE, G = _builder_vs_oracle(s, False)
    _print("c1 fp64 max|dG|", dG, "|dE|", dE, "quartets", fb.n_quartets)
    assert dG <= 1e-10 and dE <= 1e-9
    assert np.abs(G - G.T).max() == 0.0


def test_builder_config1_mixed():
    s = config1()
    fb, dG, dE, _ = _builder_vs_oracle(s, True)
    _print("c1 mixed max|dG|", dG, "|dE|", dE, "fp32 share", fb.fp32_share())
    assert dG <= 1e-8 and dE <= 1e-6


@pytest.mark.parametrize("which", ["c1", "c2-10"])
def test_builder_all_fp32_tier(which):
    """Every tile in the FP32 tier (fp32_bound above every bound), so every
    class's FP32 kernel runs, the staged ones with their own larger VRR
    groups: against the FP64 build. FP32 VRR values carry ~1e-7 relative
    error, amplified by the HRR, so the bar is 1e-3 max|G| -- far above
    that, far below the O(max|G|) error a wrong group, target or
    workspace index gives. (The mixed-precision bars are held by the
    mixed tests, where the rule keeps such tiles in FP64.)"""
    import torch
    s = config1() if which == "c1" else config2(seed=11, n_atoms=10)
    Dd = native.to_dev(s.D)
    G64 = FockBuilder(s.basis, s.D, tau=1e-12).build(Dd).clone()
    fb = FockBuilder(s.basis, s.D, tau=1e-12, mixed=True, fp32_bound=1e30)
    G32 = fb.build(Dd)
    torch.cuda.synchronize()
    assert fb.fp32_share() > 1.0 - 1e-12
    scale = G64.abs().max().item()
    err = (G32 - G64).abs().max().item()
    _print(which, "all-FP32 max|dG|", err, "max|G|", scale)
    assert err <= 1e-3 * scale


def test_builder_everything_screened_gives_zero():
    """tau above every bound: no launch at all; eager and graph builds give
    an exact zero G (F is still cleared and closed)."""
    import torch
    s = config1()
    fb = FockBuilder(s.basis, s.D, tau=1e30)
    Dd = native.to_dev(s.D)
    assert fb.launches == []
    G = fb.build(Dd).clone()
    fb.capture(Dd)
    Gr = fb.replay().clone()
    torch.cuda.synchronize()
    assert G.abs().max().item() == 0.0 and Gr.abs().max().item() == 0.0


def test_builder_single_s_function_vs_oracle():
    """One s shell (one AO, two primitives): the (ss|ss) diagonal launch of
    one pair, FP64 bars against the oracle."""
    import torch
    from felizjueves_amd.basis import Basis, Shell
    b = Basis([Shell(0, (0.1, -0.2, 0.3), (1.3, 0.4), (0.6, 0.5))])
    D = np.array([[0.7]])
    fb = FockBuilder(b, D, tau=1e-12)
    G = fb.build(native.to_dev(D))
    torch.cuda.synchronize()
    G = G.cpu().numpy()
    Go = oracle_build(fb, D)
    _print("single s G", G[0, 0], "oracle", Go[0, 0])
    assert abs(G[0, 0] - Go[0, 0]) <= 1e-12 * max(1.0, abs(Go[0, 0]))


def test_builder_repeatable():
    """Two builds of the same density agree to the atomics' reordering."""
    import torch
    s = config1()
    fb = FockBuilder(s.basis, s.D, tau=1e-12)
    Dd = native.to_dev(s.D)
    G1 = fb.build(Dd).clone()
    G2 = fb.build(Dd).clone()
    torch.cuda.synchronize()
    assert (G1 - G2).abs().max().item() < 1e-12


def test_builder_small_cluster_vs_oracle():
    """A 10-atom cluster of the working-size recipe: FP64 bars against the
    oracle's MMD-integral build (the libcint-class reference). Its shells
    pair primitives of exponent up to 1e4 with distant diffuse shells,
    where the HGP transfer cancels in FP64: the oracle's own HGP build is
    1.3e-10 from the MMD one here (and a 40-digit MMD evaluation of the
    worst quartets agrees with the MMD path to an ulp), so the HGP oracle
    is printed, not used as the bar."""
    import torch
    s = config2(seed=11, n_atoms=10)
    fb = FockBuilder(s.basis, s.D, tau=1e-12)
    G = fb.build(native.to_dev(s.D))
    torch.cuda.synchronize()
    G = G.cpu().numpy()
    Gm = oracle_build_mmd(fb, s.D)
    Gh = oracle_build(fb, s.D)
    dG = np.abs(G - Gm).max()
    dE = abs(energy(s.D, G) - energy(s.D, Gm))
    _print("c2-10 fp64 vs MMD oracle max|dG|", dG, "|dE|", dE, "vs HGP oracle", np.abs(G - Gh).max(),
           "HGP-MMD oracles", np.abs(Gh - Gm).max(), "quartets", fb.n_quartets)
    assert dG <= 1e-10 and dE <= 1e-9


def test_builder_c2_fp64_vs_reference():
    """The 40-atom working-size cluster (n_ao 440), FP64 bars at full size
    against the committed reference G (tests/golden/make_c2_reference.py:
    the oracle's HGP launches over fully oriented pairs)."""
    import torch
    from pathlib import Path
    z = np.load(Path(__file__).resolve().parent / "golden" / "c2_40atom_fp64_reference.npz",
                allow_pickle=False)
    s = config2()
    n = int(z["n_ao"])
    assert n == s.basis.n_ao and str(z["system"]) == s.name
    ref = np.zeros((n, n))
    ref[np.triu_indices(n)] = z["G_upper"]
    ref = ref + np.triu(ref, 1).T
    fb = FockBuilder(s.basis, s.D, tau=float(z["tau"]))
    G = fb.build(native.to_dev(s.D))
    torch.cuda.synchronize()
    G = G.cpu().numpy()
    dG = np.abs(G - ref).max()
    dE = abs(energy(s.D, G) - energy(s.D, ref))
    _print("c2-40 fp64 vs reference max|dG|", dG, "|dE|", dE, "shell quartets", fb.shell_quartets())
    assert dG <= 1e-10 and dE <= 1e-9
    assert np.abs(G - G.T).max() == 0.0


def _two_atom_spdf():
    """Two atoms 1.9 bohr apart, shells s, s, p, d, f each (seeded): every
    one of the 55 classes through (ff|ff) occurs."""
    from felizjueves_amd.basis import Basis
    from felizjueves_amd.systems import atom_shells, closed_shell_density
    rng = np.random.default_rng(21)
    shells = atom_shells(rng, (0.0, 0.0, 0.0), (0, 0, 1, 2, 3)) + \
        atom_shells(rng, (0.9, -1.1, 1.2), (0, 0, 1, 2, 3))
    basis = Basis(shells)
    return basis, closed_shell_density(rng, basis.n_ao, 6)


@pytest.mark.skip(reason=F_OFF)
def test_every_class_through_f_launches():
    basis, D = _two_atom_spdf()
    worst = _launch_errors(basis, D, 0.0, 1)
    _print("per-class max |dF| (f):", {k: f"{v:.1e}" for k, v in sorted(worst.items()) if 3 in k})
    assert len(worst) == 55
    assert max(worst.values()) < 1e-11 * max(1.0, np.abs(D).max())


@pytest.mark.skip(reason=F_OFF)
def test_builder_co_ccpvtz_vs_libcint():
    """CO/cc-pVTZ (f on both atoms): the native builder against libcint's
    vj - vk/2 and the oracle."""
    import torch
    fx = load_fixture("co_ccpvtz")
    fb = FockBuilder(fx.basis, fx.density, tau=1e-12)
    G = fb.build(native.to_dev(fx.density))
    torch.cuda.synchronize()
    G = G.cpu().numpy()
    ref = fx.arrays["vj"] - 0.5 * fx.arrays["vk"]
    Go = oracle_build(fb, fx.density)
    _print("co_ccpvtz gpu-libcint", np.abs(G - ref).max(), "gpu-oracle", np.abs(G - Go).max())
    assert np.abs(G - Go).max() < 1e-10
    assert np.abs(G - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())


def test_every_class_through_f_launches_loop_f():
    """Every walker launch of the two-atom s,s,p,d,f system through the
    upstream-ABI entries, the f classes on the loop-form kernel: each
    launch's F against the oracle's witness (MMD for the f classes)."""
    basis, D = _two_atom_spdf()
    worst = _launch_errors(basis, D, 0.0, 1, mmd_for_f=True)
    _print("per-class max |dF| (f, loop form):", {k: f"{v:.1e}" for k, v in sorted(worst.items()) if 3 in k})
    assert len(worst) == 55
    assert max(worst.values()) < 1e-11 * max(1.0, np.abs(D).max())


def test_builder_spdf_vs_oracle_loop_f():
    """The native builder on the two-atom s,s,p,d,f system (oriented pairs,
    every class through (ff|ff); the f classes on the loop-form kernel):
    FP64 bars against the oracle's MMD build."""
    import torch
    basis, D = _two_atom_spdf()
    fb = FockBuilder(basis, D, tau=1e-12)
    assert any(3 in L.cls for L in fb.launches)
    G = fb.build(native.to_dev(D))
    torch.cuda.synchronize()
    Gg = G.cpu().numpy()
    Go = oracle_build_mmd(fb, D)
    dG = np.abs(Gg - Go).max()
    dE = abs(energy(D, Gg) - energy(D, Go))
    _print("spdf (loop f) max|dG|", dG, "|dE|", dE, "launches", len(fb.launches))
    assert dG <= 1e-10 and dE <= 1e-9


def test_builder_co_ccpvtz_vs_libcint_loop_f():
    """CO/cc-pVTZ (f on both atoms): the native builder, f classes on the
    loop-form kernel, against libcint's vj - vk/2 and the oracle; the mixed
    build within the mixed bars."""
    import torch
    fx = load_fixture("co_ccpvtz")
    fb = FockBuilder(fx.basis, fx.density, tau=1e-12)
    G = fb.build(native.to_dev(fx.density))
    torch.cuda.synchronize()
    G = G.cpu().numpy()
    ref = fx.arrays["vj"] - 0.5 * fx.arrays["vk"]
    Go = oracle_build_mmd(fb, fx.density)
    fm = FockBuilder(fx.basis, fx.density, tau=1e-12, mixed=True)
    Gm = fm.build(native.to_dev(fx.density))
    torch.cuda.synchronize()
    Gm = Gm.cpu().numpy()
    dEm = abs(energy(fx.density, Gm) - energy(fx.density, G))
    _print("co_ccpvtz (loop f) gpu-libcint", np.abs(G - ref).max(), "gpu-oracle", np.abs(G - Go).max(),
           "mixed max|dG|", np.abs(Gm - G).max(), "|dE|", dEm)
    assert np.abs(G - Go).max() < 1e-10
    assert np.abs(G - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())
    assert np.abs(Gm - G).max() <= 1e-8 and dEm <= 1e-6


def test_batch_and_graph_with_f_loop_f():
    """Two s,s,p,d,f systems in one batch builder (multi-system work items on
    the loop-form f kernel): each system's G equals its own single build;
    then the batch captured in a graph and replayed three times, with eager
    builds and synchronisations between the replays, equals the eager
    build every time."""
    import torch
    from felizjueves_amd.basis import Basis
    from felizjueves_amd.systems import atom_shells, closed_shell_density
    systems = []
    for seed in (31, 32):
        rng = np.random.default_rng(seed)
        b = Basis(atom_shells(rng, (0.0, 0.0, 0.0), (0, 1, 3)) +
                  atom_shells(rng, (1.1, 0.4, -0.9), (0, 2, 3)))
        systems.append((b, closed_shell_density(rng, b.n_ao, 4)))
    fb = FockBuilder.batch(systems, tau=1e-12)
    assert any(3 in L.cls for L in fb.launches)
    Ds = [native.to_dev(D) for _, D in systems]
    eager = [G.clone() for G in fb.build(Ds)]
    torch.cuda.synchronize()
    for i, (b, D) in enumerate(systems):
        one = FockBuilder(b, D, tau=1e-12).build(Ds[i])
        torch.cuda.synchronize()
        d = (one - eager[i]).abs().max().item()
        _print("batch with f, system", i, "vs alone", d)
        assert d <= 1e-12
    fb.capture(Ds, n_streams=2)
    for rep in range(3):
        Gs = fb.replay()
        torch.cuda.synchronize()
        for i in range(len(systems)):
            assert (Gs[i] - eager[i]).abs().max().item() <= 1e-12, (rep, i)
        fb.build(Ds)
        torch.cuda.synchronize()


def test_graph_replays_match_eager():
    """A captured build replayed several times, with eager builds and
    synchronisations between the replays, equals the eager build every
    time (no memset node is recorded; F is cleared by a fill kernel)."""
    import torch
    s = config1()
    fb = FockBuilder(s.basis, s.D, tau=1e-12)
    Dd = native.to_dev(s.D)
    eager = fb.build(Dd).clone()
    torch.cuda.synchronize()
    fb.capture(Dd)
    for i in range(4):
        G = fb.replay().clone()
        torch.cuda.synchronize()
        assert (G - eager).abs().max().item() < 1e-12, i
        e2 = fb.build(Dd).clone()          # eager work between replays
        torch.cuda.synchronize()
        assert (e2 - eager).abs().max().item() < 1e-12


@pytest.mark.parametrize("lane_flops", [3e4, 1e30])
def test_builder_primitive_split_vs_oracle(lane_flops):
    """Quartets spread over lanes (ket primitive chunks) and work items
    (bra primitive chunks), and whole quartets per lane: FP64 bars."""
    import torch
    s = config1()
    fb = FockBuilder(s.basis, s.D, tau=1e-12, lane_flops=lane_flops)
    G = fb.build(native.to_dev(s.D))
    torch.cuda.synchronize()
    Gg = G.cpu().numpy()
    Go = oracle_build(fb, s.D)
    dG = np.abs(Gg - Go).max()
    dE = abs(energy(s.D, Gg) - energy(s.D, Go))
    _print("split", lane_flops, "lanes", sum(len(L.ket_list) for L in fb.launches), "max|dG|", dG, "|dE|", dE)
    assert dG <= 1e-10 and dE <= 1e-9


def test_multistream_build_and_graph_match_single_stream():
    import torch
    s = config1()
    fb = FockBuilder(s.basis, s.D, tau=1e-12)
    Dd = native.to_dev(s.D)
    one = fb.build(Dd).clone()
    four = fb.build(Dd, n_streams=4).clone()
    torch.cuda.synchronize()
    assert (one - four).abs().max().item() < 1e-12
    fb.capture(Dd, n_streams=4)
    for i in range(3):
        G = fb.replay().clone()
        torch.cuda.synchronize()
        assert (G - one).abs().max().item() < 1e-12, i
        fb.build(Dd)
        torch.cuda.synchronize()


@pytest.mark.skip(reason=F_OFF)
def test_builder_spdf_vs_oracle():
    """The native builder on the two-atom s,s,p,d,f system (every class
    through (ff|ff), workspace kernels): FP64 bars."""
    import torch
    basis, D = _two_atom_spdf()
    fb = FockBuilder(basis, D, tau=1e-12)
    G = fb.build(native.to_dev(D))
    torch.cuda.synchronize()
    Gg = G.cpu().numpy()
    Go = oracle_build(fb, D)
    dG = np.abs(Gg - Go).max()
    dE = abs(energy(D, Gg) - energy(D, Go))
    _print("spdf max|dG|", dG, "|dE|", dE, "launches", len(fb.launches))
    assert dG <= 1e-10 and dE <= 1e-9


def test_batch_builder_vs_oracle():
    """Three systems in one batch builder: each system's G equals the
    oracle's for that system alone."""
    import torch
    systems = [config1(seed=1), config2(seed=5, n_atoms=4), config1(seed=3)]
    fb = FockBuilder.batch([(s.basis, s.D) for s in systems], tau=1e-12)
    Gs = fb.build([native.to_dev(s.D) for s in systems])
    torch.cuda.synchronize()
    for i, s in enumerate(systems):
        one = FockBuilder(s.basis, s.D, tau=1e-12)
        Go = oracle_build(one, s.D)
        Gg = Gs[i].cpu().numpy()
        dG = np.abs(Gg - Go).max()
        _print("batch system", i, "max|dG|", dG)
        assert dG <= 1e-10
