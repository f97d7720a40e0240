"""Benchmark: Fock builds/s and quartets/s on one MI355X, mixed vs FP64.

Headline workload: the 40-atom random cluster (s,s,p,d per atom, n_ao 440,
tau 1e-12, seed 2); one step = one closed two-electron Fock build
G = J - K/2 from device-resident pair tables and density. `value` is the
production build: the autotuned plan (per-class lanes and primitive-loop
split, launches dealt to several streams), captured in a HIP graph and
replayed, mixed precision (FP32 tier for the tiles the amplified bound rule
and, when a trained selector is present, the learned selector send there).
Its accuracy against the plain FP64 build rides along, with the plain FP64
build's own rate (eager, one stream, default plan).

Also reported:
* "c3": the mixed configuration on the three-atom s,s,p,d system -- plain
  FP64 build vs the production build (tuned, streams, graph, mixed,
  learned selector), |dE|, max|dG|, FP32 share, the J and K mixed levers'
  max|dJ| and max|dK| (FP32 Hermite towers in the MMD J and K engines) and
  their G = J - K/2 against the FP64 build, and the literal bound rule
  (FP32 iff Q_ab Q_cd max|D| < 1e-6) for comparison;
* "c5": learned selector vs the bound rule on held-out systems;
* "gradient": the fused two-electron nuclear gradient (dE_2e/dR, every
  unique quartet, no screen) on the three-atom s,s,p,d system and a
  10-atom cluster: gradients/s, its kernel's roofline (HIP events around
  the merged per-class launches), max|dG| against the oracle and the
  oracle's own time on the host cores for the three-atom system;
* "split_jk": the split HGP J and K kernels (per orbit representative) on
  the three-atom system (`--workload split-df` adds the 10-atom cluster),
  ms per J / K build, their roofline, parity and the oracle's ordered sweep
  beside them;
* "density_fitting": the metric and three-center integrals on the upstream
  fixtures (against libcint) and on the 10-atom cluster with a synthetic
  aux set;
* "roofline" of the dominant kernel (HIP events around its eager launches
  on the stream it runs on) and "cpu_baseline" (the oracle on the host
  cores, a bounded sample).

Prints ONE JSON line on stdout; progress goes to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("quartets/s & Fock builds/s, mixed vs FP64, |ΔE|≤1e-6 Eh, 1xMI355X")
FP64_PEAK_TFLOPS = 78.6
TUNED = ROOT / "felizjueves_amd" / "tuned"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def timed(fn, steps, warmup):
    """Seconds for exactly `steps` calls of fn after `warmup` untimed ones,
    bracketed by device synchronisations."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t


def launch_times(fb, Dd, reps=1):
    """Per-launch durations (ms) of eager one-stream builds, HIP events on
    the stream the launches run on (torch's current stream)."""
    import torch
    from felizjueves_amd import native
    sys_ = fb._sys([Dd])
    ws = fb._workspace(0)
    out = np.zeros((reps, len(fb.launches)))
    for r in range(reps):
        fb.F.zero_()
        native.density_blocks(fb.shell_ao[0], fb.shell_size[0], Dd, fb.Dsh)
        evs = []
        for L in fb.launches:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fb.launch(L, sys_, None, ws)
            e1.record()
            evs.append((e0, e1))
        native.close(fb.F)
        torch.cuda.synchronize()
        out[r] = [a.elapsed_time(b) for a, b in evs]
    return out


def load_tuning(name, mixed: bool = False):
    """(builder keywords, stream count) of a tuned configuration; with
    `mixed`, also the class pairs where the FP32 tier was measured to pay."""
    from felizjueves_amd import autotune
    p = TUNED / name
    if not p.exists():
        return {}, 1
    cfg = autotune.load(p)
    kw = autotune.builder_kwargs(cfg)
    if mixed:
        kw.update(autotune.mixed_kwargs(cfg))
    return kw, int(cfg.get("n_streams", 1))


def load_selector(path):
    from felizjueves_amd.selector import LearnedSelector
    p = Path(path)
    if not p.with_suffix(".net.npz").exists():
        return None
    return LearnedSelector.load(p)


def cpu_baseline(fb, D, G_gpu=None):
    """The CPU oracle timed on a whole build of the same workload: every
    class pair of the builder's tables through the oracle's HGP launch
    (OpenMP over the host cores; the same quartets, screen, weights and
    digestion as the device plan), then F + F^T. The CPU G is checked
    against the GPU's FP64 G when that is given."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle import oracle as O
    from test_builder_oracle import oracle_build
    O.lib()
    t = time.perf_counter()
    G = oracle_build(fb, D)
    t = time.perf_counter() - t
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    out = {
        "value": 1.0 / t, "unit": "Fock builds/s", "cores": O.threads(), "kind": "port",
        "sample": (f"one whole build of the workload ({fb.basis.n_ao} AOs, {fb.n_quartets} computed "
                   f"quartets, {fb.flops / 1e9:.1f} GFLOP): every class pair through the oracle's HGP "
                   f"launch, {t:.1f} s"),
        "cpu_model": model, "seconds_per_build": t, "gflops": fb.flops / t / 1e9,
        "quartets_per_s": fb.n_quartets / t,
    }
    if G_gpu is not None:
        out["max_dG_vs_gpu_fp64"] = float(np.abs(G - G_gpu).max())
    return out


def stream_block(s) -> dict:
    """The upstream pair stream (f_launches, eq. 29 with n = 3) of the
    workload planned on the host (pairs.f_launches) and on the device
    (stream.f_launches_device: bounds, prune, keys, bins, sort and
    launches by the library's kernels): seconds per plan and whether the
    two plans agree. An SCF caller of the walker re-plans once per density;
    the native builder's plan does not depend on the density at all."""
    from felizjueves_amd import pairs as P
    from felizjueves_amd.stream import DeviceStream, f_launches_device
    t = time.perf_counter()
    host = P.f_launches(s.basis, s.D, tau=1e-12, n=3)
    t_host = time.perf_counter() - t
    ds = DeviceStream(s.basis)
    f_launches_device(s.basis, s.D, tau=1e-12, n=3, ds=ds)      # warm
    dev, tm = f_launches_device(s.basis, s.D, tau=1e-12, n=3, ds=ds)
    same = ([(L.bra, L.ket) for L in dev.launches] == [(L.bra, L.ket) for L in host.launches]
            and [(b.la, b.lb, b.degree, b.mag, len(b.pairs)) for b in dev.batches]
            == [(b.la, b.lb, b.degree, b.mag, len(b.pairs)) for b in host.batches])
    out = {"workload": s.name, "host_plan_s": t_host, "device_plan_s": tm["device_s"],
           "device_plan_to_fplan_s": tm["host_s"], "batches": len(dev.batches),
           "launches": len(dev.launches), "same_plan_as_host": bool(same)}
    log("stream", json.dumps(out))
    return out


GRADIENT_STREAMS = 4   # the gradient's class launches over this many streams


def gradient_block(steps: int, cpu: bool = True) -> dict:
    """The fused two-electron gradient (gradient.py, fj_gf_jobs: one merged
    launch per class) on the three-atom s,s,p,d system and the 10-atom
    cluster. Per system: seconds per gradient (HIP events on the stream
    the launches go to, around whole gradients), the loop form's FLOPs
    (gradient.flops) over that time against the FP64 peak, primitive
    quartets/s; on the three-atom system max|dG| against the oracle
    (oracle_grad_ordered) and, unless `cpu` is off, the oracle's time (the
    whole gradient on the host cores)."""
    import torch
    from felizjueves_amd import gradient as GR
    from felizjueves_amd import native
    from felizjueves_amd.systems import config1, config2
    out = {}
    for key, s in (("c1_3atom", config1()), ("cluster_10atom", config2(n_atoms=10))):
        at = GR.shell_atoms(s.basis, s.coords)
        st = GR.gf_launches(s.basis, at)
        run = GR.GpuGfRunner(st, len(s.coords), s.basis.n_ao)
        Dd = native.to_dev(s.D)
        stream = torch.cuda.current_stream()
        run.run(Dd, stream, n_streams=GRADIENT_STREAMS)
        torch.cuda.synchronize()
        n = steps if key == "c1_3atom" else 3
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            run.run(Dd, stream, n_streams=GRADIENT_STREAMS)
        e1.record(stream)
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) * 1e-3 / n
        fl = GR.flops(st)
        r = {"workload": s.name, "n_ao": s.basis.n_ao, "quartets": GR.quartets(st),
             "prim_quartets": GR.prim_quartets(st), "classes": len(run.classes),
             "stream_launches": len(st.launches), "seconds_per_gradient": t,
             "gradients_per_s": 1.0 / t, "prim_quartets_per_s": GR.prim_quartets(st) / t,
             "roofline": {"bound": "fp64_valu", "achieved": fl / t / 1e12, "peak": FP64_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": fl / t / 1e12 / FP64_PEAK_TFLOPS,
                          "gflop_per_gradient": fl / 1e9,
                          "kernel": "fj_gfjobs_k (every class; its launches back to back)"}}
        if key == "c1_3atom":
            G = run.G.cpu().numpy()
            sys.path.insert(0, str(ROOT / "tests"))
            from oracle import oracle as O
            tc = time.perf_counter()
            Go = O.grad_ordered(s.basis, at, len(s.coords), s.D)
            tc = time.perf_counter() - tc
            r["max_dG_vs_oracle"] = float(np.abs(G - Go).max())
            r["max_abs_G"] = float(np.abs(Go).max())
            if cpu:
                r["cpu_baseline"] = {"value": 1.0 / tc, "unit": "gradients/s", "cores": O.threads(),
                                     "kind": "port",
                                     "sample": ("one whole gradient of the three-atom system by the "
                                                f"oracle's ordered sweep, {tc:.1f} s")}
        log("gradient", key, json.dumps(r))
        out[key] = r
    return out


def split_jk_block(cpu: bool = True, cluster: bool = True) -> dict:
    """The split HGP J and K kernels (hjk.py: one kernel per orbit
    representative, every ordered quartet of the canonical stream, the
    upstream's split design) on the three-atom system and the 10-atom
    cluster: ms per J and per K build (hjk.HJKBuild: the launch arguments
    staged on the device beforehand, spread over hjk.STREAMS streams; HIP
    events around three replays of the build captured as one HIP graph, and
    around three eager builds; the faster is the build's time), the
    quartets passing the in-kernel screen and their algorithmic FLOPs
    (hgp.cost of the kernel's class: VRR per primitive quartet, HRR and
    digestion per quartet) over that time; on the three-atom system
    max|dJ|, max|dK| against the oracle's ordered sweep, whose time (J and
    K together, host cores) is the CPU baseline."""
    import torch
    from felizjueves_amd import hjk
    from felizjueves_amd.systems import config1, config2
    out = {}
    work = [("c1_3atom", config1())] + ([("cluster_10atom", config2(n_atoms=10))] if cluster else [])
    for key, s in work:
        r = {"workload": s.name, "n_ao": s.basis.n_ao}
        X = {}
        for screen in ("j", "k"):
            t = time.perf_counter()
            plan = hjk.hjk_launches(s.basis, s.D, tau=1e-12, screen=screen)
            t_plan = time.perf_counter() - t
            b = hjk.HJKBuild(s.basis, s.D, screen=screen, plan=plan)
            b.run(hjk.STREAMS)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = 3
            e0.record()
            for _ in range(reps):
                b.run(hjk.STREAMS)
            e1.record()
            torch.cuda.synchronize()
            ms_eager = e0.elapsed_time(e1) / reps
            b.capture(hjk.STREAMS)
            b.replay()
            torch.cuda.synchronize()
            e0.record()
            for _ in range(reps):
                b.replay()
            e1.record()
            torch.cuda.synchronize()
            ms_graph = e0.elapsed_time(e1) / reps
            # the faster of the two forms is the build's time (the graph's
            # four branches run slower than eager streams on the larger system)
            ms = min(ms_graph, ms_eager)
            X[screen] = b.X.cpu().numpy()
            fl = b.flops
            r[screen] = {"ms_per_build": ms, "timed_as": "graph" if ms_graph <= ms_eager else "eager",
                         "ms_per_build_graph": ms_graph, "ms_per_build_eager": ms_eager,
                         "launches": len(plan.launches), "quartets_visited":
                         plan.counts()["quartets"], "quartets_computed": b.quartets_computed,
                         "plan_s": t_plan,
                         "roofline": {"bound": "fp64_valu", "achieved": fl / ms / 1e9,
                                      "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                      "frac": fl / ms / 1e9 / FP64_PEAK_TFLOPS, "gflop": fl / 1e9}}
        if key == "c1_3atom":
            sys.path.insert(0, str(ROOT / "tests"))
            from oracle import oracle as O
            t = time.perf_counter()
            Jo, Ko = O.fock_ordered(s.basis, s.D)
            t = time.perf_counter() - t
            r["max_dJ_vs_oracle"] = float(np.abs(X["j"] - Jo).max())
            r["max_dK_vs_oracle"] = float(np.abs(X["k"] - Ko).max())
            if cpu:
                r["cpu_baseline"] = {"value": 1.0 / t, "unit": "J+K builds/s", "cores": O.threads(),
                                     "kind": "port",
                                     "sample": f"the oracle's ordered sweep (J and K together) of the "
                                               f"three-atom system, {t:.1f} s"}
        log("split_jk", key, json.dumps(r))
        out[key] = r
    return out


def _aux_even_tempered(coords, ls=(0, 1, 2, 3, 4), n=4, a0=0.3, ratio=2.5):
    """A synthetic auxiliary basis: per atom, per l, n even-tempered
    single-primitive shells a0 ratio^k, unit-normalised primitives."""
    from felizjueves_amd.basis import primitive_norm
    from felizjueves_amd.df import AuxBasis
    sl, sc, ex, co = [], [], [], []
    for c in coords:
        for l in ls:  # noqa: E741
            for k in range(n):
                a = a0 * ratio ** k
                sl.append(l)
                sc.append(c)
                ex.append(a)
                co.append(primitive_norm(l, a))
    return AuxBasis(sl, sc, [1] * len(sl), ex, co)


def df_block(cpu: bool = True) -> dict:
    """The density-fitting integrals (df.py: the metric (P|Q) and the
    three-center (mn|P) on the loop-form kernel) on the upstream fixtures
    (H2O cc-pVTZ-JKFIT metric, 166 aux functions; H2O STO-3G x JKFIT
    three-center) with max|d| against libcint, and on the 10-atom cluster
    with a synthetic even-tempered s..g aux set (20 shells per atom,
    1,400 aux functions): ms per call of the launches (HIP events; the
    tables staged on the device beforehand by df.DfIntegrals; the cluster's
    (mn|P) classes over four streams), integrals
    per second and the kernel's roofline; the oracle's time on the metric
    fixture is the CPU baseline."""
    import torch
    from felizjueves_amd import df
    from felizjueves_amd.basis import GOLDEN, Basis
    from felizjueves_amd.systems import config2

    def timed_call(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {}
    z = np.load(GOLDEN / "df_h2o_ccpvtz_jkfit_metric.npz", allow_pickle=False)
    aux = df.AuxBasis.from_npz(z)
    st = df.DfIntegrals(aux)
    M = st.metric()
    ms = timed_call(lambda: st.metric(M))
    M = M.cpu().numpy()
    r = {"n_aux": aux.n_aux, "ms": ms, "max_d_vs_libcint": float(np.abs(M - z["int2c2e"]).max()),
         "integrals_per_s": aux.n_aux ** 2 / 2 / ms * 1e3}
    if cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        from oracle import oracle as O
        t = time.perf_counter()
        O.df_metric(aux)
        t = time.perf_counter() - t
        r["cpu_baseline"] = {"value": 1.0 / t, "unit": "metrics/s", "cores": O.threads(), "kind": "port",
                             "sample": f"the oracle's metric of the same aux basis, {t:.2f} s"}
    out["metric_h2o_ccpvtz_jkfit"] = r
    z = np.load(GOLDEN / "df_h2o_sto3g_jkfit_3c.npz", allow_pickle=False)
    aux = df.AuxBasis.from_npz(z)
    b = Basis.from_arrays(z["shell_l"], z["shell_centre"], z["shell_nprim"], z["exps"], z["coefs"])
    st = df.DfIntegrals(aux, b)
    E3 = st.eri3()
    ms = timed_call(lambda: st.eri3(E3))
    E3 = E3.cpu().numpy()
    out["eri3_h2o_sto3g_jkfit"] = {"n_ao": b.n_ao, "n_aux": aux.n_aux, "ms": ms,
                                   "max_d_vs_libcint": float(np.abs(E3 - z["int3c2e"]).max())}
    s = config2(n_atoms=10)
    aux = _aux_even_tempered(s.coords)
    st = df.DfIntegrals(aux, s.basis)
    Mc, Ec = st.metric(), st.eri3()
    ms_m = timed_call(lambda: st.metric(Mc))
    ms_3 = timed_call(lambda: st.eri3(Ec, n_streams=4))
    fl3, flm = df.flops_eri3(s.basis, aux), df.flops_metric(aux)
    out["cluster_10atom"] = {"workload": s.name + " + even-tempered s..g aux", "n_ao": s.basis.n_ao,
                             "n_aux": aux.n_aux, "metric_ms": ms_m, "eri3_ms": ms_3,
                             "eri3_integrals_per_s": s.basis.n_ao * (s.basis.n_ao + 1) / 2 * aux.n_aux
                             / ms_3 * 1e3,
                             "roofline": {"bound": "fp64_valu", "kernel": "fj_dfloop_k (eri3)",
                                          "achieved": fl3 / ms_3 / 1e9, "peak": FP64_PEAK_TFLOPS,
                                          "unit": "TFLOP/s", "frac": fl3 / ms_3 / 1e9 / FP64_PEAK_TFLOPS,
                                          "gflop": fl3 / 1e9, "metric_tflops": flm / ms_m / 1e9}}
    log("df", json.dumps(out))
    return out


def accuracy(D, G, Gref):
    from felizjueves_amd.fock import energy
    d = G - Gref
    return {"max_dG": float(np.abs(d).max()), "dE": abs(energy(D, d))}


def production(s, Dd, tuning, n_streams, selector, steps, warmup):
    """(builder, seconds for `steps` graph replays) of the production build."""
    from felizjueves_amd.fock import FockBuilder
    fb = FockBuilder(s.basis, s.D, tau=1e-12, mixed=True, selector=selector, **tuning)
    fb.capture(Dd, n_streams=n_streams)
    t = timed(fb.replay, steps, warmup)
    return fb, t


def c3_block(steps, warmup, selector):
    """The mixed configuration on the three-atom system."""
    import torch
    from felizjueves_amd import jk, native
    from felizjueves_amd.fock import FockBuilder
    from felizjueves_amd.systems import config1
    s = config1()
    Dd = native.to_dev(s.D)
    plain = FockBuilder(s.basis, s.D, tau=1e-12, lane_flops=1e6)
    t_plain = timed(lambda: plain.build(Dd), steps, warmup)
    G64 = plain.F.cpu().numpy()
    tuning, ns = load_tuning("c1_3atom.json", mixed=True)
    fb, t_prod = production(s, Dd, tuning, ns, selector, steps, warmup)
    Gp = fb.replay().cpu().numpy()
    torch.cuda.synchronize()
    lit = FockBuilder(s.basis, s.D, tau=1e-12, mixed=True, fp32_bound=1e-6)
    Gl = lit.build(Dd).cpu().numpy()
    J64 = jk.CoulombChain(s.basis, tau=1e-12).run(s.D)
    J32 = jk.CoulombChain(s.basis, tau=1e-12, mixed=True).run(s.D)
    K64 = jk.ExchangeBuild(s.basis, tau=1e-12).run(s.D)
    K32 = jk.ExchangeBuild(s.basis, tau=1e-12, mixed=True).run(s.D)
    torch.cuda.synchronize()
    Gjk32 = (J32 - 0.5 * K32).cpu().numpy()
    out = {
        "workload": s.name, "n_ao": s.basis.n_ao, "quartets": plain.shell_quartets(),
        "plain_fp64_builds_per_s": steps / t_plain,
        "production_builds_per_s": steps / t_prod,
        "speedup_vs_plain_fp64": t_plain / t_prod,
        "production": {"n_streams": ns, "graph": True, "tuned": bool(tuning),
                       "selector": selector is not None, "fp32_share": fb.fp32_share(),
                       **accuracy(s.D, Gp, G64)},
        "literal_bound_rule_1e-6": {"fp32_share": lit.fp32_share(), **accuracy(s.D, Gl, G64)},
        "j_mixed_lever_max_dJ": float((J32 - J64).abs().max().item()),
        "k_mixed_lever_max_dK": float((K32 - K64).abs().max().item()),
        "jk_mixed_levers_vs_fused_fp64": accuracy(s.D, Gjk32, G64),
    }
    log("c3", json.dumps({k: v for k, v in out.items() if k not in ("production", "literal_bound_rule_1e-6")}))
    return out


def c5_block(steps, warmup, selector):
    """Learned selector vs the bound rule on systems outside its training
    set (the three-atom system, seed 1, and a 10-atom cluster, seed 11)."""
    import torch
    from felizjueves_amd import native
    from felizjueves_amd.fock import FockBuilder
    from felizjueves_amd.systems import config1, config2
    if selector is None:
        return None
    out = {}
    for s in (config1(), config2(seed=11, n_atoms=10)):
        Dd = native.to_dev(s.D)
        ref = FockBuilder(s.basis, s.D, tau=1e-12)
        G64 = ref.build(Dd).cpu().numpy()
        row = {}
        for name, sel in (("bound_rule", None), ("learned", selector)):
            fb = FockBuilder(s.basis, s.D, tau=1e-12, mixed=True, selector=sel)
            fb.capture(Dd, n_streams=4)
            t = timed(fb.replay, steps, warmup)
            G = fb.replay().cpu().numpy()
            torch.cuda.synchronize()
            row[name] = {"fp32_share": fb.fp32_share(), "builds_per_s": steps / t,
                         **accuracy(s.D, G, G64)}
        out[s.name] = row
    log("c5", json.dumps(out))
    return out


def jk_block(steps, warmup):
    """The MMD families (Schwarz-screened J chain: DQ, [p|q], JAB; the K
    kernel and its scatter): J and K of the three-atom system and the
    10-atom cluster, timed per call (tables built once), and G = J - K/2
    against the fused F build."""
    import torch
    from felizjueves_amd import jk, native
    from felizjueves_amd.fock import FockBuilder
    from felizjueves_amd.systems import config1, config2
    out = {}
    for s in (config1(), config2(seed=11, n_atoms=10)):
        Dd = native.to_dev(s.D)
        cj, ck = jk.CoulombChain(s.basis, tau=1e-12), jk.ExchangeBuild(s.basis, tau=1e-12)
        state = {}

        def run_j():
            state["J"] = cj.run(Dd)

        def run_k():
            state["K"] = ck.run(Dd, n_streams=4)
        tj = timed(run_j, steps, warmup) / steps
        tk = timed(run_k, steps, warmup) / steps
        G = FockBuilder(s.basis, s.D, tau=1e-12).build(Dd)
        torch.cuda.synchronize()
        dG = float((state["J"] - 0.5 * state["K"] - G).abs().max().item())
        out[s.name] = {"j_ms": 1e3 * tj, "k_ms": 1e3 * tk, "max_dG_mmd_vs_fused": dG}
    log("jk", json.dumps(out))
    return out


def c4_block(a) -> None:
    """The batched workload: the 64 C4 molecules (8-80 atoms, seeds 4 + i)
    with their s,s,p,d shells (the f shells run on the loop-form kernel, far
    too slow for this batch) in one builder: one launch per class pair (and
    tier) for all of them, dealt to streams, captured in a graph. Timed:
    the FP64 batch, the production batch (mixed precision, the learned
    selector choosing tiers per build on the device, the FP32 tier where it
    pays), and the plain FP64 GPU build of the same workload (each molecule
    built alone, default plan, eager, one stream); accuracy of the
    production batch against the FP64 one, per molecule, and of the FP64
    batch against each checked molecule built alone."""
    import torch
    from felizjueves_amd import native
    from felizjueves_amd.fock import FockBuilder, energy
    from felizjueves_amd.systems import batch_molecule
    systems = [batch_molecule(i, lmax=2) for i in range(64)]
    sy = [(s.basis, s.D) for s in systems]
    Ds = [native.to_dev(s.D) for s in systems]
    tuning, ns = load_tuning("c4_64mol_d.json")
    ns = ns if tuning else 4
    t0 = time.time()
    fb = FockBuilder.batch(sy, tau=1e-12, **tuning)
    plan_s = time.time() - t0
    log(f"c4d: 64 molecules, n_ao {sum(s.basis.n_ao for s in systems)}, quartets {fb.shell_quartets():.3e}, "
        f"{fb.flops / 1e9:.0f} GFLOP/build, plan {plan_s:.0f}s")
    fb.capture(Ds, n_streams=ns)
    t64 = timed(fb.replay, a.steps, a.warmup)
    G64 = [G.cpu().numpy() for G in fb.replay()]
    torch.cuda.synchronize()
    log(f"c4d: fp64 batch {1e3 * t64 / a.steps:.0f} ms per build")
    n_q, flops = fb.shell_quartets(), fb.flops
    del fb
    selector = load_selector(a.selector)
    tuning_mx = load_tuning("c4_64mol_d.json", mixed=True)[0]
    t0 = time.time()
    fbp = FockBuilder.batch(sy, tau=1e-12, mixed=True, selector=selector, **tuning_mx)
    plan_p = time.time() - t0
    fbp.capture(Ds, n_streams=ns)
    tp = timed(fbp.replay, a.steps, a.warmup)
    Gp = [G.cpu().numpy() for G in fbp.replay()]
    torch.cuda.synchronize()
    share = fbp.fp32_share()
    dG = max(float(np.abs(G - R).max()) for G, R in zip(Gp, G64))
    dE = max(abs(energy(s.D, G) - energy(s.D, R)) for s, G, R in zip(systems, Gp, G64))
    log(f"c4d: production batch {1e3 * tp / a.steps:.0f} ms per build, fp32 share {share:.4f}, "
        f"max|dG| {dG:.2e}, max|dE| {dE:.2e} vs the fp64 batch")
    del fbp
    # the plain FP64 GPU build of the workload: every molecule alone
    t_plain, worst, checked = 0.0, 0.0, list(range(0, 64, 9))
    for i, (s, Dd) in enumerate(zip(systems, Ds)):
        one = FockBuilder(s.basis, s.D, tau=1e-12)
        t_plain += timed(lambda: one.build(Dd), 1, 1)
        if i in checked:
            worst = max(worst, float(np.abs(one.F.cpu().numpy() - G64[i]).max()))
        del one
    log(f"c4d: plain fp64, each molecule alone: {1e3 * t_plain:.0f} ms for the 64")
    out = {
        "metric": "Fock builds/s of the 64-molecule batch (all molecules per build)",
        "value": a.steps / tp, "unit": "batched builds/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * tp / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (f32 VRR tier on selected tiles)",
        "data": "synthetic",
        "config": {"workload": "c4-64mol-8to80atoms-seed4+i-spd", "shells": "s,s,p,d (f not run)",
                   "tau": 1e-12, "n_streams": ns, "graph": True, "tuned": bool(tuning),
                   "selector": selector is not None},
        "fp64_batch_builds_per_s": a.steps / t64,
        "plain_fp64_each_molecule_alone_s": t_plain,
        "speedup_vs_plain_fp64": t_plain / (tp / a.steps),
        "fp32_share_quartets": share, "max_dG_vs_fp64_batch": dG, "max_dE_vs_fp64_batch": dE,
        "quartets_per_build": n_q, "gflop_per_build": flops / 1e9,
        "tflops_algorithmic": flops * a.steps / tp / 1e12,
        "plan_seconds_fp64": plan_s, "plan_seconds_production": plan_p,
        "max_dG_fp64_batch_vs_molecule_alone": worst, "molecules_checked": checked,
    }
    print(json.dumps(out), flush=True)


def cof_block(a) -> None:
    """The f workload: CO/cc-pVTZ (f shells on both atoms), the Fock build
    through (ff|ff) -- HGP kernels through (dd|dd), the classes with an f
    shell on the loop-form kernel -- and the MMD families through f (J
    chain, K, Schwarz), each timed on the device and checked against
    libcint's vj / vk of the committed fixture."""
    import torch
    from felizjueves_amd import jk, native
    from felizjueves_amd.basis import load_fixture
    from felizjueves_amd.fock import FockBuilder
    fx = load_fixture("co_ccpvtz")
    b, D = fx.basis, fx.density
    Dd = native.to_dev(D)
    ref = fx.arrays["vj"] - 0.5 * fx.arrays["vk"]
    fb = FockBuilder(b, D, tau=1e-12)
    n_f = sum(1 for L in fb.launches if 3 in L.cls)
    t = timed(lambda: fb.build(Dd, n_streams=a.streams), a.steps, a.warmup)
    G = fb.build(Dd, n_streams=a.streams).cpu().numpy()
    cj, ck = jk.CoulombChain(b), jk.ExchangeBuild(b)
    tj = timed(lambda: cj.run(Dd), a.steps, a.warmup)
    tk = timed(lambda: ck.run(Dd, n_streams=4), max(1, a.steps // 4), 1)
    ts = timed(lambda: jk.schwarz_matrix(b), a.steps, a.warmup)
    J, K = cj.run(Dd).cpu().numpy(), ck.run(Dd).cpu().numpy()
    torch.cuda.synchronize()
    out = {
        "metric": "Fock builds/s, CO/cc-pVTZ (f shells; f classes on the loop-form kernel)",
        "value": a.steps / t, "unit": "Fock builds/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * t / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "committed fixture",
        "config": {"workload": "co-ccpvtz", "n_ao": b.n_ao, "tau": 1e-12, "n_streams": a.streams},
        "launches": len(fb.launches), "f_class_launches": n_f,
        "quartets_per_build": fb.shell_quartets(),
        "max_dG_vs_libcint": float(np.abs(G - ref).max()),
        "j_chain_ms": 1e3 * tj / a.steps, "k_ms": 1e3 * tk / max(1, a.steps // 4),
        "schwarz_ms": 1e3 * ts / a.steps,
        "max_dJ_vs_libcint": float(np.abs(J - fx.arrays["vj"]).max()),
        "max_dK_vs_libcint": float(np.abs(K - fx.arrays["vk"]).max()),
    }
    print(json.dumps(out), flush=True)


def c4f_block(a) -> None:
    """C4 molecules with their f shells, each built alone (FP64, the
    launches dealt to --streams streams, eager): one warm-up and `--steps` timed builds per molecule, beside the
    same molecule built through d. The whole 64-molecule batch with f is not
    timed: the f classes run on the loop-form kernels, which make the
    11-atom molecule 17 times slower than through d, and the batch holds
    molecules of up to 80 atoms (45,402 AOs in all)."""
    import torch
    from felizjueves_amd import native
    from felizjueves_amd.fock import FockBuilder
    from felizjueves_amd.systems import batch_molecule
    rows = []
    for i in a.molecules:
        row = {"molecule": i}
        for lmax in (2, 3):
            s = batch_molecule(i, lmax=lmax)
            t0 = time.time()
            fb = FockBuilder(s.basis, s.D, tau=1e-12)
            plan_s = time.time() - t0
            Dd = native.to_dev(s.D)
            log(f"c4f: {s.name} lmax {lmax}: n_ao {s.basis.n_ao}, {len(fb.launches)} launches, "
                f"plan {plan_s:.1f}s; building")
            t = timed(lambda: fb.build(Dd, n_streams=a.streams), a.steps, 1)
            torch.cuda.synchronize()
            log(f"c4f: {s.name} lmax {lmax}: {1e3 * t / a.steps:.1f} ms per build")
            row.update({"name": s.name, f"n_ao_l{lmax}": s.basis.n_ao,
                        f"launches_l{lmax}": len(fb.launches),
                        f"f_class_launches_l{lmax}": sum(1 for L in fb.launches if 3 in L.cls),
                        f"quartets_l{lmax}": fb.shell_quartets(),
                        f"ms_per_build_l{lmax}": 1e3 * t / a.steps})
        rows.append(row)
    out = {
        "metric": "Fock build ms of single C4 molecules with and without their f shells",
        "value": rows[0]["ms_per_build_l3"], "unit": "ms per build (first molecule, with f)",
        "n_gpus": 1, "steps": a.steps, "warmup": 1, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "c4-single-molecules-with-f", "molecules": a.molecules, "tau": 1e-12,
                   "n_streams": a.streams, "graph": False},
        "molecules": rows,
    }
    print(json.dumps(out), flush=True)


def c4f_batch_block(a) -> None:
    """Chosen C4 molecules with their f shells in one batch builder (one
    launch per class pair for all of them, --streams streams, eager; the f classes
    on the loop-form kernel): batched builds/s, and the first molecule
    against its own single build."""
    import torch
    from felizjueves_amd import native
    from felizjueves_amd.fock import FockBuilder
    from felizjueves_amd.systems import batch_molecule
    systems = [batch_molecule(i, lmax=3) for i in a.molecules]
    t0 = time.time()
    fb = FockBuilder.batch([(s.basis, s.D) for s in systems], tau=1e-12)
    plan_s = time.time() - t0
    Ds = [native.to_dev(s.D) for s in systems]
    log(f"c4f-batch: {[s.name for s in systems]}: {len(fb.launches)} launches, "
        f"quartets {fb.shell_quartets():.3e}, plan {plan_s:.1f}s; building")
    t = timed(lambda: fb.build(Ds, n_streams=a.streams), a.steps, 1)
    Gs = [G.cpu().numpy() for G in fb.build(Ds, n_streams=a.streams)]
    torch.cuda.synchronize()
    log(f"c4f-batch: {1e3 * t / a.steps:.0f} ms per batched build; checking molecule {a.molecules[0]}")
    s0 = systems[0]
    G0 = FockBuilder(s0.basis, s0.D, tau=1e-12).build(Ds[0]).cpu().numpy()
    out = {
        "metric": "batched Fock builds/s of chosen C4 molecules with their f shells",
        "value": a.steps / t, "unit": "batched builds/s", "n_gpus": 1, "steps": a.steps,
        "warmup": 1, "ms_per_step": 1e3 * t / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "c4-molecules-with-f-batched", "molecules": a.molecules,
                   "atoms": [s.basis.n_sh // 5 for s in systems], "tau": 1e-12,
                   "n_streams": a.streams, "graph": False},
        "n_ao": [s.basis.n_ao for s in systems], "launches": len(fb.launches),
        "f_class_launches": sum(1 for L in fb.launches if 3 in L.cls),
        "quartets_per_build": fb.shell_quartets(), "plan_seconds": plan_s,
        "max_dG_first_vs_alone": float(np.abs(Gs[0] - G0).max()),
    }
    print(json.dumps(out), flush=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--atoms", type=int, default=40)
    ap.add_argument("--selector", default=str(TUNED / "selector"))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the c3 / c5 blocks")
    ap.add_argument("--streams", type=int, default=4,
                    help="cof / c4f / c4f-batch: streams the builder's launches are dealt to")
    ap.add_argument("--molecules", type=int, nargs="+", default=[7],
                    help="c4f: which C4 molecules (index i: seed 4 + i)")
    ap.add_argument("--workload", choices=("c2", "c4d", "cof", "c4f", "c4f-batch", "gradient",
                                           "split-df"),
                    default="c2",
                    help="c4d: the 64-molecule batch (8-80 atoms, seeds 4+i) with s,s,p,d shells "
                         "(with its f shells, 45,402 AOs, one build would take minutes on the "
                         "loop-form f kernels); cof: CO/cc-pVTZ through f (the loop-form kernels); c4f: "
                         "single C4 molecules with and without their f shells; c4f-batch: "
                         "chosen C4 molecules with f in one batch builder; gradient: the fused "
                         "two-electron gradient block alone; split-df: the split J / K and "
                         "density-fitting blocks alone")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for HBM counter runs: one elementwise copy of 512 MiB (the counters' "
                         "scale) and one eager build of the default plan, then exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): the Fock build is a "
              "single-GPU path here (across GPUs it would split the cost-ordered work items "
              "and sum the private F matrices with one all-reduce).", file=sys.stderr)
        sys.exit(2)

    import torch
    from felizjueves_amd import native
    from felizjueves_amd.fock import FockBuilder
    from felizjueves_amd.systems import config2

    if a.traffic_probe:
        s = config2(n_atoms=a.atoms)
        Dd = native.to_dev(s.D)
        n = 64 << 20
        src = torch.ones(n, dtype=torch.float64, device=Dd.device)
        dst = torch.empty_like(src)
        torch.add(src, 0.0, out=dst)          # reads and writes 512 MiB
        plain = FockBuilder(s.basis, s.D, tau=1e-12)
        plain.build(Dd)
        torch.cuda.synchronize()
        lt = launch_times(plain, Dd, reps=1)
        Lp = plain.launches[int(np.argmax(lt[0]))]
        print(json.dumps({"copy_bytes_each_way": n * 8, "dominant": f"fj_f_{''.join(map(str, Lp.cls))}_k",
                          "dominant_gflop": Lp.flops / 1e9}), flush=True)
        return

    if a.workload == "c4d":
        c4_block(a)
        return
    if a.workload == "cof":
        cof_block(a)
        return
    if a.workload == "c4f":
        c4f_block(a)
        return
    if a.workload == "c4f-batch":
        c4f_batch_block(a)
        return
    if a.workload == "gradient":
        print(json.dumps({"gradient": gradient_block(a.steps, cpu=not a.no_cpu)}), flush=True)
        return
    if a.workload == "split-df":
        print(json.dumps({"split_jk": split_jk_block(cpu=not a.no_cpu),
                          "density_fitting": df_block(cpu=not a.no_cpu)}), flush=True)
        return

    selector = load_selector(a.selector)
    s = config2(n_atoms=a.atoms)
    D = s.D
    Dd = native.to_dev(D)
    plain = FockBuilder(s.basis, D, tau=1e-12)
    log(f"{s.name}: n_ao {s.basis.n_ao}, quartets {plain.shell_quartets():.3e}, primitive quartets "
        f"{plain.n_prim_quartets:.3e}, {plain.flops / 1e9:.1f} GFLOP/build, plan {plain.plan_seconds:.1f}s")

    # roofline of the dominant kernel: eager one-stream launches, events
    plain.build(Dd)
    torch.cuda.synchronize()
    lt = launch_times(plain, Dd, reps=3)
    probe = int(np.argmax(lt.mean(0)))
    Lp = plain.launches[probe]
    avg_ms = float(lt[:, probe].mean())

    t_plain = timed(lambda: plain.build(Dd), a.steps, a.warmup)
    G64 = plain.F.cpu().numpy()
    log(f"plain fp64: {a.steps / t_plain:.3f} builds/s")

    tuning, ns = load_tuning("c2_40atom.json") if a.atoms == 40 else ({}, 4)
    fb64 = FockBuilder(s.basis, D, tau=1e-12, **tuning)
    fb64.capture(Dd, n_streams=ns)
    t_t64 = timed(fb64.replay, a.steps, a.warmup)
    log(f"tuned fp64 graph: {a.steps / t_t64:.3f} builds/s")

    tuning_mx = (load_tuning("c2_40atom.json", mixed=True)[0] if a.atoms == 40 else tuning)
    fbp, t_prod = production(s, Dd, tuning_mx, ns, selector, a.steps, a.warmup)
    Gp = fbp.replay().cpu().numpy()
    torch.cuda.synchronize()
    acc = accuracy(D, Gp, G64)
    log(f"production (mixed, tuned, {ns} streams, graph): {a.steps / t_prod:.3f} builds/s, "
        f"fp32 share {fbp.fp32_share():.4f}, {acc}")

    achieved = Lp.flops / (avg_ms * 1e-3) / 1e12
    roofline = {
        "bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
        "kernel": f"fj_f_{''.join(map(str, Lp.cls))}_k<double>",
        "note": ("compute-bound on FP64 VALU (no MFMA in this kernel); 78.6 TFLOP/s is MI355X's FP64 "
                 "vector peak (equal to its FP64 matrix peak); achieved = algorithmic FLOPs of the "
                 "launch (hgp.Cost: VRR + boundary per primitive quartet, HRR + digestion per "
                 "quartet, no FMA credit) / mean duration of its eager launches (default plan, one "
                 "stream, HIP events on that stream)"),
        "launch_ms": avg_ms, "launch_gflop": Lp.flops / 1e9,
    }
    tr = TUNED / f"traffic_c2_{a.atoms}atom.json"
    if tr.exists():
        t = json.loads(tr.read_text())
        if t.get("kernel") == roofline["kernel"]:
            roofline["traffic"] = t["traffic_bytes"]
            roofline["traffic_note"] = (t["source"] + f", scaled by a 512 MiB copy (fetch x{t['fetch_scale']:.3f}, "
                                        f"write x{t['write_scale']:.3f}); mostly the launch workspace")
    va = TUNED / f"valu_c2_{a.atoms}atom.json"
    if va.exists():
        t = json.loads(va.read_text())
        if t.get("kernel") == roofline["kernel"]:
            roofline["executed_over_algorithmic"] = t["executed_over_algorithmic"]
            roofline["executed_tflops"] = achieved * t["executed_over_algorithmic"]
            roofline["valu_active_frac"] = t["valu_active_frac"]
            roofline["counters_note"] = t["source"]
    extra = {}
    if not a.no_extra:
        extra["c3"] = c3_block(max(10, a.steps), a.warmup, selector)
        extra["c5"] = c5_block(max(5, a.steps // 2), a.warmup, selector)
        extra["jk"] = jk_block(3, 1)
        extra["pair_stream"] = stream_block(s)
        extra["gradient"] = gradient_block(max(5, a.steps // 2), cpu=not a.no_cpu)
        extra["split_jk"] = split_jk_block(cpu=not a.no_cpu, cluster=False)
        extra["density_fitting"] = df_block(cpu=not a.no_cpu)
    cpu = None if a.no_cpu else cpu_baseline(plain, D, G64)
    out = {
        "metric": METRIC, "value": a.steps / t_prod, "unit": "Fock builds/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * t_prod / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (f32 VRR tier on selected tiles)", "data": "synthetic",
        "config": {"workload": s.name, "n_atoms": a.atoms, "n_ao": s.basis.n_ao, "tau": 1e-12,
                   "fp32_error_per_tile": fbp.fp32_error, "n_streams": ns, "graph": True,
                   "tuned": bool(tuning), "selector": selector is not None},
        "fp64_plain_builds_per_s": a.steps / t_plain, "fp64_plain_ms_per_build": 1e3 * t_plain / a.steps,
        "fp64_tuned_graph_builds_per_s": a.steps / t_t64,
        "quartets_per_s": plain.shell_quartets() * a.steps / t_prod,
        "quartets_per_s_fp64_plain": plain.shell_quartets() * a.steps / t_plain,
        "prim_quartets_per_s": plain.n_prim_quartets * a.steps / t_prod,
        "quartets_per_build": plain.shell_quartets(), "gflop_per_build": plain.flops / 1e9,
        "tflops_algorithmic": plain.flops * a.steps / t_prod / 1e12,
        "speedup_vs_fp64_plain": t_plain / t_prod, "fp32_share_quartets": fbp.fp32_share(),
        "dE_vs_fp64": acc["dE"], "max_dG_vs_fp64": acc["max_dG"],
        **extra, "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
