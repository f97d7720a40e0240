"""Benchmark: Fock builds/s and quartets/s on one MI355X, mixed vs FP64.

Workload (default): the 40-atom random cluster (s,s,p,d per atom, n_ao 440,
tau 1e-12, seed 2), one step = one closed two-electron Fock build
G = J - K/2 from device-resident pair tables and density.  Timed: the plain
FP64 build and the mixed build (FP32 tier for the kets whose estimated FP32
error -- the six-block bound times the HRR amplification, fock.py -- is
below 1e-11 per tile), K builds each after W warm-up builds.

Prints ONE JSON line.  `value` is the mixed build's Fock builds/s; the FP64
build's figure, the quartet rates, |dE| and max|dG| of mixed vs FP64, the
roofline of the dominant kernel (HIP events on the launch stream) and the
CPU baseline (the oracle on the host cores, a bounded sample) ride along.
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


def timed_builds(fb, Dd, steps, warmup, probe=None):
    """Run warm-up then `steps` timed builds; optionally record HIP events
    around launch index `probe` of each timed build."""
    import torch
    from felizjueves_amd import native
    for _ in range(warmup):
        fb.build(Dd)
    torch.cuda.synchronize()
    evs = []
    t0 = time.perf_counter()
    for _ in range(steps):
        if probe is None:
            fb.build(Dd)
            continue
        # the same launches as FockBuilder.build, with events around `probe`
        sys = fb._sys([Dd])
        fb.F.zero_()
        native.density_blocks(fb.shell_ao[0], fb.shell_size[0], Dd, fb.Dsh)
        for i, L in enumerate(fb.launches):
            if i == probe:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
            fb.launch(L, sys, None, fb._workspace(0))
            if i == probe:
                e1.record()
                evs.append((e0, e1))
        native.close(fb.F)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = [a.elapsed_time(b) for a, b in evs]
    return dt, ms


def launch_times(fb, Dd):
    """Per-launch durations (ms) of one build, HIP events on the stream."""
    import torch
    from felizjueves_amd import native
    sys = fb._sys([Dd])
    fb.F.zero_()
    native.density_blocks(fb.shell_ao[0], fb.shell_size[0], Dd, fb.Dsh)
    evs = []
    for L in fb.launches:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fb.launch(L, sys, None, fb._workspace(0))
        e1.record()
        evs.append((e0, e1))
    native.close(fb.F)
    torch.cuda.synchronize()
    return [a.elapsed_time(b) for a, b in evs]


def cpu_baseline(fb, D, budget_s=15.0):
    """The CPU oracle (OpenMP over the host cores) on a bounded sample of
    the same workload: every `stride`-th bra pair of the costliest
    off-diagonal class pair, all its kets; FLOPs of the sample counted by
    the same screen and cost model as the device plan."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle import oracle as O
    from felizjueves_amd import pairs as P
    from felizjueves_amd.hgp import cost
    from test_builder_oracle import builder_tables_host
    host = builder_tables_host(fb)
    Ls = [L for L in fb.launches if not L.diag and L.precision == 0]
    L = max(Ls, key=lambda L: L.flops)
    X, Y = L.cls[:2], L.cls[2:]
    bra, ket = fb.tables[X], fb.tables[Y]
    c = cost(L.cls)
    Dsh = P.density_blocks(fb.basis, D)

    def sample(stride):
        idx = np.arange(0, bra.n, stride)
        fl, nq = 0.0, 0
        for b in idx:
            bound = fb._bounds(bra, ket, b)
            live = bound >= fb.tau
            nq += int(live.sum())
            fl += c.per_prim * int(bra.deg[b]) * float(ket.deg[live].sum()) + c.per_quartet * live.sum()
        hb = host[X]
        off = hb["off"]
        sub = {"prim": np.concatenate([hb["prim"][off[b]:off[b + 1]] for b in idx]),
               "off": np.concatenate([[0], np.cumsum([off[b + 1] - off[b] for b in idx])]).astype(np.int32),
               "sep": hb["sep"][idx], "G": hb["G"][idx], "sh": hb["sh"][idx], "ao": hb["ao"][idx]}
        F = np.zeros((fb.basis.n_ao,) * 2)
        t = time.perf_counter()
        O.fx_launch(L.cls, sub, host[Y], 0, fb.basis.n_sh, fb.basis.n_ao, fb.tau, Dsh, D, F)
        return time.perf_counter() - t, fl, nq, len(idx)

    # grow the sample (halving the stride) until it takes a third of the
    # budget, then size the measured one to the budget
    stride = max(1, bra.n // 8)
    t, fl, nq, nb = sample(stride)
    while t < budget_s / 3 and stride > 1:
        stride = max(1, stride // 2)
        t, fl, nq, nb = sample(stride)
    if t < budget_s and stride > 1:
        stride = max(1, int(stride * t / budget_s))
        t, fl, nq, nb = sample(stride)
    gflops = fl / t / 1e9
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return {
        "value": (fl / t) / fb.flops, "unit": "Fock builds/s", "cores": cores, "kind": "port",
        "sample": (f"oracle f-launch, class {''.join(map(str, L.cls))}, every {stride}th bra pair "
                   f"({nb} of {bra.n}) x all {ket.n} ket pairs: {nq} quartets, "
                   f"{fl / 1e9:.1f} GFLOP in {t:.1f} s = {gflops:.2f} GFLOP/s; builds/s = that rate / "
                   f"{fb.flops / 1e9:.1f} GFLOP per build"),
        "quartets_per_s": nq / t, "gflops": gflops,
    }


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--atoms", type=int, default=40)
    ap.add_argument("--threads", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): the Fock build is a "
              "single-GPU path here; see DESIGN.md for how it would shard.", file=sys.stderr)
        sys.exit(2)

    import torch
    from felizjueves_amd import native
    from felizjueves_amd.fock import FockBuilder, energy, FP64
    from felizjueves_amd.systems import config2

    s = config2(n_atoms=a.atoms)
    D = s.D
    fb64 = FockBuilder(s.basis, D, tau=1e-12, threads=a.threads)
    fbmx = FockBuilder(s.basis, D, tau=1e-12, threads=a.threads, mixed=True)
    Dd = native.to_dev(D)
    print(f"[bench] {s.name}: n_ao {s.basis.n_ao}, quartets {fb64.n_quartets:.3e}, "
          f"primitive quartets {fb64.n_prim_quartets:.3e}, {fb64.flops / 1e9:.1f} GFLOP/build, "
          f"plan {fb64.plan_seconds:.1f}s", file=sys.stderr, flush=True)

    # dominant kernel of the FP64 build
    fb64.build(Dd)
    torch.cuda.synchronize()
    lt = launch_times(fb64, Dd)
    probe = int(np.argmax(lt))
    Lp = fb64.launches[probe]

    t64, ms_probe = timed_builds(fb64, Dd, a.steps, a.warmup, probe=probe)
    G64 = fb64.F.cpu().numpy()
    print(f"[bench] fp64: {a.steps / t64:.3f} builds/s", file=sys.stderr, flush=True)
    tmx, _ = timed_builds(fbmx, Dd, a.steps, a.warmup)
    Gmx = fbmx.F.cpu().numpy()
    print(f"[bench] mixed: {a.steps / tmx:.3f} builds/s (fp32 share {fbmx.fp32_share():.3f})",
          file=sys.stderr, flush=True)

    dE = abs(energy(D, Gmx) - energy(D, G64))
    dG = float(np.abs(Gmx - G64).max())
    avg_ms = float(np.mean(ms_probe))
    achieved = Lp.flops / (avg_ms * 1e-3) / 1e12
    roofline = {
        "bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
        "kernel": f"fj_f_{''.join(map(str, Lp.cls))}_k<double>",
        "note": ("compute-bound on FP64 VALU (no MFMA in this kernel); 78.6 TFLOP/s is MI355X's FP64 "
                 "vector peak (equal to its FP64 matrix peak); achieved = algorithmic FLOPs of the "
                 "launch (upstream hgp/cost.py convention, no FMA credit) / mean launch time"),
        "launch_ms": avg_ms, "launch_gflop": Lp.flops / 1e9,
    }
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(fb64, D)
    out = {
        "metric": METRIC, "value": a.steps / tmx, "unit": "Fock builds/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * tmx / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32-tier", "data": "synthetic",
        "config": {"workload": s.name, "n_atoms": a.atoms, "n_ao": s.basis.n_ao, "tau": 1e-12,
                   "fp32_error_per_tile": fbmx.fp32_error, "threads_per_workgroup": a.threads},
        "fp64_builds_per_s": a.steps / t64, "fp64_ms_per_build": 1e3 * t64 / a.steps,
        "quartets_per_s_mixed": fbmx.n_quartets * a.steps / tmx,
        "quartets_per_s_fp64": fb64.n_quartets * a.steps / t64,
        "prim_quartets_per_s_fp64": fb64.n_prim_quartets * a.steps / t64,
        "quartets_per_build": fb64.n_quartets, "gflop_per_build": fb64.flops / 1e9,
        "fp64_tflops_algorithmic": fb64.flops * a.steps / t64 / 1e12,
        "mixed_speedup_vs_fp64": t64 / tmx, "fp32_share_quartets": fbmx.fp32_share(),
        "dE_mixed_vs_fp64": dE, "max_dG_mixed_vs_fp64": dG,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
