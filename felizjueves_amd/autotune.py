"""Autotuner of the batched Fock build: per class pair, the lanes per work
item (threads) and the primitive-loop split (lane_flops), then the number
of streams the launches are dealt to.

Each candidate plan is timed on the device: HIP events around every launch
of a whole build (median over repeats), summed per class pair over its
tiers. The fastest candidate per class pair is kept; the tuned plan is then
timed as whole captured builds for each stream count. The choice is a pure
function of the measured times, written to / read from a JSON file so that
a tuned configuration can be reused without measuring again.
"""

from __future__ import annotations

import json
import time
from pathlib import Path

import numpy as np

from . import native
from .fock import FockBuilder

THREADS = (64, 128, 256)
LANE_FLOPS = (3.0e4, 1.0e5, 3.0e5, 1.0e6)
STREAMS = (1, 2, 4)


def launch_times(fb: FockBuilder, D_devs, reps: int = 3) -> dict:
    """{class pair: median over reps of the summed launch times (ms)} of
    one eager build, each launch bracketed by events on the stream it runs
    on."""
    import torch
    fb.build(D_devs)
    torch.cuda.synchronize()
    sys = fb._sys(D_devs)
    ws = fb._workspace(0)
    per: dict = {}
    for _ in range(reps):
        evs = []
        for L in fb.launches:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fb.launch(L, sys, None, ws)
            e1.record()
            evs.append((L.cls, e0, e1))
        torch.cuda.synchronize()
        tot: dict = {}
        for c, a, b in evs:
            tot[c] = tot.get(c, 0.0) + a.elapsed_time(b)
        for c, t in tot.items():
            per.setdefault(c, []).append(t)
    return {c: float(np.median(v)) for c, v in per.items()}


def build_time(fb: FockBuilder, D_devs, n_streams: int, steps: int = 10) -> float:
    """ms per captured-graph build with n_streams."""
    import torch
    fb.capture(D_devs, n_streams=n_streams)
    fb.replay()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        fb.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / steps * 1e3


def tune(systems, threads=THREADS, lane_flops=LANE_FLOPS, streams=STREAMS, log=None, **kw) -> dict:
    """Tune the build of `systems` [(basis, D), ...] (other FockBuilder
    keywords in kw). Returns {"threads": {cls: t}, "lane_flops": {cls: lf},
    "n_streams": s, "ms": {...}} with class pairs as 'labcd' strings."""
    D_devs = [native.to_dev(D) for _, D in systems]
    best: dict = {}
    table = {}
    for t in threads:
        for lf in lane_flops:
            fb = FockBuilder.batch(systems, threads=t, lane_flops=lf, **kw)
            ms = launch_times(fb, D_devs)
            table[f"{t}/{lf:g}"] = {"".join(map(str, c)): v for c, v in ms.items()}
            for c, v in ms.items():
                if c not in best or v < best[c][0]:
                    best[c] = (v, t, lf)
            if log:
                log(t, lf, sum(ms.values()))
            del fb
    th = {c: b[1] for c, b in best.items()}
    lfs = {c: b[2] for c, b in best.items()}
    fb = FockBuilder.batch(systems, threads=th, lane_flops=lfs, **kw)
    st = {s: build_time(fb, D_devs, s) for s in streams}
    n_streams = min(st, key=st.get)
    if log:
        log("streams", st)
    return {"threads": {"".join(map(str, c)): v for c, v in th.items()},
            "lane_flops": {"".join(map(str, c)): v for c, v in lfs.items()},
            "n_streams": n_streams, "graph_ms": {str(k): v for k, v in st.items()},
            "launch_ms": table}


def tier_pays(systems, tuned: dict, margin: float = 0.02, log=None, **kw) -> list:
    """The class pairs whose mixed launch (per-item FP32 / FP64 tiers, kw
    the mixed builder's keywords, e.g. the selector) is measured faster than
    their FP64 launch at the tuned threads and split, by more than
    `margin`: a mixed build sends work to the FP32 tier only there (on
    MI355X, FP32 VALU without packed math issues at the FP64 rate, so the
    tier pays only where its smaller register footprint does)."""
    D_devs = [native.to_dev(D) for _, D in systems]
    kwt = builder_kwargs(tuned)
    t64 = launch_times(FockBuilder.batch(systems, **kwt), D_devs, reps=5)
    tmx = launch_times(FockBuilder.batch(systems, mixed=True, **kwt, **kw), D_devs, reps=5)
    out = sorted(c for c in tmx if c in t64 and tmx[c] < (1.0 - margin) * t64[c])
    if log:
        log({"".join(map(str, c)): (round(t64[c], 3), round(tmx.get(c, float("nan")), 3))
             for c in sorted(t64)})
    return ["".join(map(str, c)) for c in out]


def save(cfg: dict, path: Path) -> None:
    Path(path).write_text(json.dumps(cfg, indent=1, sort_keys=True))


def load(path: Path) -> dict:
    return json.loads(Path(path).read_text())


def builder_kwargs(cfg: dict) -> dict:
    """FockBuilder keywords of a tuned configuration."""
    key = lambda s: tuple(int(ch) for ch in s)  # noqa: E731
    out = {"threads": {key(k): v for k, v in cfg["threads"].items()},
           "lane_flops": {key(k): v for k, v in cfg["lane_flops"].items()}}
    return out


def mixed_kwargs(cfg: dict) -> dict:
    """The mixed builder's extra keywords of a tuned configuration: the class
    pairs where the FP32 tier was measured to pay (none recorded: every
    class pair)."""
    if "fp32_classes" not in cfg:
        return {}
    return {"fp32_classes": [tuple(int(ch) for ch in c) for c in cfg["fp32_classes"]]}


def main() -> None:
    """Tune a workload (--tune: per class pair threads and split, then the
    stream count) and / or measure where the FP32 tier pays for the tuned
    configuration (`fp32_classes`), writing the configuration file.

        python -m felizjueves_amd.autotune tuned/c2_40atom.json --workload c2
        python -m felizjueves_amd.autotune tuned/c4_64mol_d.json --workload c4 --tune --every 4

    c4: the batched workload's molecules (shells through d), every `every`-th
    of the 64 (a subset tunes the per-class choices in a fraction of the
    time; the choices are per class pair and carry to the whole batch)."""
    import argparse
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("config", help="tuned configuration file (JSON) to update or create")
    ap.add_argument("--workload", choices=("c1", "c2", "c4"), required=True)
    ap.add_argument("--tune", action="store_true", help="run the threads / split / streams search first")
    ap.add_argument("--every", type=int, default=4, help="c4: tune on every N-th molecule")
    ap.add_argument("--lane-flops", type=float, nargs="+", default=list(LANE_FLOPS),
                    help="the primitive-loop split candidates (FLOPs per lane)")
    ap.add_argument("--selector", default=str(Path(__file__).resolve().parent / "tuned" / "selector"))
    ap.add_argument("--out", default=None, help="write the updated configuration here instead")
    a = ap.parse_args()
    from .selector import LearnedSelector
    from .systems import batch_molecule, config1, config2
    if a.workload == "c4":
        systems = [(m.basis, m.D) for m in (batch_molecule(i, lmax=2) for i in range(0, 64, a.every))]
    else:
        s = config1() if a.workload == "c1" else config2()
        systems = [(s.basis, s.D)]
    log = lambda *m: print("[autotune]", *m, flush=True)  # noqa: E731
    if a.tune:
        t0 = time.perf_counter()
        cfg = tune(systems, lane_flops=tuple(a.lane_flops), tau=1e-12, log=log)
        cfg["tuned_on"] = (f"{a.workload}: {len(systems)} systems" +
                           (f" (every {a.every}-th of the 64 batch molecules, shells through d)"
                            if a.workload == "c4" else ""))
        log(f"tuned in {time.perf_counter() - t0:.0f} s: streams {cfg['n_streams']}, graph ms {cfg['graph_ms']}")
    else:
        cfg = load(a.config)
    sel = LearnedSelector.load(a.selector)
    pays = tier_pays(systems, cfg, selector=sel, tau=1e-12,
                     log=lambda m: print("[tier_pays] fp64 ms, mixed ms:", m, flush=True))
    cfg["fp32_classes"] = pays
    save(cfg, a.out or a.config)
    print(f"[tier_pays] {len(pays)} class pairs: {pays}", flush=True)
    sys.exit(0)


if __name__ == "__main__":
    main()
