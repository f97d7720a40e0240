"""Build libfj_mi355x.so for gfx950: generate the per-class HIP units, compile
them in parallel with hipcc, link one shared library in-tree.

    python -m felizjueves_amd.build [--lmax 2] [--jobs N]
"""

from __future__ import annotations

import argparse
import os
import subprocess
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

from . import gen_hip, gen_mmd
from .hgp import builder_classes

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "libfj_mi355x.so"
ARCH = "gfx950"
LMAX = 2         # generated HGP fused F kernels through (dd|dd)
MMD_LMAX = 3     # Schwarz, DQ, [p|q] and JAB through f
LOOP_LMAX = 3    # fused F classes past LMAX run on the loop-form kernel (fj_floop.h)
MMD_K_LMAX = 2   # straight-line K kernels through d; the f classes in loop form (fj_kloop.h)
SPLIT_LMAX = 2   # split HGP J / K kernels (one per orbit) through (dd|dd)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    # packed FP32 VALU forms are kept out of every code object built here
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-variable", "-Wno-unused-parameter",
]


def sources(lmax: int = LMAX) -> list[Path]:
    classes = builder_classes(lmax)
    files = gen_hip.generate(classes)
    # the classes with an f shell: entries onto the loop-form fused kernel
    loop = [c for c in builder_classes(LOOP_LMAX) if max(c) > lmax]
    if loop:
        p = gen_hip.GEN / "f_loop_entries.hip"
        text = gen_hip.emit_loop_entries(loop)
        if not p.exists() or p.read_text() != text:
            p.write_text(text)
        files.append(p)
    built = set(classes) | set(loop)
    gen_hip.write_header(tuple(c for c in builder_classes(LOOP_LMAX) if c in built))
    p = gen_hip.GEN / "f_any_table.h"
    text = gen_hip.emit_any_table(tuple(c for c in builder_classes(LOOP_LMAX) if c in built))
    if not p.exists() or p.read_text() != text:
        p.write_text(text)
    # the split J / K kernels: one per orbit representative through d
    from . import orbits
    reps = orbits.canonical_classes(4 * SPLIT_LMAX, SPLIT_LMAX)
    split = gen_hip.generate_split([(r, sum(1 << g for g in orbits.orientations(r))) for r in reps])
    gen_hip.write_split_header(reps)
    files += split
    files.append(gen_hip.write_df())   # density fitting: metric and three-center entries
    files.append(gen_hip.write_gf())   # the fused two-electron gradient's entries
    mmd = gen_mmd.generate(MMD_LMAX, k_lmax=MMD_K_LMAX)
    gen_mmd.write_header(MMD_LMAX, MMD_K_LMAX)
    return [CSRC / "fj_runtime.hip", CSRC / "fj_mlp.hip", CSRC / "fj_plan.hip", CSRC / "fj_kloop.hip",
            CSRC / "fj_noalloc.hip", CSRC / "fj_dfloop.hip", CSRC / "fj_stream.hip", CSRC / "fj_gfloop.hip", *files, *mmd]


def _compile(src: Path, extra: list[str], force: bool = False) -> tuple[Path, float, str]:
    OBJ.mkdir(exist_ok=True)
    obj = OBJ / (src.stem + ".o")
    deps = [src, CSRC / "fj_device.h", HERE.parent / "include" / "fj_mi355x.h",
            CSRC / "gen" / "fj_boys_table.h"]
    hdr = src.parent / (src.stem.replace("_r32", "") + ".h")
    if hdr.exists():
        deps.append(hdr)
    if src.name in ("fj_kloop.hip", "fj_dfloop.hip", "fj_gfloop.hip"):
        deps.append(CSRC / "fj_kloop.h")
    if src.name == "fj_gfloop.hip":
        deps.append(CSRC / "fj_gfloop.h")
    if src.name == "fj_dfloop.hip":
        deps.append(CSRC / "fj_dfloop.h")
    if src.name == "fj_kloop.hip":
        deps.append(CSRC / "fj_floop.h")
    if src.name == "fj_noalloc.hip":
        deps.append(CSRC / "gen" / "f_any_table.h")
    if not force and obj.exists() and all(obj.stat().st_mtime >= d.stat().st_mtime
                                          for d in deps if d.exists()):
        return obj, 0.0, ""
    t = time.time()
    r = subprocess.run([HIPCC, *FLAGS, *extra, "-Rpass-analysis=kernel-resource-usage", "-c",
                        str(src), "-o", str(obj)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr[-4000:]}")
    obj.with_suffix(".res").write_text(r.stderr)
    return obj, time.time() - t, r.stderr


def build(lmax: int = LMAX, jobs: int | None = None, extra: list[str] | None = None,
          verbose: bool = True, force: bool = False) -> Path:
    """Generate, compile (every unit when `force`, else the out-of-date ones)
    and link the library; refuse it if a kernel needs too much scratch."""
    srcs = sources(lmax)
    jobs = jobs or min(16, os.cpu_count() or 4)
    t0 = time.time()
    # biggest units first so the long compiles start early
    srcs.sort(key=lambda p: -p.stat().st_size)
    with ThreadPoolExecutor(jobs) as ex:
        res = list(ex.map(lambda s: _compile(s, extra or [], force), srcs))
    if verbose:
        slow = sorted(res, key=lambda r: -r[1])[:3]
        print(f"[fj build] {len(res)} units in {time.time() - t0:.1f}s; slowest: " +
              ", ".join(f"{p.name} {dt:.1f}s" for p, dt, _ in slow), flush=True)
    objs = [str(p) for p, _, _ in res]
    check_resources(objs)
    tmp = LIB.with_suffix(".so.tmp")
    r = subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    return LIB


# Private scratch per lane that every built kernel stays within: the runtime
# reserves it for every wavefront a queue can hold, on every queue in use.
SCRATCH_MAX = 6 * 1024


def resources(objs=None) -> dict:
    """{kernel: (VGPRs, scratch bytes per lane, occupancy)} from hipcc's
    kernel resource remarks of the given objects (default: every one in
    the build folder)."""
    import re
    out = {}
    files = [Path(o).with_suffix(".res") for o in objs] if objs is not None else sorted(OBJ.glob("*.res"))
    for f in files:
        if not f.exists():
            continue
        name = None
        cur: dict = {}
        for line in f.read_text().splitlines():
            m = re.search(r"Function Name: (\S+)", line)
            if m:
                name = m.group(1)
                cur = out.setdefault(name, {})
                continue
            m = re.search(r"(VGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]): (\d+)", line)
            if m and name:
                cur[m.group(1).split()[0]] = int(m.group(2))
            m = re.search(r"Dynamic Stack: (True|False)", line)
            if m and name:
                cur["DynamicStack"] = m.group(1) == "True"
    return out


def check_resources(objs=None) -> None:
    """Refuse a build in which any kernel asks for more private scratch per
    lane than SCRATCH_MAX (the runtime fails such launches)."""
    res = resources(objs)
    big = {k: v["ScratchSize"] for k, v in res.items() if v.get("ScratchSize", 0) > SCRATCH_MAX}
    if big:
        raise RuntimeError(f"kernels over {SCRATCH_MAX} B of scratch per lane: {big}")
    dyn = [k for k, v in res.items() if v.get("DynamicStack")]
    if dyn:
        raise RuntimeError(f"kernels with a dynamic stack (unknown scratch size): {dyn}")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lmax", type=int, default=LMAX)
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    print(build(a.lmax, a.jobs))


if __name__ == "__main__":
    main()
