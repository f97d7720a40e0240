"""The generated F kernels run unchanged on the CPU (CPU only): the kernel
source of a class is compiled with g++ against an emulation of the device
header (tests/emu: one thread per lane, barriers for __syncthreads and the
wavefront shuffles) under AddressSanitizer, launched on walker launches of
the two-atom s,s,p,d,f system, and compared with the oracle's f-launch.

This checks the generator's kernels -- including the f classes, which the
library does not build for the GPU -- for out-of-bounds accesses and for
parity, independently of the GPU toolchain."""

import os
import shutil
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

from felizjueves_amd import gen_hip
from felizjueves_amd import pairs as P
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent
EMU = ROOT / "tests" / "emu"

# (class, launch index): fused f classes, an f workspace class with both
# HRRs (its VRR groups as separate functions), a d workspace class
CASES = [((0, 0, 3, 3), 0), ((0, 1, 3, 3), 2), ((0, 0, 2, 3), 1), ((1, 2, 2, 2), 0)]


def _system():
    from felizjueves_amd.basis import Basis
    from felizjueves_amd.systems import atom_shells, closed_shell_density
    rng = np.random.default_rng(21)
    shells = atom_shells(rng, (0.0, 0.0, 0.0), (0, 0, 1, 2, 3)) + \
        atom_shells(rng, (0.9, -1.1, 1.2), (0, 0, 1, 2, 3))
    basis = Basis(shells)
    return basis, closed_shell_density(rng, basis.n_ao, 6)


def _dump(path: Path, a: dict, cls) -> None:
    with open(path, "wb") as f:
        np.array([a["n_bra"], a["n_ket"], a["k_ab"], a["k_cd"], a["n_sh"], a["n_ao"], a["diag"]],
                 dtype=np.int32).tofile(f)
        np.array([a["tau"]], dtype=np.float64).tofile(f)

        def w(x, dt):
            x = np.ascontiguousarray(x, dtype=dt).ravel()
            np.array([x.size], dtype=np.int64).tofile(f)
            x.tofile(f)
        w(a["bra_prim"], np.float64)
        if cls[1]:
            w(a["AB"], np.float64)
        w(a["G_ab"], np.float64)
        w(a["bra_sh"], np.int32)
        w(a["bra_ao"], np.int32)
        w(a["ket_prim"], np.float64)
        if cls[3]:
            w(a["CD"], np.float64)
        w(a["G_cd"], np.float64)
        w(a["ket_sh"], np.int32)
        w(a["ket_ao"], np.int32)
        w(a["Dsh"], np.float64)
        w(a["D"], np.float64)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
@pytest.mark.parametrize("cls,which", CASES)
def test_generated_kernel_on_cpu_vs_oracle(cls, which):
    basis, D = _system()
    plan = P.f_launches(basis, D, tau=0.0, n=1)
    Dsh = P.density_blocks(basis, D)
    L = [L for L in plan.launches if tuple(L.cls) == cls][which]
    a = P.launch_args(basis, plan, L, D, Dsh)
    name = "".join(map(str, cls))
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        gen_hip.generate([cls], td / "gen")
        shutil.copy(EMU / "fj_device_emu.h", td / "fj_device.h")
        exe = td / f"emu_{name}"
        cmd = ["g++", "-std=c++20", "-O1", "-g0", "-fsanitize=address", f"-I{ROOT / 'include'}",
               f"-I{td}", f'-DCLS_HEADER="gen/f_{name}.h"', f"-DHAS_AB={int(cls[1] > 0)}",
               f"-DHAS_CD={int(cls[3] > 0)}", f"-DWS_DOUBLES={gen_hip.ws_doubles(cls)}",
               f"-DKERNEL=fj_f_{name}_k", str(EMU / "emu_driver.cpp"), "-o", str(exe), "-lpthread"]
        subprocess.run(cmd, check=True, cwd=td, capture_output=True)
        _dump(td / "case.bin", a, cls)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
        r = subprocess.run([str(exe), str(td / "case.bin"), str(td / "F.bin")], cwd=td, env=env,
                           capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-3000:]
        F = np.fromfile(td / "F.bin").reshape(basis.n_ao, basis.n_ao)
    Fo = np.zeros_like(F)
    O.f_launch(L.cls, a, Fo)
    err = np.abs(F - Fo).max()
    assert err <= 1e-12 * max(1.0, np.abs(Fo).max()), (cls, err)
