"""Write tests/golden/c2_40atom_fp64_reference.npz: the FP64 two-electron
Fock matrix G = J - K/2 of the 40-atom working-size cluster (seed 2, tau
1e-12), computed by the CPU oracle's HGP launches over fully oriented pair
tables (every primitive pair built on the centre it sits near), the upper
triangle stored row by row. That oracle build agrees with the oracle's MMD
build (which has no HRR and so no cancellation) to 1.7e-12 on the 10-atom
cluster; the MMD build itself is too slow at 40 atoms. About 5 minutes on
8 CPU threads. Test infrastructure: run by hand, not by the tests.

    python tests/golden/make_c2_reference.py
"""

import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from felizjueves_amd import fock  # noqa: E402
from felizjueves_amd.fock import FockBuilder  # noqa: E402
from felizjueves_amd.systems import config2  # noqa: E402
from test_builder_oracle import oracle_build  # noqa: E402

OUT = Path(__file__).resolve().parent / "c2_40atom_fp64_reference.npz"


def main() -> None:
    s = config2()
    t = time.time()
    fock.ORIENT_AMP_MAX = 0.0          # split every mixed pair by centre
    fb = FockBuilder(s.basis, s.D, tau=1e-12, planner="python")
    G = oracle_build(fb, s.D)
    iu = np.triu_indices(s.basis.n_ao)
    np.savez_compressed(OUT, G_upper=G[iu], n_ao=s.basis.n_ao, system=s.name, tau=1e-12)
    print(f"{OUT.name}: n_ao {s.basis.n_ao}, {time.time() - t:.0f} s", flush=True)


if __name__ == "__main__":
    main()
