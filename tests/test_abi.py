"""The C-ABI: the library loads and exports every entry include/fj_mi355x.h
declares (CPU only: no compute call), and the header's generated section
matches the generator."""

import ctypes
import re
from pathlib import Path

import pytest

from felizjueves_amd import gen_hip, native
from felizjueves_amd.hgp import builder_classes, fock_classes

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "fj_mi355x.h"


def declared() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*int\s+(fj_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entries():
    names = declared()
    assert "fj_version" in names and "fj_close" in names
    for cls in fock_classes(2):
        n = "".join(map(str, cls))
        assert f"fj_f_{n}" in names and f"fj_fx_{n}" in names
    # the builder's classes (both pair orders, through (ff|ff): the f classes
    # on the loop-form kernel) have the builder entries; the upstream-ABI
    # entry exists for the canonical classes only
    for cls in builder_classes(3):
        n = "".join(map(str, cls))
        assert f"fj_fx_{n}" in names and f"fj_fxm_{n}" in names
        assert (f"fj_f_{n}" in names) == (cls in fock_classes(3))


def test_header_matches_generator():
    cur = HEADER.read_text()
    classes = tuple(c for c in builder_classes(3)
                    if f"fj_fx_{''.join(map(str, c))}(" in cur)
    assert gen_hip.header_text(classes, cur) == cur


@pytest.mark.skipif(not native.LIB_PATH.exists(), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(native.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_upstream_parameter_lists():
    """fj_f_<cls> takes AB only when lb > 0 and CD only when ld > 0, in the
    upstream order (eri/fock/levers/seams.py:228-235)."""
    text = HEADER.read_text()
    for cls in fock_classes(3):
        n = "".join(map(str, cls))
        m = re.search(rf"int fj_f_{n}\((.*?)\);", text, flags=re.S)
        params = [p.strip().split()[-1].lstrip("*") for p in m.group(1).split(",")]
        want = ["stream", "threads", "n_bra", "n_ket", "k_ab", "k_cd", "n_sh", "n_ao", "diag",
                "tau", "bra_prim", *(["AB"] if cls[1] else []), "G_ab", "bra_sh", "bra_ao",
                "ket_prim", *(["CD"] if cls[3] else []), "G_cd", "ket_sh", "ket_ao", "Dsh",
                "D", "F"]
        assert params == want, (n, params)
