"""fj_f_noalloc (the upstream-ABI fused-F launch on a fixed device array,
no allocation) against fj_f_<cls> (library-owned workspace) through the
C-ABI: every walker launch of a 10-atom cluster, then of the smaller
three-atom system (the second call on the shared array with a smaller
input). Both run the same FP64 kernels; the no-alloc launch cuts the
persistent grid to the workgroups its array holds, so only the order of
F's atomic additions differs: the bar is 1e-13 relative per launch."""

import ctypes as C

import numpy as np
import pytest

from felizjueves_amd import native
from felizjueves_amd import pairs as P
from felizjueves_amd.systems import config1, config2

pytestmark = pytest.mark.gpu

ARGS = ("n_bra", "n_ket", "k_ab", "k_cd", "n_sh", "n_ao", "diag")


def _noalloc(cls, threads, a, F):
    fn = native.lib().fj_f_noalloc
    fn.restype = C.c_int
    p = lambda k: C.c_void_p(native.ptr(a[k]) if k in a else None)  # noqa: E731
    rc = fn(C.c_void_p(native.stream_handle()), *[C.c_int(x) for x in cls], C.c_int(threads),
            *[C.c_int(a[k]) for k in ARGS], C.c_double(a["tau"]), p("bra_prim"), p("AB"), p("G_ab"),
            p("bra_sh"), p("bra_ao"), p("ket_prim"), p("CD"), p("G_cd"), p("ket_sh"), p("ket_ao"),
            p("Dsh"), p("D"), C.c_void_p(native.ptr(F)))
    native.check(rc, "fj_f_noalloc")


def _compare(s, threads=64):
    import torch
    plan = P.f_launches(s.basis, s.D, tau=1e-12, n=3)
    Dsh = P.density_blocks(s.basis, s.D)
    n = s.basis.n_ao
    worst, n_launch, staged = 0.0, 0, 0
    for L in plan.launches:
        a = {k: (native.to_dev(v) if isinstance(v, np.ndarray) else v)
             for k, v in P.launch_args(s.basis, plan, L, s.D, Dsh).items()}
        Fa = torch.zeros((n, n), dtype=torch.float64, device=native.device())
        Fb = torch.zeros_like(Fa)
        native.call_f(L.cls, threads, a, Fa)
        _noalloc(L.cls, threads, a, Fb)
        torch.cuda.synchronize()
        scale = max(1.0, Fa.abs().max().item())
        err = (Fa - Fb).abs().max().item() / scale
        worst = max(worst, err)
        n_launch += 1
        staged += native.workspace_bytes(L.cls, threads, 1) > 0
    print(f"{s.name}: {n_launch} launches ({staged} with a workspace class), worst relative "
          f"|F_noalloc - F| {worst:.2e}", flush=True)
    assert staged > 0
    assert worst <= 1e-13


def test_noalloc_matches_the_allocating_entry_twice():
    """Two calls in a row on the shared array: a 10-atom cluster, then the
    smaller three-atom system."""
    assert native.lib().fj_f_noalloc_scratch_bytes() == 64 << 20
    _compare(config2(seed=11, n_atoms=10))
    _compare(config1())


def test_noalloc_refuses_a_bad_class():
    """A class outside the built ones is refused before any launch."""
    fn = native.lib().fj_f_noalloc
    fn.restype = C.c_int
    rc = fn(C.c_void_p(None), 4, 0, 0, 0, C.c_int(64), *[C.c_int(1)] * 7, C.c_double(0.0),
            *[C.c_void_p(None)] * 13)
    assert rc == 9004   # FJ_ERR_CLASS
