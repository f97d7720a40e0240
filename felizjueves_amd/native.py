"""ctypes binding of libfj_mi355x.so (the gfx950 HIP library).

The product path runs only through this library: if it is missing, or no
GPU is present, every call raises -- there is no CPU fallback.
Device buffers are torch tensors on ``cuda:0`` (PyTorch-ROCm is used for
device memory and streams only).
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libfj_mi355x.so"

_lib = None


class FjSide(C.Structure):
    """Mirror of ``FjSide`` in include/fj_mi355x.h."""

    _fields_ = [("prim", C.c_void_p), ("off", C.c_void_p), ("k", C.c_int),
                ("sep", C.c_void_p), ("G", C.c_void_p), ("sh", C.c_void_p),
                ("ao", C.c_void_p), ("n", C.c_int)]


class FjSys(C.Structure):
    """Mirror of ``FjSys`` in include/fj_mi355x.h (device pointers)."""

    _fields_ = [("D", C.c_void_p), ("Dsh", C.c_void_p), ("F", C.c_void_p),
                ("n_ao", C.c_int), ("n_sh", C.c_int)]


def sys_table(systems) -> "torch.Tensor":
    """A device array of FjSys from (D, Dsh, F) device tensors (square)."""
    arr = (FjSys * len(systems))()
    for i, (D, Dsh, F) in enumerate(systems):
        arr[i] = FjSys(D.data_ptr(), Dsh.data_ptr(), F.data_ptr(), D.shape[0], Dsh.shape[0])
    raw = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
    return to_dev(raw)


class FjError(RuntimeError):
    pass


ERRORS = {9001: "FJ_ERR_THREADS", 9002: "FJ_ERR_DEGREE", 9003: "FJ_ERR_SHAPE",
          9004: "FJ_ERR_CLASS", 9005: "FJ_ERR_ARG"}


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise FjError(f"{what} returned {rc} ({ERRORS.get(rc, 'hipError_t')})")


def lib():
    """Load the library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FjError(f"{LIB_PATH} is missing: run `python -m felizjueves_amd.build` "
                          "(the HIP library is the only implementation)")
        # torch ships its own HIP runtime (SONAME libamdhip64.so.7); load it
        # first so this library binds to that one instance instead of pulling
        # a second HIP/HSA runtime into the process.
        import torch  # noqa: F401
        L = C.CDLL(str(LIB_PATH))
        L.fj_version.restype = C.c_int
        _lib = L
    return _lib


def has(symbol: str) -> bool:
    return hasattr(lib(), symbol)


def stream_handle(stream=None) -> int:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def ptr(t) -> int | None:
    if t is None:
        return None
    return int(t.data_ptr())


def device():
    import torch
    if not torch.cuda.is_available():
        raise FjError("no GPU: the Fock build runs only on the gfx950 device")
    return torch.device("cuda:0")


def to_dev(a, dtype=None):
    import torch
    t = torch.as_tensor(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device()).contiguous()


def call_f(cls, threads: int, args: dict, F, stream=None) -> None:
    """One launch of the upstream-ABI entry fj_f_<cls> on device tensors
    (`args` by upstream parameter name)."""
    la, lb, lc, ld = cls
    fn = getattr(lib(), "fj_f_" + "".join(map(str, cls)))
    a = args
    vals = [C.c_void_p(stream_handle(stream)), C.c_int(threads), C.c_int(a["n_bra"]),
            C.c_int(a["n_ket"]), C.c_int(a["k_ab"]), C.c_int(a["k_cd"]), C.c_int(a["n_sh"]),
            C.c_int(a["n_ao"]), C.c_int(a["diag"]), C.c_double(a["tau"]),
            C.c_void_p(ptr(a["bra_prim"]))]
    if lb:
        vals.append(C.c_void_p(ptr(a["AB"])))
    vals += [C.c_void_p(ptr(a["G_ab"])), C.c_void_p(ptr(a["bra_sh"])),
             C.c_void_p(ptr(a["bra_ao"])), C.c_void_p(ptr(a["ket_prim"]))]
    if ld:
        vals.append(C.c_void_p(ptr(a["CD"])))
    vals += [C.c_void_p(ptr(a["G_cd"])), C.c_void_p(ptr(a["ket_sh"])),
             C.c_void_p(ptr(a["ket_ao"])), C.c_void_p(ptr(a["Dsh"])),
             C.c_void_p(ptr(a["D"])), C.c_void_p(ptr(F))]
    fn.restype = C.c_int
    check(fn(*vals), fn.__name__)


def call_fx(cls, threads: int, precision: int, bra: FjSide, ket: FjSide, work, n_work: int,
            sys, diag: int, tau: float, stream=None, ket_list=None, ws=None) -> None:
    """One launch of fj_fx_<cls>: `work` a device int32 [n_work, 4] tensor of
    (bra, first ket, system, ket end), `sys` a device FjSys table; with
    `ket_list` the (first, end) range indexes that device int32 list; `ws`
    a device tensor used as the launch workspace (classes that need one)."""
    fn = getattr(lib(), "fj_fx_" + "".join(map(str, cls)))
    fn.restype = C.c_int
    ws_bytes = 0 if ws is None else ws.numel() * ws.element_size()
    rc = fn(C.c_void_p(stream_handle(stream)), C.c_int(threads), C.c_int(precision),
            C.byref(bra), C.byref(ket), C.c_void_p(ptr(work)), C.c_int(n_work),
            C.c_void_p(ptr(ket_list)), C.c_void_p(ptr(sys)), C.c_int(diag), C.c_double(tau),
            C.c_void_p(ptr(ws)), C.c_longlong(ws_bytes))
    check(rc, fn.__name__)


def call_fxm(cls, threads: int, bra: FjSide, ket: FjSide, work, n_work: int, n_fp64: int,
             sys, diag: int, tau: float, stream=None, ket_list=None, ws=None, tier=None) -> None:
    """One mixed-tier launch of fj_fxm_<cls>: with `tier` (a device int32
    array, 0 FP64 / 1 FP32 per work item) each item in its tier, else items
    [0, n_fp64) in FP64 and the rest in FP32 (otherwise as call_fx)."""
    fn = getattr(lib(), "fj_fxm_" + "".join(map(str, cls)))
    fn.restype = C.c_int
    ws_bytes = 0 if ws is None else ws.numel() * ws.element_size()
    rc = fn(C.c_void_p(stream_handle(stream)), C.c_int(threads), C.byref(bra), C.byref(ket),
            C.c_void_p(ptr(work)), C.c_int(n_work), C.c_int(n_fp64), C.c_void_p(ptr(tier)),
            C.c_void_p(ptr(ket_list)), C.c_void_p(ptr(sys)), C.c_int(diag), C.c_double(tau),
            C.c_void_p(ptr(ws)), C.c_longlong(ws_bytes))
    check(rc, fn.__name__)


class FjTierSel(C.Structure):
    """Mirror of ``FjTierSel`` in include/fj_mi355x.h (device pointers)."""

    _fields_ = [("work", C.c_void_p), ("ket_list", C.c_void_p), ("n_items", C.c_int),
                ("diag", C.c_int), ("tau", C.c_double), ("amp_bra", C.c_void_p),
                ("amp_ket", C.c_void_p), ("est_scale", C.c_double), ("fp32_error", C.c_double),
                ("fp32_bound", C.c_double), ("feat", C.c_void_p), ("inside", C.c_void_p),
                ("n_feat", C.c_int), ("d_mean", C.c_double), ("d_std", C.c_double),
                ("d_lo", C.c_double), ("d_hi", C.c_double), ("Wb", C.c_void_p), ("P", C.c_void_p),
                ("margin", C.c_double), ("log_budget", C.c_double), ("tier", C.c_void_p)]


class FjTierJob(C.Structure):
    """Mirror of ``FjTierJob`` in include/fj_mi355x.h."""

    _fields_ = [("bra", FjSide), ("ket", FjSide), ("sel", FjTierSel), ("first", C.c_int),
                ("pad", C.c_int)]


def tier_jobs(jobs) -> "torch.Tensor":
    """A device array of FjTierJob from (bra FjSide, ket FjSide, FjTierSel,
    first item) tuples."""
    arr = (FjTierJob * len(jobs))()
    for i, (b, k, sel, first) in enumerate(jobs):
        arr[i] = FjTierJob(b, k, sel, first, 0)
    return to_dev(np.frombuffer(bytes(arr), dtype=np.uint8).copy())


def tier_select_jobs(jobs_dev, n_jobs: int, n_items: int, sys, net=None, stream=None) -> None:
    """Every mixed launch's per-build tiers in one launch (fj_tier_select_jobs)."""
    fn = lib().fj_tier_select_jobs
    fn.restype = C.c_int
    check(fn(C.c_void_p(stream_handle(stream)), C.c_int(n_jobs), C.c_void_p(ptr(jobs_dev)),
             C.c_int(n_items), C.c_void_p(ptr(sys)), C.c_int(net.F if net is not None else 0),
             C.c_void_p(ptr(net.Wb) if net is not None else None),
             C.c_void_p(ptr(net.P) if net is not None else None)), "fj_tier_select_jobs")


def tier_select(bra: FjSide, ket: FjSide, sys, sel: FjTierSel, stream=None) -> None:
    """The per-build tiers of one mixed launch, on the device (fj_tier_select)."""
    fn = lib().fj_tier_select
    fn.restype = C.c_int
    check(fn(C.c_void_p(stream_handle(stream)), C.byref(bra), C.byref(ket), C.c_void_p(ptr(sys)),
             C.byref(sel)), "fj_tier_select")


def workspace_bytes(cls, threads: int, n_work: int) -> int:
    """Bytes of launch workspace fj_fx_<cls> uses for n_work items."""
    fn = lib().fj_fx_workspace_bytes
    fn.restype = C.c_longlong
    return int(fn(*[C.c_int(x) for x in cls], C.c_int(threads), C.c_int(n_work)))


def close(F, stream=None) -> None:
    L = lib()
    L.fj_close.restype = C.c_int
    check(L.fj_close(C.c_void_p(stream_handle(stream)), C.c_int(F.shape[0]),
                     C.c_void_p(ptr(F))), "fj_close")


def density_blocks(shell_ao, shell_size, D, Dsh, stream=None) -> None:
    L = lib()
    L.fj_density_blocks.restype = C.c_int
    check(L.fj_density_blocks(C.c_void_p(stream_handle(stream)), C.c_int(Dsh.shape[0]),
                              C.c_int(D.shape[0]), C.c_void_p(ptr(shell_ao)),
                              C.c_void_p(ptr(shell_size)), C.c_void_p(ptr(D)),
                              C.c_void_p(ptr(Dsh))), "fj_density_blocks")


def boys(M: int, T) -> np.ndarray:
    """Evaluate the device Boys function at host values T (test entry)."""
    import torch
    Td = to_dev(np.asarray(T, dtype=np.float64))
    out = torch.zeros((Td.shape[0], M + 1), dtype=torch.float64, device=device())
    L = lib()
    L.fj_boys_eval.restype = C.c_int
    check(L.fj_boys_eval(C.c_void_p(stream_handle()), C.c_int(M), C.c_int(Td.shape[0]),
                         C.c_void_p(ptr(Td)), C.c_void_p(ptr(out))), "fj_boys_eval")
    torch.cuda.synchronize()
    return out.cpu().numpy()
