"""ctypes binding of the native planner (csrc/fj_plan.hip, host code):
per class pair the screen, the tier split of the bound rule, the lanes and
the work items, the same as FockBuilder's Python planner produces."""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import native


class FjPlanSide(C.Structure):
    _fields_ = [("G", C.c_void_p), ("deg", C.c_void_p), ("sh", C.c_void_p), ("mol", C.c_void_p),
                ("amp", C.c_void_p), ("n", C.c_int)]


class FjPlanResult(C.Structure):
    _fields_ = [("work", C.POINTER(C.c_int) * 2), ("n_work", C.c_longlong * 2),
                ("kets", C.POINTER(C.c_int) * 2), ("n_kets", C.c_longlong * 2),
                ("n_quartets", C.c_double * 2), ("n_prim", C.c_double * 2),
                ("flops", C.c_double * 2), ("n_cold", C.c_longlong)]


class Side:
    """Host arrays of one side table, kept alive with their FjPlanSide."""

    def __init__(self, t):
        self.G = np.ascontiguousarray(t.G, dtype=np.float64)
        self.deg = np.ascontiguousarray(t.deg, dtype=np.int32)
        self.sh = np.ascontiguousarray(t.sh, dtype=np.int32)
        self.mol = np.ascontiguousarray(t.mol, dtype=np.int32)
        self.amp = np.ascontiguousarray(t.amp, dtype=np.float64)
        self.s = FjPlanSide(self.G.ctypes.data, self.deg.ctypes.data, self.sh.ctypes.data,
                            self.mol.ctypes.data, self.amp.ctypes.data, len(self.G))


class Systems:
    """The per-system density block maxima, as the planner reads them."""

    def __init__(self, Dsh_list):
        self.Dsh = [np.ascontiguousarray(d, dtype=np.float64) for d in Dsh_list]
        self.ptrs = (C.c_void_p * len(self.Dsh))(*[d.ctypes.data for d in self.Dsh])
        self.n_sh = np.array([d.shape[0] for d in self.Dsh], dtype=np.int32)


def plan_pairs(bra: Side, ket: Side, ket_table, systems: Systems, *, diag: int, tau: float,
               mixed: bool, fp32_error: float, fp32_bound, est_scale: float, threads: int,
               lane_flops: float, per_prim: float, per_quartet: float, max_prim: int):
    """One class pair. lane_flops <= 0: counting only. Returns, per tier,
    (work [n, 4] int32 or None, ket list int32 or None, (n_quartets,
    n_prim, flops)), and the number of cold items (the last ones of the
    FP64 tier's work)."""
    n_mol = len(systems.Dsh)
    lo = np.zeros(n_mol, dtype=np.int32)
    hi = np.zeros(n_mol, dtype=np.int32)
    for m, (a, b) in ket_table.mol_range.items():
        lo[m], hi[m] = a, b
    fn = native.lib().fj_plan_pairs
    fn.restype = C.c_int
    res = FjPlanResult()
    rc = fn(C.byref(bra.s), C.byref(ket.s), C.c_void_p(lo.ctypes.data), C.c_void_p(hi.ctypes.data),
            systems.ptrs, C.c_void_p(systems.n_sh.ctypes.data), C.c_int(diag), C.c_double(tau),
            C.c_int(int(mixed)), C.c_double(fp32_error),
            C.c_double(-1.0 if fp32_bound is None else fp32_bound), C.c_double(est_scale),
            C.c_int(threads), C.c_double(lane_flops), C.c_double(per_prim), C.c_double(per_quartet),
            C.c_int(max_prim), C.byref(res))
    native.check(rc, "fj_plan_pairs")
    out = []
    try:
        for t in (0, 1):
            stats = (int(res.n_quartets[t]), int(res.n_prim[t]), float(res.flops[t]))
            if lane_flops > 0 and res.n_work[t]:
                w = np.ctypeslib.as_array(res.work[t], shape=(res.n_work[t] * 4,)).reshape(-1, 4).copy()
                k = np.ctypeslib.as_array(res.kets[t], shape=(res.n_kets[t],)).copy()
                out.append((w, k, stats))
            else:
                out.append((None, None, stats))
        n_cold = int(res.n_cold)
    finally:
        native.lib().fj_plan_free(C.byref(res))
    return out, n_cold
