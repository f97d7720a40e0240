"""The fused-F pair stream on the device (upstream `eri/fock/launch_f.py:
390-470`, `f_launches`): the same plan as `pairs.f_launches` -- the
primitive-pair bounds, their sums G, the prune below eps = tau / (2 G_max),
the density block maxima, the keys G dmax, the eq. 29 magnitudes, the bins
(la, lb, degree, magnitude) with their pairs by key descending, and the
launches I <= J whose magnitude indices sum below n -- computed by the
library's kernels (`fj_stream_plan`, `fj_stream_launches`), so an SCF
caller re-plans for each density without a host pass over the pairs.

    f_launches_device(basis, D, *, tau, n, ket_break=False) -> (FPlan, timings)

The host hands over the basis arrays, the unique shell pairs (index work),
the magnitudes tau^(i/n) and D on the device, and reads back the plan; the
FPlan it returns is the one `walk.walk_f` takes (the pairs' staged records
are filled from the basis for the kept primitive pairs, as the upstream's
walker packs them).
"""

from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import native
from . import pairs as P
from .basis import Basis


class FjStreamIn(C.Structure):
    _fields_ = [("n_pair", C.c_int), ("n_prim", C.c_int), ("n_sh", C.c_int), ("n", C.c_int),
                ("ket_break", C.c_int), ("tau", C.c_double)] + \
               [(k, C.c_void_p) for k in ("pair_ab", "prim_off", "shell_l", "centre", "nprim", "pstart",
                                         "exps", "coefs", "Dsh", "mags")]


class FjStreamOut(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("g", "G", "dmax", "key", "degree", "mag", "keep", "order",
                                         "batch_start", "batch_mag", "batch_gmax_bits", "gmax_bits",
                                         "eps", "n_batch", "n_kept")]


class DeviceStream:
    """The device tables of one basis for repeated plans (one per density)."""

    def __init__(self, basis: Basis):
        import torch
        self.basis = basis
        pairs = P.unique_pairs(basis)
        self.pairs = pairs
        npr = np.array([basis.shell_nprim[a] * basis.shell_nprim[b] for a, b in pairs], dtype=np.int64)
        self.prim_off = np.concatenate([[0], np.cumsum(npr)]).astype(np.int32)
        self.n_pair, self.n_prim = len(pairs), int(self.prim_off[-1])
        dev = native.device()
        td = native.to_dev
        self.d = {"pair_ab": td(np.array(pairs, dtype=np.int32)), "prim_off": td(self.prim_off),
                  "shell_l": td(basis.shell_l.astype(np.int32)), "centre": td(basis.shell_centre),
                  "nprim": td(basis.shell_nprim.astype(np.int32)),
                  "pstart": td(basis.shell_pstart.astype(np.int32)), "exps": td(basis.exps),
                  "coefs": td(basis.coefs), "shell_ao": td(basis.shell_ao.astype(np.int32)),
                  "shell_size": td(basis.shell_size.astype(np.int32))}
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        u64 = dict(dtype=torch.int64, device=dev)
        np_, nk = self.n_pair, self.n_prim
        self.o = {"g": torch.empty(nk, **f64), "G": torch.empty(np_, **f64),
                  "dmax": torch.empty(np_, **f64), "key": torch.empty(np_, **f64),
                  "degree": torch.empty(np_, **i32), "mag": torch.empty(np_, **i32),
                  "keep": torch.empty(nk, **i32), "order": torch.empty(np_, **i32),
                  "batch_start": torch.empty(np_, **i32), "batch_mag": torch.empty(np_, **i32),
                  "batch_gmax_bits": torch.empty(np_, **u64), "gmax_bits": torch.empty(1, **u64),
                  "eps": torch.empty(1, **f64), "n_batch": torch.empty(1, **i32),
                  "n_kept": torch.empty(1, **i32)}
        self.Dsh = torch.empty((basis.n_sh, basis.n_sh), **f64)
        lib = native.lib()
        lib.fj_stream_workspace_bytes.restype = C.c_longlong
        wsb = int(lib.fj_stream_workspace_bytes(C.c_int(np_), C.c_int(nk)))
        self.ws = torch.empty(max(1, wsb), dtype=torch.uint8, device=dev)
        self.out = FjStreamOut(*[native.ptr(self.o[k]) for k, _ in FjStreamOut._fields_])
        self._lt = {}

    def plan(self, D_dev, *, tau: float, n: int, ket_break: bool = False, stream=None):
        """(device outputs, host copies) of the plan for the device density D_dev."""
        import torch
        lib = native.lib()
        native.density_blocks(self.d["shell_ao"], self.d["shell_size"], D_dev, self.Dsh, stream)
        mags = P.magnitudes(tau, n) if tau > 0 else tuple([1.0] + [0.0] * (n - 1))
        self.mags = mags
        mags_d = native.to_dev(np.array(mags, dtype=np.float64))
        fin = FjStreamIn(self.n_pair, self.n_prim, self.basis.n_sh, n, int(ket_break), float(tau),
                         *[native.ptr(self.d[k]) for k in ("pair_ab", "prim_off", "shell_l", "centre",
                                                           "nprim", "pstart", "exps", "coefs")],
                         native.ptr(self.Dsh), native.ptr(mags_d))
        lib.fj_stream_plan.restype = C.c_int
        native.check(lib.fj_stream_plan(C.c_void_p(native.stream_handle(stream)), C.byref(fin),
                                        C.byref(self.out), C.c_void_p(native.ptr(self.ws)),
                                        C.c_longlong(self.ws.numel())), "fj_stream_plan")
        nb = int(self.o["n_batch"].cpu()[0])
        launches = np.zeros((0, 2), dtype=np.int32)
        if nb:
            if nb not in self._lt:
                self._lt = {nb: (torch.empty(nb + 1, dtype=torch.int32, device=mags_d.device),
                                 torch.empty((nb * (nb + 1) // 2, 2), dtype=torch.int32,
                                             device=mags_d.device),
                                 torch.empty(1, dtype=torch.int32, device=mags_d.device))}
            row, lst, cnt = self._lt[nb]
            lib.fj_stream_launches.restype = C.c_int
            native.check(lib.fj_stream_launches(C.c_void_p(native.stream_handle(stream)), C.c_int(nb),
                                                C.c_int(n), C.c_void_p(native.ptr(self.o["batch_mag"])),
                                                C.c_void_p(native.ptr(row)), C.c_void_p(native.ptr(lst)),
                                                C.c_void_p(native.ptr(cnt))), "fj_stream_launches")
            launches = lst[:int(cnt.cpu()[0])].cpu().numpy()
        host = {k: v.cpu().numpy() for k, v in self.o.items()}
        host["n_batch"], host["launches"] = nb, launches
        return host


def to_fplan(ds: DeviceStream, h: dict, tau: float, n: int) -> P.FPlan:
    """The upstream FPlan from the device plan's outputs (no arithmetic:
    the values are the device's; the kept primitive pairs' staged records
    come from the basis, as the walker packs them)."""
    b = ds.basis
    nk = int(h["n_kept"][0])
    big = int(h["degree"].max()) if len(h["degree"]) else 0
    if big > P.MAX_PRIM_PAIRS:
        raise ValueError(f"a pair keeps {big} primitive pairs, above MAX_PRIM_PAIRS = "
                         f"{P.MAX_PRIM_PAIRS}")
    order = h["order"][:nk]
    starts = list(h["batch_start"][:h["n_batch"]]) + [nk]
    mags = ds.mags
    tables = {}
    batches = []
    for bi in range(h["n_batch"]):
        idx = order[starts[bi]:starts[bi + 1]]
        ps = []
        for p in idx:
            a, bb = ds.pairs[p]
            k0 = int(ds.prim_off[p])
            keep = h["keep"][k0:k0 + int(h["degree"][p])].astype(np.int64)
            cls = (int(b.shell_l[a]), int(b.shell_l[bb]))
            if (a, bb) not in tables:
                pp = P.primitive_pairs(b, [(a, bb)])
                tables[(a, bb)] = np.concatenate([pp.P, pp.PA, pp.zeta[:, None], pp.U[:, None]], axis=1)
            rec = np.ascontiguousarray(tables[(a, bb)][keep])
            ps.append(P.FPair(a, bb, cls[0], cls[1], int(ds.prim_off[p + 1] - k0), keep,
                              float(h["G"][p]), float(h["dmax"][p]), float(h["key"][p]), int(h["mag"][p]),
                              rec))
        m = int(h["batch_mag"][bi])
        batches.append(P.FBatch(ps[0].la, ps[0].lb, ps[0].degree, m, mags[m], ps))
    launches = [P.FLaunch(int(i), int(j), (batches[i].la, batches[i].lb, batches[j].la, batches[j].lb))
                for i, j in h["launches"]]
    kept = {(int(i), int(j)) for i, j in h["launches"]}
    screened = [(i, j) for i in range(len(batches)) for j in range(i, len(batches)) if (i, j) not in kept]
    deleted = [ds.pairs[p] for p in range(ds.n_pair) if h["degree"][p] == 0]
    gmax = float(np.frombuffer(h["gmax_bits"].tobytes(), dtype=np.float64)[0])
    return P.FPlan(tau, n, mags, float(h["eps"][0]), gmax, batches, launches, screened, deleted,
                   ds.n_prim)


def f_launches_device(basis: Basis, D, *, tau: float, n: int, ket_break: bool = False,
                      stream=None, ds: DeviceStream | None = None):
    """(FPlan, {"device_s": plan on the device incl. read-back, "host_s":
    building the FPlan objects}) for the density D (host or device)."""
    import torch
    if n < 1:
        raise ValueError("n must be >= 1")
    D_dev = D if isinstance(D, torch.Tensor) else native.to_dev(np.asarray(D, dtype=np.float64))
    ds = ds or DeviceStream(basis)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    h = ds.plan(D_dev, tau=tau, n=n, ket_break=ket_break, stream=stream)
    t1 = time.perf_counter()
    plan = to_fplan(ds, h, tau, n)
    t2 = time.perf_counter()
    return plan, {"device_s": t1 - t0, "host_s": t2 - t1}
