"""Walk a pair-stream plan with an injected launcher (upstream
`eri/fock/walk_f.py:306-412`): check the plan, form the density block
maxima, stage each batch pair, run every launch into one F, close F.

The launcher is injected, as upstream's ``run``: ``run(cls, grid, block,
args, F)`` adds one launch's contribution into ``F`` (its own array type);
``runner.zeros(n_ao)``, ``runner.put(args)`` and ``runner.host(F)`` move
data. ``GpuRunner`` drives the per-class C-ABI entries of the HIP library.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import pairs as P
from .basis import Basis

THREADS = 64


@dataclass
class FWalk:
    F: np.ndarray                # closed
    n_launches: int
    n_quartets_launched: int


def _check_plan(plan: P.FPlan, basis: Basis) -> None:
    want = set(P.unique_pairs(basis))
    seen = [(p.a, p.b) for b in plan.batches for p in b.pairs] + [tuple(x) for x in plan.deleted]
    if len(seen) != len(set(seen)) or set(seen) != want:
        raise ValueError("walk_f refuses the plan: the pairs are not every canonical "
                         "shell pair exactly once")
    nb = len(plan.batches)
    got = [(L.bra, L.ket) for L in plan.launches] + [tuple(x) for x in plan.screened]
    if len(got) != len(set(got)) or set(got) != {(i, j) for i in range(nb) for j in range(i, nb)}:
        raise ValueError("walk_f refuses the plan: launches and screened are not every "
                         "batch pair I <= J once")
    for b in plan.batches:
        for p in b.pairs:
            if (int(basis.shell_l[p.a]), int(basis.shell_l[p.b])) != (b.la, b.lb) or p.degree != b.degree:
                raise ValueError(f"walk_f refuses the plan: pair {(p.a, p.b)} is not of its batch")


PRIM_CHUNK = 2   # the sub-pair packing the walker uses by default (pairs.pack)


def walk_f(plan: P.FPlan, basis: Basis, D: np.ndarray, *, run, threads: int = THREADS,
           prim_chunk: int = PRIM_CHUNK) -> FWalk:
    """prim_chunk: each side's pairs packed as sub-pairs of at most that many
    primitive pairs (pairs.pack; 0 packs whole pairs, the upstream's
    packing): the same F to rounding, with the long primitive loops of
    contracted pairs spread over lanes."""
    _check_plan(plan, basis)
    if not np.array_equal(D, D.T):
        raise ValueError("the density must be exactly symmetric")
    Dsh = P.density_blocks(basis, D)
    F = run.zeros(basis.n_ao)
    nq = 0
    for L in plan.launches:
        args = P.launch_args(basis, plan, L, D, Dsh, prim_chunk)
        nb, nk = len(plan.batches[L.bra].pairs), len(plan.batches[L.ket].pairs)
        nq += nb * (nb + 1) // 2 if args["diag"] else nb * nk
        run(L.cls, args["n_bra"], (threads, 1), run.put(args), F)
    Fh = run.host(F)
    return FWalk(P.close(Fh), len(plan.launches), nq)


class GpuRunner:
    """Launch each batch pair through the upstream-ABI entry fj_f_<cls>."""

    def __init__(self):
        from . import native
        self.native = native
        self.dev = native.device()

    def zeros(self, n):
        import torch
        return torch.zeros((n, n), dtype=torch.float64, device=self.dev)

    def put(self, args: dict) -> dict:
        out = {}
        for k, v in args.items():
            if isinstance(v, np.ndarray):
                out[k] = self.native.to_dev(v)
            else:
                out[k] = v
        return out

    def __call__(self, cls, grid, block, args, F):
        self.native.call_f(cls, block[0], args, F)

    def host(self, F) -> np.ndarray:
        import torch
        torch.cuda.synchronize()
        return F.cpu().numpy()
