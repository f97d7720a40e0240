"""The 8-fold ERI symmetry orbits of an ordered class and their cheapest
member (upstream `hgp/orbits.py:1-184`), for the split J / K build's
canonical class mode.

The real ERI's permutational symmetry,

    (ab|cd) = (ba|cd) = (ab|dc) = (ba|dc) = (cd|ab) = (dc|ab) = (cd|ba) = (dc|ba),

makes the ordered classes that one quartet visits in a split build (every
ORDERED quartet) one orbit. One kernel is built per orbit, on its
representative, and each ordered quartet is deposited through the
permutation that carries its class onto the representative.

`GROUP` lists the eight elements as shell-position maps g: the class
`apply(g, cls)` has `cls[g[j]]` in position j. The table and its order are
the upstream's, so a `perm` index names the same orientation on both sides.
The representative is the member with the smallest `rank_key` = (HRR
statements with the boundary reads, pruned VRR nodes, the class tuple), the
upstream's order; the DAGs are `hgp.py`'s, whose counts equal the
upstream's (e.g. (dp|ss) 34 HRR statements against (pd|ss)'s 60, (ds|ps)
43 VRR nodes against (ps|ds)'s 52).
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass
from functools import lru_cache

from .basis import cartesians
from .hgp import hrr_dag, vrr_dag

GROUP: tuple[tuple[int, int, int, int], ...] = (
    (0, 1, 2, 3), (1, 0, 2, 3), (0, 1, 3, 2), (1, 0, 3, 2),
    (2, 3, 0, 1), (3, 2, 0, 1), (2, 3, 1, 0), (3, 2, 1, 0),
)


def _check(cls) -> tuple[int, int, int, int]:
    cls = tuple(cls)
    if len(cls) != 4 or any(not isinstance(x, int) or x < 0 for x in cls):
        raise ValueError(f"class {cls!r} is not four ints >= 0 (la, lb, lc, ld)")
    return cls


def apply(g, cls) -> tuple[int, int, int, int]:
    """g . cls: position j holds cls[g[j]]."""
    return tuple(cls[i] for i in g)


def orbit(cls) -> tuple[tuple[int, int, int, int], ...]:
    """The distinct members of cls's orbit, sorted."""
    cls = _check(cls)
    return tuple(sorted({apply(g, cls) for g in GROUP}))


@lru_cache(maxsize=None)
def rank_key(cls) -> tuple:
    """(HRR statements incl. boundary reads, VRR nodes, cls)."""
    cls = _check(cls)
    order, _, _, _ = hrr_dag(cls)
    return (len(order), vrr_dag(cls).n_nodes, cls)


def representative(cls) -> tuple[int, int, int, int]:
    """The orbit member with the smallest rank_key."""
    return min(orbit(cls), key=rank_key)


def inverse(g) -> tuple[int, int, int, int]:
    inv = [0] * 4
    for j, i in enumerate(g):
        inv[i] = j
    return tuple(inv)


@dataclass(frozen=True)
class Orientation:
    """How the ordered class `cls` is served by its representative `rep`:
    `shells[j]` is the ordered quartet's position that goes to the
    representative's position j (rep == apply(shells, cls)); `axes` is its
    inverse, the ordered block being rep_block.transpose(axes)."""

    cls: tuple
    rep: tuple
    shells: tuple
    axes: tuple

    @property
    def perm(self) -> int:
        """The index of `shells` in GROUP (the kernels' perm argument)."""
        return GROUP.index(self.shells)

    @property
    def gather(self) -> tuple[int, ...]:
        """Flat indices: ordered.flat[k] == rep.flat[gather[k]], both blocks
        [a][b][c][d] row-major."""
        dims = [len(cartesians(x)) for x in self.rep]
        strides = [dims[1] * dims[2] * dims[3], dims[2] * dims[3], dims[3], 1]
        out = []
        for idx in itertools.product(*(range(len(cartesians(x))) for x in self.cls)):
            out.append(sum(idx[self.shells[j]] * strides[j] for j in range(4)))
        return tuple(out)


@lru_cache(maxsize=None)
def to_representative(cls) -> Orientation:
    """The first GROUP element carrying cls onto representative(cls)."""
    cls = _check(cls)
    rep = representative(cls)
    g = next(g for g in GROUP if apply(g, cls) == rep)
    return Orientation(cls, rep, g, inverse(g))


def ordered_classes(lmax_total: int, lmax_shell: int | None = 2) -> tuple:
    """Every ordered (la, lb, lc, ld) with total <= lmax_total and each
    l <= lmax_shell (None: no cap), lexicographic."""
    top = lmax_total if lmax_shell is None else min(lmax_total, lmax_shell)
    return tuple(c for c in itertools.product(range(top + 1), repeat=4) if sum(c) <= lmax_total)


def canonical_classes(lmax_total: int, lmax_shell: int | None = 2) -> tuple:
    """One representative per orbit of ordered_classes, sorted."""
    return tuple(sorted({representative(c) for c in ordered_classes(lmax_total, lmax_shell)}))


def orientations(cls) -> tuple[int, ...]:
    """The GROUP indices a representative's kernel deposits through: the
    orientation of each distinct member of the orbit, sorted."""
    return tuple(sorted({to_representative(m).perm for m in orbit(cls)}))
