"""Contracted Cartesian shells and the basis, in the upstream convention.

Convention (upstream `src/felizjueves/ao/shell.py:1-40`): a shell's
coefficients multiply UNNORMALISED Cartesian primitives
``x^i y^j z^k exp(-alpha r^2)``, ``i + j + k = l``, centred on the shell; any
normalisation is already inside the coefficients. Components are enumerated
``i`` descending, then ``j`` descending (``ao/shell.py:68-74``): xx, xy, xz,
yy, yz, zz at l = 2. AO offsets follow shell order.

The basis is held as flat numpy arrays, the layout the C-ABI and the oracle
read:

    shell_l[n_sh] (i4)   shell_centre[n_sh,3]   shell_nprim[n_sh] (i4)
    shell_pstart[n_sh] (i4, first primitive)    exps[], coefs[]
    shell_ao[n_sh] (i4, first AO)               n_ao
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path
from typing import Sequence

import numpy as np

L_MAX = 3  # f shells: (ff|ff) is the top fused-F class built here


def n_cart(l: int) -> int:  # noqa: E741
    """(l+1)(l+2)/2 Cartesian components of one shell."""
    return (l + 1) * (l + 2) // 2


def cartesians(l: int) -> tuple[tuple[int, int, int], ...]:  # noqa: E741
    """Every (i, j, k), i + j + k = l: i descending, then j descending
    (upstream `ao/shell.py:68-74`)."""
    return tuple((i, j, l - i - j) for i in range(l, -1, -1)
                 for j in range(l - i, -1, -1))


def _dfact_odd(l: int) -> int:  # noqa: E741
    out = 1
    for k in range(1, 2 * l, 2):
        out *= k
    return out


def primitive_norm(l: int, alpha: float) -> float:  # noqa: E741
    """N with unit self-overlap of the axis component x^l exp(-alpha r^2)
    (upstream `ao/shell.py:84-95`):
    ``(2a/pi)^(3/4) (4a)^(l/2) / sqrt((2l-1)!!)``."""
    return ((2.0 * alpha / math.pi) ** 0.75 * (4.0 * alpha) ** (l / 2.0)
            / math.sqrt(_dfact_odd(l)))


@dataclass(frozen=True)
class Shell:
    l: int  # noqa: E741
    centre: tuple[float, float, float]
    exponents: tuple[float, ...]
    coefficients: tuple[float, ...]

    def __post_init__(self) -> None:
        if type(self.l) is not int or not 0 <= self.l <= L_MAX:
            raise ValueError(f"l = {self.l!r} is not an int in 0..{L_MAX}")
        if len(self.centre) != 3:
            raise ValueError("centre is not three floats")
        if not self.exponents or len(self.exponents) != len(self.coefficients):
            raise ValueError("exponents and coefficients differ in length or are empty")
        if not all(a > 0 and math.isfinite(a) for a in self.exponents):
            raise ValueError("exponents must be finite and > 0")

    @property
    def n_cart(self) -> int:
        return n_cart(self.l)


class Basis:
    """Shells in AO order, flattened to the arrays the device path reads."""

    def __init__(self, shells: Sequence[Shell]):
        shells = tuple(shells)
        if not shells:
            raise ValueError("a basis needs at least one shell")
        self.shells = shells
        self.n_sh = len(shells)
        self.shell_l = np.array([s.l for s in shells], dtype=np.int32)
        self.shell_centre = np.array([s.centre for s in shells], dtype=np.float64)
        self.shell_nprim = np.array([len(s.exponents) for s in shells], dtype=np.int32)
        self.shell_pstart = np.concatenate(
            [[0], np.cumsum(self.shell_nprim)[:-1]]).astype(np.int32)
        self.exps = np.array([x for s in shells for x in s.exponents], dtype=np.float64)
        self.coefs = np.array([x for s in shells for x in s.coefficients], dtype=np.float64)
        sizes = np.array([s.n_cart for s in shells], dtype=np.int32)
        self.shell_ao = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
        self.shell_size = sizes
        self.n_ao = int(sizes.sum())

    @property
    def l_max(self) -> int:
        return int(self.shell_l.max())

    @classmethod
    def from_arrays(cls, shell_l, shell_centre, shell_nprim, exps, coefs) -> "Basis":
        shells, k = [], 0
        for l, c, n in zip(shell_l, shell_centre, shell_nprim):
            n = int(n)
            shells.append(Shell(int(l), tuple(float(x) for x in c),
                                tuple(float(x) for x in exps[k:k + n]),
                                tuple(float(x) for x in coefs[k:k + n])))
            k += n
        return cls(shells)


GOLDEN = Path(__file__).resolve().parent.parent / "tests" / "golden"


@dataclass(frozen=True)
class Fixture:
    """One libcint golden set (tests/golden/<name>.npz)."""

    name: str
    basis: Basis
    density: np.ndarray
    arrays: dict

    def eri_blocks(self) -> dict[tuple[int, int, int, int], np.ndarray]:
        if "eri_keys" not in self.arrays:
            return {}
        keys, off, vals = (self.arrays["eri_keys"], self.arrays["eri_off"],
                           self.arrays["eri_vals"])
        out = {}
        for i, q in enumerate(keys):
            q = tuple(int(x) for x in q)
            shape = tuple(self.basis.shells[s].n_cart for s in q)
            out[q] = vals[off[i]:off[i + 1]].reshape(shape)
        return out


def load_fixture(name: str, root: Path = GOLDEN) -> Fixture:
    with np.load(root / f"{name}.npz", allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    basis = Basis.from_arrays(arrays["shell_l"], arrays["shell_centre"],
                              arrays["shell_nprim"], arrays["exps"], arrays["coefs"])
    return Fixture(name, basis, arrays["density"], arrays)


FOCK_FIXTURES = ("h2_sto3g", "h2o_sto3g", "h2o_631g", "h2o_ccpvdz",
                 "h10_chain_ccpvdz", "co_ccpvtz")
