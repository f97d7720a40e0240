"""Seeded synthetic molecules and densities for the benchmark workloads.

Every system is built from `numpy.random.default_rng(seed)`; the seed is part
of the workload's name. Geometry is in bohr (1 Angstrom = ANGSTROM bohr).

Shell recipe per atom (the same for every workload): contracted Cartesian
shells ``s, s, p, d`` (plus ``f`` where the workload asks for it), each with a
contraction depth K ~ U{1..6}, exponents log-uniform in [0.1, 1e4], positive
raw weights U(0.1, 1) times `primitive_norm`, and the contraction scaled so
its axis component has unit self-overlap.

Density: ``D = 2 C C^T`` with ``C`` the Q factor of a QR decomposition of a
seeded Gaussian ``n_ao x nocc`` matrix (orthonormal columns), so D is
symmetric and idempotent up to the factor 2 of a closed-shell density.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .basis import Basis, Shell, primitive_norm, _dfact_odd

ANGSTROM = 1.8897261254578281  # bohr per Angstrom

EXP_LO, EXP_HI = 0.1, 1.0e4
K_MAX = 6


def _self_overlap_axis(l: int, a: float, b: float) -> float:  # noqa: E741
    """<x^l e^{-a r^2} | x^l e^{-b r^2}> for unnormalised primitives."""
    p = a + b
    return (math.pi / p) ** 1.5 * _dfact_odd(l) / (2.0 * p) ** l


def random_shell(rng: np.random.Generator, l: int, centre) -> Shell:  # noqa: E741
    k = int(rng.integers(1, K_MAX + 1))
    exps = np.exp(rng.uniform(math.log(EXP_LO), math.log(EXP_HI), size=k))
    exps = np.sort(exps)[::-1]
    raw = rng.uniform(0.1, 1.0, size=k)
    c = np.array([w * primitive_norm(l, a) for w, a in zip(raw, exps)])
    s = sum(c[i] * c[j] * _self_overlap_axis(l, exps[i], exps[j])
            for i in range(k) for j in range(k))
    c = c / math.sqrt(s)
    return Shell(l, tuple(float(x) for x in centre), tuple(float(x) for x in exps),
                 tuple(float(x) for x in c))


def atom_shells(rng, centre, ls=(0, 0, 1, 2)) -> list[Shell]:
    return [random_shell(rng, l, centre) for l in ls]


def chain_geometry(rng, n_atoms: int, lo=0.9, hi=1.1) -> np.ndarray:
    """Atom 0 at the origin, each next one U(lo, hi) Angstrom from the
    previous in a uniformly random direction."""
    pos = [np.zeros(3)]
    for _ in range(n_atoms - 1):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        pos.append(pos[-1] + v * rng.uniform(lo, hi) * ANGSTROM)
    return np.array(pos)


def cluster_geometry(rng, n_atoms: int, min_dist=1.0, volume_per_atom=10.0) -> np.ndarray:
    """Uniform in a cube of volume ``n_atoms * volume_per_atom`` Angstrom^3,
    rejection to a minimum distance of ``min_dist`` Angstrom."""
    side = (n_atoms * volume_per_atom) ** (1.0 / 3.0)
    pos: list[np.ndarray] = []
    tries = 0
    while len(pos) < n_atoms:
        tries += 1
        if tries > 1_000_000:
            raise RuntimeError("rejection sampling did not place every atom")
        x = rng.uniform(0.0, side, size=3)
        if all(np.linalg.norm(x - y) >= min_dist for y in pos):
            pos.append(x)
    return np.array(pos) * ANGSTROM


def closed_shell_density(rng, n_ao: int, nocc: int) -> np.ndarray:
    nocc = min(nocc, n_ao)
    q, _ = np.linalg.qr(rng.normal(size=(n_ao, nocc)))
    D = 2.0 * q @ q.T
    return 0.5 * (D + D.T)  # exactly symmetric


@dataclass(frozen=True)
class System:
    name: str
    seed: int
    basis: Basis
    D: np.ndarray
    coords: np.ndarray


def build_system(name: str, seed: int, coords: np.ndarray, ls, nocc: int,
                 rng: np.random.Generator) -> System:
    shells = []
    for c in coords:
        shells += atom_shells(rng, c, ls)
    basis = Basis(shells)
    D = closed_shell_density(rng, basis.n_ao, nocc)
    return System(name, seed, basis, D, coords)


def config1(seed: int = 1) -> System:
    """3 atoms 0.9-1.1 Angstrom apart; s,s,p,d per atom (n_ao 33); nocc 5."""
    rng = np.random.default_rng(seed)
    coords = chain_geometry(rng, 3)
    return build_system(f"c1-3atom-spd-seed{seed}", seed, coords, (0, 0, 1, 2), 5, rng)


def config2(seed: int = 2, n_atoms: int = 40) -> System:
    """40-atom random cluster, min 1.0 Angstrom; s,s,p,d per atom (n_ao 440);
    nocc 100."""
    rng = np.random.default_rng(seed)
    coords = cluster_geometry(rng, n_atoms)
    return build_system(f"c2-{n_atoms}atom-spd-seed{seed}", seed, coords,
                        (0, 0, 1, 2), 100, rng)


def batch_molecule(i: int, lmax: int = 3) -> System:
    """Molecule i of the batched workload: atoms log-uniform in [8, 80]
    (seed 4 + i), random cluster, shells s,s,p,d(,f); nocc = n_atoms."""
    seed = 4 + i
    rng = np.random.default_rng(seed)
    n_atoms = int(round(math.exp(rng.uniform(math.log(8), math.log(80)))))
    coords = cluster_geometry(rng, n_atoms)
    ls = tuple(l for l in (0, 0, 1, 2, 3) if l <= lmax)  # noqa: E741
    return build_system(f"c4-mol{i}-{n_atoms}atom-seed{seed}", seed, coords, ls,
                        n_atoms, rng)
