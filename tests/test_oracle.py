"""Pins of the CPU oracle and of the host-side math (CPU only)."""

import itertools
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

from felizjueves_amd import pairs as P
from felizjueves_amd.basis import cartesians, load_fixture
from felizjueves_amd.systems import config1
from oracle import oracle as O


# --- Boys -------------------------------------------------------------------

def kummer_50(m: int, T: float) -> float:
    """F_m(T) = e^-T sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)), 50 digits."""
    getcontext().prec = 50
    T = Decimal(T)
    term = Decimal(1) / (2 * m + 1)
    s = term
    k = 1
    while True:
        term = term * 2 * T / (2 * m + 2 * k + 1)
        s += term
        if term < s * Decimal("1e-45"):
            break
        k += 1
    return float((-T).exp() * s)


@pytest.mark.parametrize("T", [0.0, 1e-9, 0.3, 1.7, 9.9, 24.5, 44.0, 46.0, 120.0, 700.0])
def test_boys_vs_50_digit_series(T):
    F = O.boys(16, T)
    for m in (0, 1, 5, 8, 12, 16):
        ref = kummer_50(m, T)
        assert abs(F[m] - ref) <= 2e-15 * ref, (m, T, F[m], ref)


def test_boys_closed_forms():
    F0 = O.boys(12, 0.0)
    assert np.allclose(F0, [1.0 / (2 * m + 1) for m in range(13)], rtol=1e-15, atol=0)
    for T in (0.01, 2.0, 30.0, 60.0, 1e4):
        f = O.boys(0, T)[0]
        assert abs(f - 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))) < 1e-15 * f
    # the downward recursion holds between orders
    for T in (0.5, 7.0, 41.0):
        F = O.boys(10, T)
        for m in range(10):
            assert abs(F[m] - (2 * T * F[m + 1] + math.exp(-T)) / (2 * m + 1)) < 1e-14 * F[m]


# --- ERIs against libcint -----------------------------------------------------

@pytest.mark.parametrize("name", ["h2o_631g", "hf_ccpvtz"])
@pytest.mark.parametrize("method", ["hgp", "mmd"])
def test_eri_blocks_vs_libcint(name, method):
    fx = load_fixture(name)
    blocks = fx.eri_blocks()
    assert blocks
    ls = set()
    for q, ref in blocks.items():
        got = O.eri(fx.basis, q, method)
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(got - ref).max() <= 2e-13 * scale, (q, np.abs(got - ref).max(), scale)
        ls.add(tuple(int(fx.basis.shell_l[s]) for s in q))
    if name == "hf_ccpvtz":
        assert (3, 3, 3, 3) in ls  # (ff|ff) is pinned


@pytest.mark.parametrize("name", ["h2_sto3g", "h2o_sto3g", "h2o_631g"])
def test_schwarz_vs_libcint(name):
    fx = load_fixture(name)
    Q = O.schwarz(fx.basis)
    ref = fx.arrays["int2e_diag"]
    assert np.abs(Q * Q - ref).max() < 1e-12 * max(1.0, ref.max())


@pytest.mark.parametrize("name", ["h2o_sto3g", "h2o_631g", "h2o_ccpvdz"])
def test_ordered_sweep_vs_libcint(name):
    fx = load_fixture(name)
    J, K = O.fock_ordered(fx.basis, fx.density)
    assert np.abs(J - fx.arrays["vj"]).max() < 1e-12 * max(1.0, np.abs(fx.arrays["vj"]).max())
    assert np.abs(K - fx.arrays["vk"]).max() < 1e-12 * max(1.0, np.abs(fx.arrays["vk"]).max())


def test_mmd_equals_hgp_random_quartets():
    s = config1()
    rng = np.random.default_rng(3)
    for _ in range(30):
        q = tuple(int(x) for x in rng.integers(0, s.basis.n_sh, 4))
        a, b = O.eri(s.basis, q, "hgp"), O.eri(s.basis, q, "mmd")
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max()), q


# --- the screening bound bounds -------------------------------------------------

def test_contracted_bound_bounds_every_eri():
    """|(ab|cd)| <= G_ab G_cd with G the contracted primitive-pair bound
    (upstream ao/pairs.py: the Cauchy-Schwarz argument)."""
    s = config1()
    b = s.basis
    G = {}
    for cls, (prs, pp) in P.pair_tables(b).items():
        for k, pr in enumerate(prs):
            G[pr] = math.fsum(pp.bound[pp.pair == k])
    prs = list(G)
    rng = np.random.default_rng(5)
    for _ in range(60):
        p, q = prs[rng.integers(len(prs))], prs[rng.integers(len(prs))]
        x = O.eri(b, (*p, *q), "mmd")
        assert np.abs(x).max() <= G[p] * G[q] * (1 + 1e-12), (p, q)


# --- Hermite E: the pointwise identity -----------------------------------------

def _hermite_poly(t: int, x: np.ndarray) -> np.ndarray:
    h0, h1 = np.ones_like(x), 2 * x
    if t == 0:
        return h0
    for k in range(1, t):
        h0, h1 = h1, 2 * x * h1 - 2 * k * h0
    return h1


def test_hermite_expansion_identity():
    """x_A^i x_B^j e^{-a x_A^2} e^{-b x_B^2} = sum_t E^{ij}_t Lambda_t(x),
    Lambda_t = zeta^{t/2} H_t(sqrt(zeta)(x - P)) e^{-zeta (x-P)^2}
    (upstream mmd/hermite.py's defining identity)."""
    a, b, A, B = 0.7, 1.3, -0.4, 0.9
    z = a + b
    Pc = (a * A + b * B) / z
    x = np.linspace(-2, 2, 41)
    for i in range(4):
        for j in range(4):
            E = P.e_axis(i, j, np.array([z]), np.array([Pc - A]), np.array([Pc - B]))[0, i, j]
            lhs = (x - A) ** i * (x - B) ** j * np.exp(-a * (x - A) ** 2 - b * (x - B) ** 2)
            rhs = sum(E[t] * z ** (t / 2) * _hermite_poly(t, math.sqrt(z) * (x - Pc))
                      * np.exp(-z * (x - Pc) ** 2) for t in range(i + j + 1))
            assert np.abs(lhs - rhs).max() < 1e-13


# --- the digestion weights, derived -------------------------------------------

def test_digestion_weights_derived_from_the_orbit():
    """Enumerate the 8 index permutations preserving a real ERI; count, per
    unordered block, the members that deposit IN ORDER: Coulomb 2 x 1,
    exchange 1 x (1/2) -- the constants the kernels and the oracle use."""
    gens = [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]
    group = {(0, 1, 2, 3)}
    todo = [(0, 1, 2, 3)]
    while todo:
        g = todo.pop()
        for h in gens:
            gh = tuple(g[i] for i in h)
            if gh not in group:
                group.add(gh)
                todo.append(gh)
    assert len(group) == 8
    blocks = [(0, 1), (2, 3), (0, 2), (0, 3), (1, 2), (1, 3)]
    inorder = {blk: [0, 0] for blk in blocks}
    for g in group:
        for kind, out in ((0, (g[0], g[1])), (1, (g[0], g[2]))):
            if out == tuple(sorted(out)):
                inorder[tuple(sorted(out))][kind] += 1
    w = {blk: 1.0 * inorder[blk][0] - 0.5 * inorder[blk][1] for blk in blocks}
    assert w[(0, 1)] == w[(2, 3)] == 2.0
    assert all(w[blk] == -0.5 for blk in blocks[2:])
    # degeneracy = 1/|Stab| = (1/2)^{[a=b] + [c=d] + [P=Q]}
    for q in itertools.product(range(3), repeat=4):
        a, b_, c, d = q
        if not (a <= b_ and c <= d and (a, b_) <= (c, d)):
            continue
        stab = sum(1 for g in group if tuple(q[i] for i in g) == q)
        want = 0.5 ** ((a == b_) + (c == d) + ((a, b_) == (c, d)))
        assert 1.0 / stab == want, q
