"""HIP generator for the McMurchie-Davidson kernel families (gfx950, wave64):
Schwarz, DQ, [p|q], JAB and K.

Restated mathematics (upstream files cited per family):

* Hermite coefficients per axis (upstream `mmd/hermite.py:45-79`):
  ``E^{00}_0 = exp(zeta X_PA X_PB)``,
  ``E^{i+1,j}_t = E^{ij}_{t-1}/(2 zeta) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}``,
  the ``j`` raise with ``X_PB``; ``E^{ab}_t = E_x E_y E_z``, structurally
  zero where ``t_d > a_d + b_d``. Rows ``(a, b)`` in ``cartesians(la) x
  cartesians(lb)`` order, columns in ``hermite_order(la + lb)``.
* The Hermite integrals (upstream `reference/pq.py:54-133`):
  ``[0]^(m) = (2/sqrt(pi)) theta (-2 theta^2)^m F_m(theta^2 |PQ|^2)``,
  ``[r]^(m) = PQ_d [r - 1_d]^(m+1) + (r_d - 1) [r - 2_d]^(m+1)`` on the first
  nonzero axis d of r, and ``[p|q] = (-1)^|q| [p + q]^(0)``.

The dressings (``s = c_a c_b (pi/zeta)^{3/2}``) are host-supplied per record,
as upstream (`ao/pairs.py:238`, `ao/density.py:151`).
"""

from __future__ import annotations

from pathlib import Path

from .basis import cartesians, n_cart
from .gen_hip import GEN, _Emit

AX = "xyz"


def hermite_order(L: int):
    return [(tx, ty, n - tx - ty) for n in range(L + 1)
            for tx in range(n, -1, -1) for ty in range(n - tx, -1, -1)]


def n_herm(L: int) -> int:
    return (L + 1) * (L + 2) * (L + 3) // 6


def _tname(t) -> str:
    return f"{t[0]}_{t[1]}_{t[2]}"


# ---------------------------------------------------------------------------
# pieces


def emit_tower(E: _Emit, L: int, ind: int, need=None, f32: bool = False) -> dict:
    """[t]^(0) for |t| <= L (or the set `need`) as locals r<t>; expects
    doubles PQx PQy PQz, rho and T in scope. Returns {t: name}.

    f32: the J mixed lever (upstream eri/j_coulomb/levers/mixed.py:82-191):
    PQ, theta = sqrt(rho), T, the [0]^(m) and every [t]^(m) in float; F_m
    evaluated in double on the widened float T and narrowed."""
    targets = list(need) if need is not None else hermite_order(L)
    nodes: dict = {}
    todo = [(t, 0) for t in targets]
    while todo:
        t, m = todo.pop()
        if (t, m) in nodes:
            continue
        if t == (0, 0, 0):
            nodes[(t, m)] = ()
            continue
        d = next(i for i, c in enumerate(t) if c)
        t1 = tuple(c - (i == d) for i, c in enumerate(t))
        ch = [(t1, m + 1)]
        if t[d] >= 2:
            ch.append((tuple(c - 2 * (i == d) for i, c in enumerate(t)), m + 1))
        nodes[(t, m)] = (d, ch)
        todo += ch
    mmax = max(m for _, m in nodes)
    real = "float" if f32 else "double"
    if f32:
        E("const float PQxf = (float)PQx, PQyf = (float)PQy, PQzf = (float)PQz;", ind)
        E("const float thf = (float)sqrt(rho), t2f = thf * thf;", ind)
        E("const float Tf = t2f * ((PQxf * PQxf + PQyf * PQyf) + PQzf * PQzf);", ind)
        E(f"double Fm[{mmax + 1}];", ind)
        E(f"fj_boys<{mmax}>((double)Tf, Fm);", ind)
        E("const float m2r = -2.0f * t2f;", ind)
        E("float wm = (float)FJ_TWO_OVER_SQRT_PI * thf;", ind)
        for m in range(mmax + 1):
            if ((0, 0, 0), m) in nodes:
                E(f"const float R0_0_0m{m} = wm * (float)Fm[{m}];", ind)
            if m < mmax:
                E("wm *= m2r;", ind)
    else:
        E(f"double Fm[{mmax + 1}];", ind)
        E(f"fj_boys<{mmax}>(T, Fm);", ind)
        E("const double w0 = FJ_TWO_OVER_SQRT_PI * sqrt(rho), m2r = -2.0 * rho;", ind)
        E("double wm = w0;", ind)
        for m in range(mmax + 1):
            if ((0, 0, 0), m) in nodes:
                E(f"const double R0_0_0m{m} = wm * Fm[{m}];", ind)
            if m < mmax:
                E("wm *= m2r;", ind)
    sfx = "f" if f32 else ""
    order = sorted(nodes, key=lambda n: (-n[1], sum(n[0])))
    for n in order:
        r = nodes[n]
        t, m = n
        if not r:
            continue
        d, ch = r
        nm = f"R{_tname(t)}m{m}"
        e = f"PQ{AX[d]}{sfx} * R{_tname(ch[0][0])}m{ch[0][1]}"
        if len(ch) > 1:
            k = t[d] - 1
            e += (f" + R{_tname(ch[1][0])}m{ch[1][1]}" if k == 1 else
                  f" + {float(k)}{sfx} * R{_tname(ch[1][0])}m{ch[1][1]}")
        E(f"const {real} {nm} = {e};", ind)
    if f32:
        return {t: f"(double)R{_tname(t)}m0" for t in targets}
    return {t: f"R{_tname(t)}m0" for t in targets}


def emit_eaxis(E: _Emit, la: int, lb: int, zeta: str, pa: str, pb: str, pre: str,
               ind: int) -> None:
    """Per-axis E^{ij}_t, i <= la, j <= lb, as locals {pre}{axis}{i}_{j}_{t}."""
    E(f"const double {pre}h = 0.5 / {zeta};", ind)
    for d in range(3):
        ax = AX[d]
        nm = lambda i, j, t: f"{pre}{ax}{i}_{j}_{t}"  # noqa: E731
        E(f"const double {nm(0, 0, 0)} = exp({zeta} * {pa}{ax} * {pb}{ax});", ind)
        for i in range(la + 1):
            if i > 0:
                for t in range(i + 1):
                    terms = []
                    if t >= 1:
                        terms.append(f"{pre}h * {nm(i - 1, 0, t - 1)}")
                    if t <= i - 1:
                        terms.append(f"{pa}{ax} * {nm(i - 1, 0, t)}")
                    if t + 1 <= i - 1:
                        terms.append(f"{float(t + 1)} * {nm(i - 1, 0, t + 1)}")
                    E(f"const double {nm(i, 0, t)} = {' + '.join(terms)};", ind)
            for j in range(1, lb + 1):
                for t in range(i + j + 1):
                    terms = []
                    if t >= 1:
                        terms.append(f"{pre}h * {nm(i, j - 1, t - 1)}")
                    if t <= i + j - 1:
                        terms.append(f"{pb}{ax} * {nm(i, j - 1, t)}")
                    if t + 1 <= i + j - 1:
                        terms.append(f"{float(t + 1)} * {nm(i, j - 1, t + 1)}")
                    E(f"const double {nm(i, j, t)} = {' + '.join(terms)};", ind)


def e_entry(pre: str, a, b, t) -> str | None:
    """The product E_x E_y E_z of row (a, b), column t; None if structurally 0."""
    if any(t[d] > a[d] + b[d] for d in range(3)):
        return None
    return " * ".join(f"{pre}{AX[d]}{a[d]}_{b[d]}_{t[d]}" for d in range(3))


def _rows(la, lb):
    return [(a, b) for a in cartesians(la) for b in cartesians(lb)]


def _vec(prefix: str, arr: str, idx: str, ind: int, E: _Emit) -> None:
    E(f"const double {prefix}x = {arr}[3 * {idx}], {prefix}y = {arr}[3 * {idx} + 1], "
      f"{prefix}z = {arr}[3 * {idx} + 2];", ind)


# ---------------------------------------------------------------------------
# Schwarz (upstream eri/kernel_schwarz.py:1-101): one thread per shell pair,
# (mn|mn) = sum_{i,k recs} s_i s_k sum_p E_i[r,p] sum_q (-1)^|q| [p+q] E_k[r,q]


def emit_schwarz(la: int, lb: int) -> str:
    n = f"{la}{lb}"
    L = 2 * (la + lb)
    E = _Emit()
    E(f"// generated by felizjueves_amd/gen_mmd.py -- Schwarz class ({la}{lb})", 0)
    E('#include "../fj_device.h"', 0)
    E("", 0)
    E(f"__global__ void __launch_bounds__(256) fj_schwarz_{n}_k(int n_pair, int n_ao,", 0)
    E("const double* __restrict__ zeta_ab, const double* __restrict__ P,")
    E("const double* __restrict__ XPA, const double* __restrict__ XPB,")
    E("const double* __restrict__ s_ab, const int* __restrict__ rec0,")
    E("const int* __restrict__ rec1, const int* __restrict__ off_a,")
    E("const int* __restrict__ off_b, double* __restrict__ Q) {")
    E("const int pair = blockIdx.x * blockDim.x + threadIdx.x;")
    E("if (pair >= n_pair) return;")
    rows = _rows(la, lb)
    for i in range(len(rows)):
        E(f"double acc{i} = 0.0;")
    E("for (int i = rec0[pair]; i < rec1[pair]; ++i) {")
    E("const double zi = zeta_ab[i], si = s_ab[i];", 2)
    _vec("Pi", "P", "i", 2, E)
    _vec("Ai", "XPA", "i", 2, E)
    _vec("Bi", "XPB", "i", 2, E)
    emit_eaxis(E, la, lb, "zi", "Ai", "Bi", "Ei", 2)
    E("for (int k = rec0[pair]; k < rec1[pair]; ++k) {", 2)
    E("const double zk = zeta_ab[k], sk = s_ab[k];", 3)
    _vec("Pk", "P", "k", 3, E)
    _vec("Ak", "XPA", "k", 3, E)
    _vec("Bk", "XPB", "k", 3, E)
    emit_eaxis(E, la, lb, "zk", "Ak", "Bk", "Ek", 3)
    E("const double PQx = Pix - Pkx, PQy = Piy - Pky, PQz = Piz - Pkz;", 3)
    E("const double rho = zi * zk / (zi + zk);", 3)
    E("const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);", 3)
    R = emit_tower(E, L, 3)
    for ri, (a, b) in enumerate(rows):
        box = [a[d] + b[d] for d in range(3)]
        ps = [t for t in hermite_order(sum(box)) if all(t[d] <= box[d] for d in range(3))]
        E("{", 3)
        E("double v = 0.0;", 4)
        for p in ps:
            terms = []
            for q in ps:
                r = tuple(p[d] + q[d] for d in range(3))
                sg = "-" if sum(q) % 2 else "+"
                terms.append(f"{sg} {R[r]} * ({e_entry('Ek', a, b, q)})")
            w = " ".join(terms).lstrip("+ ")
            E(f"v += ({e_entry('Ei', a, b, p)}) * ({w});", 4)
        E(f"acc{ri} += si * sk * v;", 4)
        E("}", 3)
    E("}", 2)
    E("}")
    E("const int oa = off_a[pair], ob = off_b[pair];")
    nb = n_cart(lb)
    for ri, (a, b) in enumerate(rows):
        i, j = ri // nb, ri % nb
        E(f"{{ const double q = sqrt(fmax(acc{ri}, 0.0));")
        if la == lb:
            E(f"  if (oa != ob || {i} >= {j}) {{ Q[(size_t)(oa + {i}) * n_ao + ob + {j}] = q; "
              f"Q[(size_t)(ob + {j}) * n_ao + oa + {i}] = q; }} }}")
        else:
            E(f"  Q[(size_t)(oa + {i}) * n_ao + ob + {j}] = q; Q[(size_t)(ob + {j}) * n_ao + oa + {i}] = q; }}")
    E("}", 0)
    E("", 0)
    E(f'extern "C" int fj_schwarz_{n}(hipStream_t s, int threads, int n_pair, int n_rec, int n_ao,', 0)
    E("const double* zeta_ab, const double* P, const double* XPA, const double* XPB,")
    E("const double* s_ab, const int* rec0, const int* rec1, const int* off_a,")
    E("const int* off_b, double* schwarz_bound) {")
    E("if (threads <= 0) threads = 64;")
    E("if (threads % 64 || threads > 256) return FJ_ERR_THREADS;")
    E("if (n_pair < 0 || n_rec < 0 || n_ao <= 0) return FJ_ERR_SHAPE;")
    E("if (n_pair == 0) return 0;")
    E(f"hipLaunchKernelGGL(fj_schwarz_{n}_k, dim3((n_pair + threads - 1) / threads), dim3(threads), 0, s,")
    E("n_pair, n_ao, zeta_ab, P, XPA, XPB, s_ab, rec0, rec1, off_a, off_b, schwarz_bound);", 2)
    E("return (int)hipGetLastError();")
    E("}", 0)
    return E.text()


# ---------------------------------------------------------------------------
# DQ (upstream eri/j_coulomb/kernel_dq.py:1-110): one thread per ket record,
# D[k,q] = w_k sum_r Dcd[pair_k, r] (s_k E[r,q]),  Dmax_k = max_q |D[k,q]|


def emit_dq(lc: int, ld: int) -> str:
    n = f"{lc}_{ld}"
    Lq = lc + ld
    qs = hermite_order(Lq)
    rows = _rows(lc, ld)
    E = _Emit()
    E(f"// generated by felizjueves_amd/gen_mmd.py -- DQ class ({lc},{ld})", 0)
    E('#include "../fj_device.h"', 0)
    E("", 0)
    E(f"__global__ void __launch_bounds__(256) fj_dq_{n}_k(int n_rec, const double* __restrict__ eta,", 0)
    E("const double* __restrict__ XPA, const double* __restrict__ XPB,")
    E("const double* __restrict__ s, const double* __restrict__ w,")
    E("const int* __restrict__ pair, const double* __restrict__ Dcd,")
    E("double* __restrict__ D, double* __restrict__ Dmax) {")
    E("const int k = blockIdx.x * blockDim.x + threadIdx.x;")
    E("if (k >= n_rec) return;")
    E("const double z = eta[k], sk = s[k], wk = w[k];")
    _vec("A", "XPA", "k", 1, E)
    _vec("B", "XPB", "k", 1, E)
    emit_eaxis(E, lc, ld, "z", "A", "B", "E", 1)
    E(f"const double* __restrict__ dc = Dcd + (size_t)pair[k] * {len(rows)};")
    for ri in range(len(rows)):
        E(f"const double x{ri} = dc[{ri}];")
    E("double dm = 0.0;")
    for qi, q in enumerate(qs):
        terms = []
        for ri, (a, b) in enumerate(rows):
            e = e_entry("E", a, b, q)
            if e is not None:
                terms.append(f"x{ri} * (sk * ({e}))")
        E(f"{{ const double v = wk * ({' + '.join(terms) if terms else '0.0'}); "
          f"D[(size_t)k * {len(qs)} + {qi}] = v; dm = fmax(dm, fabs(v)); }}")
    E("Dmax[k] = dm;")
    E("}", 0)
    E("", 0)
    E(f'extern "C" int fj_dq_{n}(hipStream_t s_, int threads, int n_rec, int n_pair,', 0)
    E("const double* eta, const double* XPA, const double* XPB, const double* s,")
    E("const double* w, const int* pair, const double* Dcd, double* D, double* Dmax) {")
    E("if (threads <= 0) threads = 64;")
    E("if (threads % 64 || threads > 256) return FJ_ERR_THREADS;")
    E("if (n_rec < 0 || n_pair < 0) return FJ_ERR_SHAPE;")
    E("if (n_rec == 0) return 0;")
    E(f"hipLaunchKernelGGL(fj_dq_{n}_k, dim3((n_rec + threads - 1) / threads), dim3(threads), 0, s_,")
    E("n_rec, eta, XPA, XPB, s, w, pair, Dcd, D, Dmax);", 2)
    E("return (int)hipGetLastError();")
    E("}", 0)
    return E.text()


# ---------------------------------------------------------------------------
# [p|q] (upstream mmd/kernel.py:1-60,397-487): one wavefront per bra record,
# kets strided over its lanes, sorted descending by Iq Dmax so a lane stops
# at its first ket under tau; J_p = sum_k sum_q (-1)^|q| [p+q]^(0) D_q.


def emit_pq(lp: int, lq: int, f32: bool = False) -> str:
    """fj_pq_<n> (FP64) or fj_pq32_<n> (the J mixed lever: the Hermite
    recurrence in FP32 inside the FP64 boundary -- D_q, the J_p accumulator,
    the wavefront reduction and the write stay FP64)."""
    L = lp + lq
    n = f"L{L}_p{lp}_q{lq}"
    kn = ("pq32_" if f32 else "pq_") + n
    ps, qs = hermite_order(lp), hermite_order(lq)
    E = _Emit()
    E(f"// generated by felizjueves_amd/gen_mmd.py -- [p|q] class (lp={lp}, lq={lq})", 0)
    E('#include "../fj_device.h"', 0)
    E("", 0)
    E(f"__global__ void __launch_bounds__(256) fj_{kn}_k(int n_bra, int n_ket, double tau,", 0)
    E("const double* __restrict__ P, const double* __restrict__ zeta,")
    E("const double* __restrict__ Ip, const double* __restrict__ Q,")
    E("const double* __restrict__ eta, const double* __restrict__ Iq,")
    E("const double* __restrict__ Dmax, const double* __restrict__ D,")
    E("double* __restrict__ J) {")
    E("const int bra = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);")
    E("const int lane = threadIdx.x & 63;")
    E("if (bra >= n_bra) return;   // uniform per wavefront")
    E("const double z = zeta[bra], ip = Ip[bra];")
    _vec("P", "P", "bra", 1, E)
    for i in range(len(ps)):
        E(f"double j{i} = 0.0;")
    E("for (int k = lane; k < n_ket; k += 64) {")
    E("if (Dmax[k] * (ip * Iq[k]) < tau) break;   // kets sorted: the rest are smaller", 2)
    E("const double e = eta[k];", 2)
    E("const double PQx = Px - Q[3 * k], PQy = Py - Q[3 * k + 1], PQz = Pz - Q[3 * k + 2];", 2)
    E("const double rho = z * e / (z + e);", 2)
    E("const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);", 2)
    R = emit_tower(E, L, 2, f32=f32)
    E(f"const double* __restrict__ dq = D + (size_t)k * {len(qs)};", 2)
    for qi in range(len(qs)):
        E(f"const double d{qi} = dq[{qi}];", 2)
    for pi_, p in enumerate(ps):
        terms = []
        for qi, q in enumerate(qs):
            r = tuple(p[d] + q[d] for d in range(3))
            terms.append(("- " if sum(q) % 2 else "+ ") + f"{R[r]} * d{qi}")
        E(f"j{pi_} += {' '.join(terms).lstrip('+ ')};", 2)
    E("}")
    for i in range(len(ps)):
        E(f"{{ const double v = fj_wave_sum(j{i}); if (lane == 0) J[(size_t)bra * {len(ps)} + {i}] = v; }}")
    E("}", 0)
    E("", 0)
    E(f'extern "C" int fj_{kn}(hipStream_t s, int threads, int n_bra, int n_ket, double tau,', 0)
    E("const double* P, const double* zeta, const double* Ip, const double* Q,")
    E("const double* eta, const double* Iq, const double* Dmax, const double* D, double* J) {")
    E("if (threads <= 0) threads = 256;")
    E("if (threads % 64 || threads > 256) return FJ_ERR_THREADS;")
    E("if (n_bra < 0 || n_ket < 0) return FJ_ERR_SHAPE;")
    E("if (n_bra == 0) return 0;")
    E("const int waves = threads / 64;")
    E(f"hipLaunchKernelGGL(fj_{kn}_k, dim3((n_bra + waves - 1) / waves), dim3(threads), 0, s,")
    E("n_bra, n_ket, tau, P, zeta, Ip, Q, eta, Iq, Dmax, D, J);", 2)
    E("return (int)hipGetLastError();")
    E("}", 0)
    return E.text()


# ---------------------------------------------------------------------------
# JAB (upstream eri/j_coulomb/kernel_jab.py:1-80): one thread per bra record,
# v_r = sum_p (s E[r,p]) Jp[row, p], scattered with the mirror / diagonal rule


def emit_jab(la: int, lb: int) -> str:
    n = f"{la}{lb}"
    ps = hermite_order(la + lb)
    rows = _rows(la, lb)
    nb = n_cart(lb)
    E = _Emit()
    E(f"// generated by felizjueves_amd/gen_mmd.py -- JAB class ({la}{lb})", 0)
    E('#include "../fj_device.h"', 0)
    E("", 0)
    E(f"__global__ void __launch_bounds__(256) fj_jab_{n}_k(int n_rec, int n_ao,", 0)
    E("const double* __restrict__ zeta_ab, const double* __restrict__ XPA,")
    E("const double* __restrict__ XPB, const double* __restrict__ s,")
    E("const int* __restrict__ off_a, const int* __restrict__ off_b,")
    E("const int* __restrict__ row, const double* __restrict__ Jp, double* __restrict__ J) {")
    E("const int k = blockIdx.x * blockDim.x + threadIdx.x;")
    E("if (k >= n_rec) return;")
    E("const double z = zeta_ab[k], sk = s[k];")
    _vec("A", "XPA", "k", 1, E)
    _vec("B", "XPB", "k", 1, E)
    emit_eaxis(E, la, lb, "z", "A", "B", "E", 1)
    E(f"const double* __restrict__ jp = Jp + (size_t)row[k] * {len(ps)};")
    for pi_ in range(len(ps)):
        E(f"const double g{pi_} = jp[{pi_}];")
    E("const int oa = off_a[k], ob = off_b[k];")
    for ri, (a, b) in enumerate(rows):
        terms = []
        for pi_, p in enumerate(ps):
            e = e_entry("E", a, b, p)
            if e is not None:
                terms.append(f"(sk * ({e})) * g{pi_}")
        i, j = ri // nb, ri % nb
        E(f"{{ const double v = {' + '.join(terms)};")
        if la == lb:
            E(f"  if (oa != ob) {{ unsafeAtomicAdd(J + (size_t)(oa + {i}) * n_ao + ob + {j}, v); "
              f"unsafeAtomicAdd(J + (size_t)(ob + {j}) * n_ao + oa + {i}, v); }}")
            if i > j:
                E(f"  else {{ unsafeAtomicAdd(J + (size_t)(oa + {i}) * n_ao + oa + {j}, v); "
                  f"unsafeAtomicAdd(J + (size_t)(oa + {j}) * n_ao + oa + {i}, v); }}")
            elif i == j:
                E(f"  else unsafeAtomicAdd(J + (size_t)(oa + {i}) * n_ao + oa + {j}, v);")
            E("}")
        else:
            E(f"  unsafeAtomicAdd(J + (size_t)(oa + {i}) * n_ao + ob + {j}, v); "
              f"unsafeAtomicAdd(J + (size_t)(ob + {j}) * n_ao + oa + {i}, v); }}")
    E("}", 0)
    E("", 0)
    E(f'extern "C" int fj_jab_{n}(hipStream_t s_, int threads, int n_rec, int n_ao, int n_rows,', 0)
    E("const double* zeta_ab, const double* XPA, const double* XPB, const double* s,")
    E("const int* off_a, const int* off_b, const int* row, const double* Jp, double* J) {")
    E("if (threads <= 0) threads = 64;")
    E("if (threads % 64 || threads > 256) return FJ_ERR_THREADS;")
    E("if (n_rec < 0 || n_rows < 0 || n_ao <= 0) return FJ_ERR_SHAPE;")
    E("if (n_rec == 0) return 0;")
    E(f"hipLaunchKernelGGL(fj_jab_{n}_k, dim3((n_rec + threads - 1) / threads), dim3(threads), 0, s_,")
    E("n_rec, n_ao, zeta_ab, XPA, XPB, s, off_a, off_b, row, Jp, J);", 2)
    E("return (int)hipGetLastError();")
    E("}", 0)
    return E.text()


# ---------------------------------------------------------------------------
# K (upstream eri/k_exchange/kernel_k.py:1-120): one wavefront per bra record
# (a, c), kets (b, d) over its lanes, sorted descending by Iq; per quartet
#   I[p] = sum_q (-1)^|q| [p+q] (s_bd E^bd[bd, q])      per (b, d) row
#   K[a, b] += Dcd[dc + dd, (c, d)] sum_p (s_ac E^ac[ac, p]) I[p]
# the per-bra K block wave-reduced and stored by lane 0.


def emit_k(la: int, lc: int, lb: int, ld: int, f32: bool = False) -> str:
    """The K kernel of class (la lc | lb ld). f32: the K mixed lever's
    compute arm (upstream eri/k_exchange/levers/mixed.py:1-60, `kmixed`): the
    per-(bra, ket) region -- PQ, theta, T, the [0]^(m) and every [t]^(m) --
    in float, F_m in double on the widened float T; the E blocks, D_cd, the
    stage-1 I[p] and stage-2 folds and the K accumulation stay double (a
    tower value enters them widened). Entry fj_k32_<cls>."""
    n = f"{la}{lc}{lb}{ld}"
    kn = "k32" if f32 else "k"
    lp, lq = la + lc, lb + ld
    L = lp + lq
    ps, qs = hermite_order(lp), hermite_order(lq)
    na, nc, nb, nd = n_cart(la), n_cart(lc), n_cart(lb), n_cart(ld)
    E = _Emit()
    E(f"// generated by felizjueves_amd/gen_mmd.py -- K class ({la}{lc}|{lb}{ld})"
      + (", the FP32 tower (K mixed lever)" if f32 else ""), 0)
    E('#include "../fj_device.h"', 0)
    E("", 0)
    E(f"__global__ void __launch_bounds__(256) fj_{kn}_{n}_k(int n_bra, int n_ket, double tau, double maxD,", 0)
    E("const double* __restrict__ P, const double* __restrict__ zeta_ac,")
    E("const double* __restrict__ XPA, const double* __restrict__ XPC,")
    E("const double* __restrict__ s_ac, const double* __restrict__ Ip,")
    E("const int* __restrict__ dc, const double* __restrict__ Q,")
    E("const double* __restrict__ eta, const double* __restrict__ XQB,")
    E("const double* __restrict__ XQD, const double* __restrict__ s_bd,")
    E("const double* __restrict__ Iq, const int* __restrict__ dd,")
    E("const double* __restrict__ Dcd, double* __restrict__ K) {")
    E("const int bra = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);")
    E("const int lane = threadIdx.x & 63;")
    E("if (bra >= n_bra) return;   // uniform per wavefront")
    E("const double z = zeta_ac[bra], sa = s_ac[bra], ip = Ip[bra];")
    _vec("P", "P", "bra", 1, E)
    _vec("A", "XPA", "bra", 1, E)
    _vec("C", "XPC", "bra", 1, E)
    emit_eaxis(E, la, lc, "z", "A", "C", "Ea", 1)
    for a in range(na):
        for b in range(nb):
            E(f"double k{a}_{b} = 0.0;")
    E("const int dcb = dc[bra];")
    E("for (int k = lane; k < n_ket; k += 64) {")
    E("if (ip * Iq[k] * maxD < tau) break;   // kets sorted by Iq: the rest are smaller", 2)
    E("const double e = eta[k], sb = s_bd[k];", 2)
    _vec("B", "XQB", "k", 2, E)
    _vec("Dd", "XQD", "k", 2, E)
    emit_eaxis(E, lb, ld, "e", "B", "Dd", "Eb", 2)
    E("const double PQx = Px - Q[3 * k], PQy = Py - Q[3 * k + 1], PQz = Pz - Q[3 * k + 2];", 2)
    E("const double rho = z * e / (z + e);", 2)
    E("const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);", 2)
    R = emit_tower(E, L, 2, f32=f32)
    E(f"const double* __restrict__ dm = Dcd + (size_t)(dcb + dd[k]) * {nc * nd};", 2)
    for b_i, b in enumerate(cartesians(lb)):
        for d_i, d in enumerate(cartesians(ld)):
            E("{", 2)
            # stage 1: I[p] for this (b, d) row
            for pi_, p in enumerate(ps):
                terms = []
                for q in qs:
                    e = e_entry("Eb", b, d, q)
                    if e is None:
                        continue
                    r = tuple(p[x] + q[x] for x in range(3))
                    terms.append(("- " if sum(q) % 2 else "+ ") + f"{R[r]} * (sb * ({e}))")
                E(f"const double I{pi_} = {' '.join(terms).lstrip('+ ') if terms else '0.0'};", 3)
            # stages 2 and 3
            for a_i, a in enumerate(cartesians(la)):
                for c_i, c in enumerate(cartesians(lc)):
                    terms = []
                    for pi_, p in enumerate(ps):
                        e = e_entry("Ea", a, c, p)
                        if e is not None:
                            terms.append(f"(sa * ({e})) * I{pi_}")
                    E(f"k{a_i}_{b_i} += dm[{c_i * nd + d_i}] * ({' + '.join(terms)});", 3)
            E("}", 2)
    E("}")
    for a in range(na):
        for b in range(nb):
            E(f"{{ const double v = fj_wave_sum(k{a}_{b}); if (lane == 0) K[(size_t)bra * {na * nb} + {a * nb + b}] = v; }}")
    E("}", 0)
    E("", 0)
    E(f'extern "C" int fj_{kn}_{n}(hipStream_t s, int threads, int n_bra, int n_ket, int n_dc,', 0)
    E("double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,")
    E("const double* XPC, const double* s_ac, const double* Ip, const int* dc,")
    E("const double* Q, const double* eta, const double* XQB, const double* XQD,")
    E("const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K) {")
    E("if (threads <= 0) threads = 256;")
    E("if (threads % 64 || threads > 256) return FJ_ERR_THREADS;")
    E("if (n_bra < 0 || n_ket < 0 || n_dc < 0) return FJ_ERR_SHAPE;")
    E("if (n_bra == 0) return 0;")
    E("const int waves = threads / 64;")
    E(f"hipLaunchKernelGGL(fj_{kn}_{n}_k, dim3((n_bra + waves - 1) / waves), dim3(threads), 0, s,")
    E("n_bra, n_ket, tau, maxD, P, zeta_ac, XPA, XPC, s_ac, Ip, dc, Q, eta, XQB, XQD, s_bd, Iq, dd, Dcd, K);", 2)
    E("return (int)hipGetLastError();")
    E("}", 0)
    return E.text()


# ---------------------------------------------------------------------------

def pair_classes(lmax: int):
    return [(a, b) for a in range(lmax + 1) for b in range(a, lmax + 1)]


def k_classes(lmax: int):
    r = range(lmax + 1)
    return [(a, c, b, d) for a in r for c in r for b in r for d in r]


# Straight-line Schwarz units for pairs with la + lb <= this: the (ab|ab)
# tower of a (df) or (ff) pair keeps too much live (8.5 / 19 KB of scratch
# per lane, above the library's 6 KB cap); those two classes run on the
# loop-form kernel (csrc/fj_kloop.h).
SCHWARZ_LSUM_MAX = 4


def units(lmax: int, k_lmax: int | None = None) -> dict[str, str]:
    """Every unit through angular momentum lmax. The K classes are
    straight-line kernels only through k_lmax (their Hermite towers grow as
    L^6 in source: (ff|ff) alone would be 18 MB); the classes past it get
    entries onto the loop-form kernel (csrc/fj_kloop.h)."""
    k_lmax = lmax if k_lmax is None else k_lmax
    out = {}
    for la, lb in pair_classes(lmax):
        if la + lb <= SCHWARZ_LSUM_MAX:
            out[f"schwarz_{la}{lb}.hip"] = emit_schwarz(la, lb)
        out[f"dq_{la}_{lb}.hip"] = emit_dq(la, lb)
        out[f"jab_{la}{lb}.hip"] = emit_jab(la, lb)
    big = [c for c in pair_classes(lmax) if sum(c) > SCHWARZ_LSUM_MAX]
    if big:
        out["schwarz_loop_entries.hip"] = emit_schwarz_loop_entries(big)
    for lp in range(2 * lmax + 1):
        for lq in range(2 * lmax + 1):
            out[f"pq_L{lp + lq}_p{lp}_q{lq}.hip"] = emit_pq(lp, lq)
            out[f"pq32_L{lp + lq}_p{lp}_q{lq}.hip"] = emit_pq(lp, lq, f32=True)
    for c in k_classes(k_lmax):
        out[f"k_{''.join(map(str, c))}.hip"] = emit_k(*c)
        out[f"k32_{''.join(map(str, c))}.hip"] = emit_k(*c, f32=True)
    loop = [c for c in k_classes(lmax) if max(c) > k_lmax]
    if loop:
        out["k_loop_entries.hip"] = emit_k_loop_entries(loop)
    return out


K_ARGS = ("int n_bra, int n_ket, int n_dc, double tau, double maxD, const double* P,\n"
          "    const double* zeta_ac, const double* XPA, const double* XPC, const double* s_ac,\n"
          "    const double* Ip, const int* dc, const double* Q, const double* eta,\n"
          "    const double* XQB, const double* XQD, const double* s_bd, const double* Iq,\n"
          "    const int* dd, const double* Dcd, double* K")
K_NAMES = ("n_bra, n_ket, n_dc, tau, maxD, P, zeta_ac, XPA, XPC, s_ac, Ip, dc, Q, eta, XQB, XQD, "
           "s_bd, Iq, dd, Dcd, K")


S_ARGS = ("int n_pair, int n_rec, int n_ao, const double* zeta_ab, const double* P,\n"
          "    const double* XPA, const double* XPB, const double* s_ab, const int* rec0,\n"
          "    const int* rec1, const int* off_a, const int* off_b, double* schwarz_bound")
S_NAMES = "n_pair, n_rec, n_ao, zeta_ab, P, XPA, XPB, s_ab, rec0, rec1, off_a, off_b, schwarz_bound"


def emit_schwarz_loop_entries(classes) -> str:
    """The fj_schwarz_<la><lb> entries of the pair classes past the
    straight-line Schwarz kernels: the same parameter list, launched
    through the loop-form kernel (csrc/fj_kloop.h)."""
    E = _Emit()
    E("// generated by felizjueves_amd/gen_mmd.py -- Schwarz classes in loop form", 0)
    E('#include "../fj_device.h"', 0)
    E("", 0)
    E("int fj_schwarz_loop_launch(hipStream_t s, int la, int lb, int threads,", 0)
    E(S_ARGS.replace("\n    ", " ") + ");")
    for la, lb in classes:
        E("", 0)
        E(f'extern "C" int fj_schwarz_{la}{lb}(hipStream_t s, int threads, '
          + S_ARGS.replace("\n    ", " ") + ") {", 0)
        E(f"return fj_schwarz_loop_launch(s, {la}, {lb}, threads, {S_NAMES});")
        E("}", 0)
    return E.text()


def emit_k_loop_entries(classes) -> str:
    """The fj_k_<cls> entries of the classes past the straight-line K
    kernels: the same parameter list, launched through the loop-form kernel
    (csrc/fj_kloop.h) with the class as arguments."""
    E = _Emit()
    E("// generated by felizjueves_amd/gen_mmd.py -- K classes in loop form", 0)
    E('#include "../fj_device.h"', 0)
    E("", 0)
    E("int fj_kloop_launch(hipStream_t s, int la, int lc, int lb, int ld, int threads,", 0)
    E(K_ARGS.replace("\n    ", " ") + ");")
    for c in classes:
        n = "".join(map(str, c))
        E("", 0)
        E(f'extern "C" int fj_k_{n}(hipStream_t s, int threads, ' + K_ARGS.replace("\n    ", " ") + ") {", 0)
        E(f"return fj_kloop_launch(s, {c[0]}, {c[1]}, {c[2]}, {c[3]}, threads, {K_NAMES});")
        E("}", 0)
    return E.text()


def generate(lmax: int, out: Path = GEN, k_lmax: int | None = None) -> list[Path]:
    out.mkdir(parents=True, exist_ok=True)
    files = []
    for name, text in units(lmax, k_lmax).items():
        p = out / name
        if not p.exists() or p.read_text() != text:
            p.write_text(text)
        files.append(p)
    return files


def header_decls(lmax: int, k_lmax: int | None = None) -> str:
    k_lmax = lmax if k_lmax is None else k_lmax
    out = []
    for la, lb in pair_classes(lmax):
        out.append(
            f"int fj_schwarz_{la}{lb}(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_ao,\n"
            f"    const double* zeta_ab, const double* P, const double* XPA, const double* XPB,\n"
            f"    const double* s_ab, const int* rec0, const int* rec1, const int* off_a,\n"
            f"    const int* off_b, double* schwarz_bound);")
        out.append(
            f"int fj_dq_{la}_{lb}(fj_stream_t stream, int threads, int n_rec, int n_pair,\n"
            f"    const double* eta, const double* XPA, const double* XPB, const double* s,\n"
            f"    const double* w, const int* pair, const double* Dcd, double* D, double* Dmax);")
        out.append(
            f"int fj_jab_{la}{lb}(fj_stream_t stream, int threads, int n_rec, int n_ao, int n_rows,\n"
            f"    const double* zeta_ab, const double* XPA, const double* XPB, const double* s,\n"
            f"    const int* off_a, const int* off_b, const int* row, const double* Jp, double* J);")
    for lp in range(2 * lmax + 1):
        for lq in range(2 * lmax + 1):
            for kn in ("pq", "pq32"):
                out.append(
                    f"int fj_{kn}_L{lp + lq}_p{lp}_q{lq}(fj_stream_t stream, int threads, int n_bra,"
                    f" int n_ket, double tau,\n"
                    f"    const double* P, const double* zeta, const double* Ip, const double* Q,\n"
                    f"    const double* eta, const double* Iq, const double* Dmax, const double* D,"
                    f" double* J);")
    for c in k_classes(lmax):
        out.append(
            f"int fj_k_{''.join(map(str, c))}(fj_stream_t stream, int threads, int n_bra, int n_ket,"
            f" int n_dc,\n"
            f"    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,\n"
            f"    const double* XPC, const double* s_ac, const double* Ip, const int* dc,\n"
            f"    const double* Q, const double* eta, const double* XQB, const double* XQD,\n"
            f"    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);")
    for c in k_classes(k_lmax):
        out.append(
            f"int fj_k32_{''.join(map(str, c))}(fj_stream_t stream, int threads, int n_bra, int n_ket,"
            f" int n_dc,\n"
            f"    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,\n"
            f"    const double* XPC, const double* s_ac, const double* Ip, const int* dc,\n"
            f"    const double* Q, const double* eta, const double* XQB, const double* XQD,\n"
            f"    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);")
    return "\n".join(out) + "\n"


HEADER = Path(__file__).resolve().parent.parent / "include" / "fj_mi355x.h"
BEGIN, END = ("/* BEGIN MMD CLASSES (written by felizjueves_amd/gen_mmd.py) */",
              "/* END MMD CLASSES */")


def header_text(lmax: int, current: str, k_lmax: int | None = None) -> str:
    i, j = current.index(BEGIN), current.index(END)
    return current[:i + len(BEGIN)] + "\n" + header_decls(lmax, k_lmax) + current[j:]


def write_header(lmax: int, k_lmax: int | None = None) -> None:
    cur = HEADER.read_text()
    new = header_text(lmax, cur, k_lmax)
    if new != cur:
        HEADER.write_text(new)
