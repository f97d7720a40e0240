/*
 * fj_oracle.c -- CPU FP64 oracle for the two-electron Fock-build path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * the library built from it (oracle/liboracle.so).  The product path
 * (felizjueves_amd/) never links, loads or calls it.
 *
 * Each function restates the upstream algorithm it cites (paths relative to
 * the upstream FelizJueves checkout).  Nothing is copied; the arithmetic is
 * written from the stated recurrences.
 *
 *   oracle_boys          F_m(T), m = 0..M            reference/boys.py:184-218
 *   oracle_eri_mmd       (ab|cd) by McMurchie-Davidson  mmd/hermite.py:45-117,
 *                                                     reference/pq.py:54-133
 *   oracle_eri_hgp       (ab|cd) by HGP VRR + HRR       hgp/vrr.py:152-178,
 *                                                     hgp/hrr.py:101-120,
 *                                                     hgp/recurrence.py:308-335
 *   oracle_f_launch      one fused-F launch (same args as the device entry)
 *   oracle_fx_launch_mmd the same launch with MMD integrals (no HRR)
 *                                                     eri/fock/kernel_f.py:1-165,
 *                                                     levers/seams.py:400-770
 *   oracle_fock_ordered  J, K by EVERY ordered quartet, weight 1
 *                                                     eri/fock/fock.py:126-176
 *   oracle_schwarz       Q_mn = sqrt(max((mn|mn), 0))   eri/schwarz.py:64-106
 *   oracle_grad_ordered  dE_2e/dR by EVERY ordered quartet, the first
 *                        centre differentiated  gradient/kernel_gf.py:12-60
 *
 * Parity pins (tests/test_oracle_*.py): libcint int2e blocks, vj, vk and
 * int2e_diag of the tracked fixtures; Boys closed forms and a 50-digit
 * Kummer series; MMD vs HGP per quartet; the unique-quartet digestion vs the
 * ordered sweep.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TWO_OVER_SQRT_PI 1.1283791670955126
#define PI 3.14159265358979323846

/* ---------------------------------------------------------------------- */
/* Cartesian index helpers: components of |e| = n enumerated i desc, j desc */

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }
static int ncart_upto(int n) { return n * (n + 1) * (n + 2) / 6; } /* |e| < n */
static int cidx(int i, int j, int k) {                             /* within level */
    int l = i + j + k;
    return (l - i) * (l - i + 1) / 2 + (l - i - j);
}
static int fidx(int i, int j, int k) { return ncart_upto(i + j + k) + cidx(i, j, k); }

typedef struct { int c[3]; } cart3;
static cart3 cart_of(int l, int idx) {
    cart3 r;
    int n = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) {
            if (n == idx) { r.c[0] = i; r.c[1] = j; r.c[2] = l - i - j; return r; }
            ++n;
        }
    r.c[0] = r.c[1] = r.c[2] = -1;
    return r;
}

/* ---------------------------------------------------------------------- */
/* Boys function.  Series (Kummer) in long double below T_SERIES, then the
 * downward recursion F_{m-1} = (2T F_m + e^-T)/(2m-1); above it the
 * erf-closed F_0 and the upward recursion F_{m+1} = ((2m+1) F_m - e^-T)/(2T),
 * stable for T > m. */
#define T_SERIES 45.0L

void oracle_boys(int M, double Td, double *F) {
    long double T = Td;
    long double e = expl(-T);
    if (T < T_SERIES) {
        /* F_M(T) = e^-T sum_k (2T)^k / ((2M+1)(2M+3)...(2M+2k+1)) */
        long double term = 1.0L / (2 * M + 1), sum = term;
        for (int k = 1; k < 100000; ++k) {
            term *= 2.0L * T / (2 * M + 2 * k + 1);
            sum += term;
            if (term < sum * 1e-21L) break;
        }
        long double f = e * sum;
        F[M] = (double)f;
        for (int m = M; m > 0; --m) {
            f = (2.0L * T * f + e) / (2 * m - 1);
            F[m - 1] = (double)f;
        }
    } else {
        long double f = 0.5L * sqrtl(PI / T) * erfl(sqrtl(T));
        F[0] = (double)f;
        for (int m = 0; m < M; ++m) {
            f = ((2 * m + 1) * f - e) / (2.0L * T);
            F[m + 1] = (double)f;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* McMurchie-Davidson.  E^{ij}_t per axis (mmd/hermite.py:45-79):
 *   E^{00}_0 = exp(-mu X_AB^2) = exp(zeta X_PA X_PB)
 *   E^{i+1,j}_t = E^{ij}_{t-1}/(2 zeta) + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}
 *   E^{i,j+1}_t = E^{ij}_{t-1}/(2 zeta) + X_PB E^{ij}_t + (t+1) E^{ij}_{t+1}
 * stored E[i][j][t], i <= la, j <= lb, t <= i+j.                             */
#define LM 8  /* per-shell l up to 7 (f is 3) */
static void e_axis(int la, int lb, double zeta, double pa, double pb,
                   double E[LM][LM][2 * LM]) {
    memset(E, 0, sizeof(double) * LM * LM * 2 * LM);
    double h = 0.5 / zeta;
    E[0][0][0] = exp(zeta * pa * pb);
    for (int i = 0; i <= la; ++i) {
        if (i > 0)
            for (int t = 0; t <= i; ++t) {
                double v = pa * E[i - 1][0][t];
                if (t > 0) v += h * E[i - 1][0][t - 1];
                if (t + 1 <= i - 1) v += (t + 1) * E[i - 1][0][t + 1];
                E[i][0][t] = v;
            }
        for (int j = 1; j <= lb; ++j)
            for (int t = 0; t <= i + j; ++t) {
                double v = (t <= i + j - 1) ? pb * E[i][j - 1][t] : 0.0;
                if (t > 0) v += h * E[i][j - 1][t - 1];
                if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1];
                E[i][j][t] = v;
            }
    }
}

/* R_{tuv}^{(0)} for t+u+v <= L (reference/pq.py:94-133): the raising
 * recursion [r]^(m) = PQ_d [r-1_d]^(m+1) + (r_d - 1)[r-2_d]^(m+1) from
 * [0]^(m) = (-2 theta^2)^m F_m(T) (the (2/sqrt(pi)) theta is applied by the
 * caller).  R[n][t][u][v] laid out flat, n = 0..L.                           */
static void r_tower(int L, double X, double Y, double Z, double rho,
                    double *R /* [(L+1)^4] */) {
    int n1 = L + 1;
#define RI(n, t, u, v) ((((n) * n1 + (t)) * n1 + (u)) * n1 + (v))
    double F[64];
    oracle_boys(L, rho * (X * X + Y * Y + Z * Z), F);
    double f = 1.0;
    for (int n = 0; n <= L; ++n) {
        R[RI(n, 0, 0, 0)] = f * F[n];
        f *= -2.0 * rho;
    }
    for (int n = L - 1; n >= 0; --n) {
        int top = L - n;
        for (int t = 0; t <= top; ++t)
            for (int u = 0; u <= top - t; ++u)
                for (int v = 0; v <= top - t - u; ++v) {
                    if (t + u + v == 0) continue;
                    double val;
                    if (t > 0) {
                        val = X * R[RI(n + 1, t - 1, u, v)];
                        if (t > 1) val += (t - 1) * R[RI(n + 1, t - 2, u, v)];
                    } else if (u > 0) {
                        val = Y * R[RI(n + 1, t, u - 1, v)];
                        if (u > 1) val += (u - 1) * R[RI(n + 1, t, u - 2, v)];
                    } else {
                        val = Z * R[RI(n + 1, t, u, v - 1)];
                        if (v > 1) val += (v - 1) * R[RI(n + 1, t, u, v - 2)];
                    }
                    R[RI(n, t, u, v)] = val;
                }
    }
#undef RI
}

/* One contracted shell (unnormalised Cartesian primitives, coefficients
 * carry every normalisation: ao/shell.py:1-40). */
typedef struct {
    int l;
    const double *A;
    int np;
    const double *ex, *co;
} shell_t;

static shell_t mkshell(int s, const int *shell_l, const double *centre,
                       const int *nprim, const int *pstart, const double *exps,
                       const double *coefs) {
    shell_t r = {shell_l[s], centre + 3 * s, nprim[s], exps + pstart[s], coefs + pstart[s]};
    return r;
}

/* (ab|cd) by MMD, out[i][j][k][l] row-major over cartesians(la..ld). */
static void eri_mmd_sh(const shell_t *a, const shell_t *b, const shell_t *c,
                       const shell_t *d, double *out) {
    int na = ncart(a->l), nb = ncart(b->l), nc = ncart(c->l), nd = ncart(d->l);
    int L = a->l + b->l + c->l + d->l, n1 = L + 1;
    memset(out, 0, sizeof(double) * na * nb * nc * nd);
    double *R = malloc(sizeof(double) * n1 * n1 * n1 * n1);
    static __thread double Eab[3][LM][LM][2 * LM], Ecd[3][LM][LM][2 * LM];
    for (int i = 0; i < a->np; ++i)
        for (int j = 0; j < b->np; ++j) {
            double al = a->ex[i], be = b->ex[j], p = al + be;
            double P[3];
            for (int x = 0; x < 3; ++x) {
                P[x] = (al * a->A[x] + be * b->A[x]) / p;
                e_axis(a->l, b->l, p, P[x] - a->A[x], P[x] - b->A[x], Eab[x]);
            }
            for (int k = 0; k < c->np; ++k)
                for (int m = 0; m < d->np; ++m) {
                    double ga = c->ex[k], de = d->ex[m], q = ga + de;
                    double Q[3];
                    for (int x = 0; x < 3; ++x) {
                        Q[x] = (ga * c->A[x] + de * d->A[x]) / q;
                        e_axis(c->l, d->l, q, Q[x] - c->A[x], Q[x] - d->A[x], Ecd[x]);
                    }
                    double rho = p * q / (p + q);
                    r_tower(L, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], rho, R);
                    double pref = 2.0 * pow(PI, 2.5) / (p * q * sqrt(p + q)) *
                                  a->co[i] * b->co[j] * c->co[k] * d->co[m];
                    for (int ia = 0; ia < na; ++ia) {
                        cart3 ca = cart_of(a->l, ia);
                        for (int ib = 0; ib < nb; ++ib) {
                            cart3 cb = cart_of(b->l, ib);
                            for (int ic = 0; ic < nc; ++ic) {
                                cart3 cc = cart_of(c->l, ic);
                                for (int id = 0; id < nd; ++id) {
                                    cart3 cd = cart_of(d->l, id);
                                    double s = 0.0;
                                    int tx = ca.c[0] + cb.c[0], ty = ca.c[1] + cb.c[1],
                                        tz = ca.c[2] + cb.c[2];
                                    int ux = cc.c[0] + cd.c[0], uy = cc.c[1] + cd.c[1],
                                        uz = cc.c[2] + cd.c[2];
                                    for (int t = 0; t <= tx; ++t)
                                        for (int u = 0; u <= ty; ++u)
                                            for (int v = 0; v <= tz; ++v) {
                                                double e1 = Eab[0][ca.c[0]][cb.c[0]][t] *
                                                            Eab[1][ca.c[1]][cb.c[1]][u] *
                                                            Eab[2][ca.c[2]][cb.c[2]][v];
                                                if (e1 == 0.0) continue;
                                                for (int tt = 0; tt <= ux; ++tt)
                                                    for (int uu = 0; uu <= uy; ++uu)
                                                        for (int vv = 0; vv <= uz; ++vv) {
                                                            double e2 =
                                                                Ecd[0][cc.c[0]][cd.c[0]][tt] *
                                                                Ecd[1][cc.c[1]][cd.c[1]][uu] *
                                                                Ecd[2][cc.c[2]][cd.c[2]][vv];
                                                            double sg = ((tt + uu + vv) & 1) ? -1.0 : 1.0;
                                                            s += e1 * sg * e2 *
                                                                 R[((0 * n1 + t + tt) * n1 + u + uu) * n1 + v + vv];
                                                        }
                                            }
                                    out[((ia * nb + ib) * nc + ic) * nd + id] += pref * s;
                                }
                            }
                        }
                    }
                }
        }
    free(R);
}

void oracle_eri_mmd(int sa, int sb, int sc, int sd, const int *shell_l,
                    const double *centre, const int *nprim, const int *pstart,
                    const double *exps, const double *coefs, double *out) {
    shell_t a = mkshell(sa, shell_l, centre, nprim, pstart, exps, coefs);
    shell_t b = mkshell(sb, shell_l, centre, nprim, pstart, exps, coefs);
    shell_t c = mkshell(sc, shell_l, centre, nprim, pstart, exps, coefs);
    shell_t d = mkshell(sd, shell_l, centre, nprim, pstart, exps, coefs);
    eri_mmd_sh(&a, &b, &c, &d, out);
}

/* ---------------------------------------------------------------------- */
/* HGP.  One primitive quartet's VRR (hgp/vrr.py:1-60,152-178), forward
 * build over |e| <= E, |f| <= Fk, m <= L - |e| - |f|, reduced along the
 * first nonzero axis of f (eq 26) else of e (eq 25), seed
 * [00|00]^(m) = U_P U_Q (2/sqrt(pi)) theta F_m(T) (hgp/recurrence.py:62-94),
 * W - Q = +p/(p+q) PQ (the derived sign, hgp/recurrence.py:53-61).
 * Adds [e0|f0]^(0) into the contracted boundary C[e][f] for
 * la <= |e| <= E, lc <= |f| <= Fk (full-table indices).                     */
typedef struct {
    double P[3], PA[3], p, U;
} prim_t;

static void hgp_vrr_add(int la, int E, int lc, int Fk, const prim_t *bp,
                        const prim_t *kp, double *V /* scratch */, double *C) {
    int L = E + Fk, nm = L + 1;
    int ne = ncart_upto(E + 1), nf = ncart_upto(Fk + 1);
#define VI(ei, fi, m) (((ei) * nf + (fi)) * nm + (m))
    double p = bp->p, q = kp->p, ppq = p + q, rho = p * q / ppq, theta = sqrt(rho);
    double PQ[3], WP[3], WQ[3];
    for (int x = 0; x < 3; ++x) {
        PQ[x] = bp->P[x] - kp->P[x];
        WP[x] = -(q / ppq) * PQ[x];
        WQ[x] = (p / ppq) * PQ[x];
    }
    double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
    double S = bp->U * kp->U * TWO_OVER_SQRT_PI * theta;
    double Fm[64];
    oracle_boys(L, T, Fm);
    double i2p = 0.5 / p, i2q = 0.5 / q, i2pq = 0.5 / ppq;
    double i2p_r = i2p * (rho / p), i2q_r = i2q * (rho / q);
    for (int m = 0; m <= L; ++m) V[VI(0, 0, m)] = S * Fm[m];
    /* bra phase, f = 0 */
    for (int n = 1; n <= E; ++n)
        for (int c = 0; c < ncart(n); ++c) {
            cart3 e = cart_of(n, c);
            int d = e.c[0] ? 0 : (e.c[1] ? 1 : 2);
            int g[3] = {e.c[0], e.c[1], e.c[2]};
            g[d] -= 1;
            int gi = fidx(g[0], g[1], g[2]), ei = fidx(e.c[0], e.c[1], e.c[2]);
            int hi = -1;
            if (g[d] > 0) { int h[3] = {g[0], g[1], g[2]}; h[d] -= 1; hi = fidx(h[0], h[1], h[2]); }
            for (int m = 0; m <= L - n; ++m) {
                double v = bp->PA[d] * V[VI(gi, 0, m)] + WP[d] * V[VI(gi, 0, m + 1)];
                if (hi >= 0) v += g[d] * (i2p * V[VI(hi, 0, m)] - i2p_r * V[VI(hi, 0, m + 1)]);
                V[VI(ei, 0, m)] = v;
            }
        }
    /* ket phase by |f| ascending */
    for (int nfv = 1; nfv <= Fk; ++nfv)
        for (int cf = 0; cf < ncart(nfv); ++cf) {
            cart3 f = cart_of(nfv, cf);
            int d = f.c[0] ? 0 : (f.c[1] ? 1 : 2);
            int g[3] = {f.c[0], f.c[1], f.c[2]};
            g[d] -= 1;
            int fi = fidx(f.c[0], f.c[1], f.c[2]), gi = fidx(g[0], g[1], g[2]);
            int hi = -1;
            if (g[d] > 0) { int h[3] = {g[0], g[1], g[2]}; h[d] -= 1; hi = fidx(h[0], h[1], h[2]); }
            for (int n = 0; n <= E; ++n)
                for (int c = 0; c < ncart(n); ++c) {
                    cart3 e = cart_of(n, c);
                    int ei = fidx(e.c[0], e.c[1], e.c[2]);
                    int emi = -1;
                    if (e.c[d] > 0) { int t[3] = {e.c[0], e.c[1], e.c[2]}; t[d] -= 1; emi = fidx(t[0], t[1], t[2]); }
                    for (int m = 0; m <= L - n - nfv; ++m) {
                        double v = kp->PA[d] * V[VI(ei, gi, m)] + WQ[d] * V[VI(ei, gi, m + 1)];
                        if (hi >= 0) v += g[d] * (i2q * V[VI(ei, hi, m)] - i2q_r * V[VI(ei, hi, m + 1)]);
                        if (emi >= 0) v += e.c[d] * i2pq * V[VI(emi, gi, m + 1)];
                        V[VI(ei, fi, m)] = v;
                    }
                }
        }
    for (int ei = ncart_upto(la); ei < ne; ++ei)
        for (int fi = ncart_upto(lc); fi < nf; ++fi) C[ei * nf + fi] += V[VI(ei, fi, 0)];
#undef VI
}

/* HRR on contracted values (hgp/hrr.py:1-60,101-120):
 *   (a, b+1_k | = (a+1_k, b | + AB_k (a, b |     ket twin with CD.
 * C is [ne][nf] over full tables; out is (ab|cd) row-major.  The bra is
 * transferred first, then the ket.                                        */
static void hgp_hrr(int la, int lb, int lc, int ld, const double AB[3],
                    const double CD[3], const double *C, double *out) {
    int E = la + lb, Fk = lc + ld;
    int nf = ncart_upto(Fk + 1);
    int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    /* bra: X[j] over (a full idx, b idx in level j, f) with |a| in [la, E-j] */
    int nef = ncart_upto(E + 1);
    double *cur = malloc(sizeof(double) * nef * 1 * nf);
    memcpy(cur, C, sizeof(double) * nef * nf); /* j = 0: b = 0 */
    for (int j = 1; j <= lb; ++j) {
        int nbj = ncart(j), nbp = ncart(j - 1);
        double *nxt = calloc((size_t)nef * nbj * nf, sizeof(double));
        for (int ib = 0; ib < nbj; ++ib) {
            cart3 b = cart_of(j, ib);
            int k = b.c[0] ? 0 : (b.c[1] ? 1 : 2);
            int bm[3] = {b.c[0], b.c[1], b.c[2]};
            bm[k] -= 1;
            int ibm = cidx(bm[0], bm[1], bm[2]);
            for (int n = la; n <= E - j; ++n)
                for (int c = 0; c < ncart(n); ++c) {
                    cart3 a = cart_of(n, c);
                    int ai = fidx(a.c[0], a.c[1], a.c[2]);
                    int ap[3] = {a.c[0], a.c[1], a.c[2]};
                    ap[k] += 1;
                    int api = fidx(ap[0], ap[1], ap[2]);
                    for (int f = 0; f < nf; ++f)
                        nxt[(ai * nbj + ib) * nf + f] =
                            cur[(api * nbp + ibm) * nf + f] + AB[k] * cur[(ai * nbp + ibm) * nf + f];
                }
        }
        free(cur);
        cur = nxt;
    }
    /* cur: [nef][nb][nf] with |a| = la valid.  Now the ket for each (a, b). */
    double *ab = malloc(sizeof(double) * na * nb * nf);
    for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib)
            for (int f = 0; f < nf; ++f)
                ab[(ia * nb + ib) * nf + f] = cur[((ncart_upto(la) + ia) * nb + ib) * nf + f];
    free(cur);
    int nabv = na * nb;
    double *kc = malloc(sizeof(double) * nabv * nf);
    memcpy(kc, ab, sizeof(double) * nabv * nf);
    for (int j = 1; j <= ld; ++j) {
        int ndj = ncart(j), ndp = ncart(j - 1);
        double *nxt = calloc((size_t)nabv * nf * ndj, sizeof(double));
        for (int id = 0; id < ndj; ++id) {
            cart3 dd = cart_of(j, id);
            int k = dd.c[0] ? 0 : (dd.c[1] ? 1 : 2);
            int dm[3] = {dd.c[0], dd.c[1], dd.c[2]};
            dm[k] -= 1;
            int idm = cidx(dm[0], dm[1], dm[2]);
            for (int n = lc; n <= Fk - j; ++n)
                for (int c = 0; c < ncart(n); ++c) {
                    cart3 cc = cart_of(n, c);
                    int ci = fidx(cc.c[0], cc.c[1], cc.c[2]);
                    int cp[3] = {cc.c[0], cc.c[1], cc.c[2]};
                    cp[k] += 1;
                    int cpi = fidx(cp[0], cp[1], cp[2]);
                    for (int x = 0; x < nabv; ++x)
                        nxt[(x * nf + ci) * ndj + id] =
                            kc[(x * nf + cpi) * ndp + idm] + CD[k] * kc[(x * nf + ci) * ndp + idm];
                }
        }
        free(kc);
        kc = nxt;
    }
    for (int x = 0; x < nabv; ++x)
        for (int ic = 0; ic < nc; ++ic)
            for (int id = 0; id < nd; ++id)
                out[(x * nc + ic) * nd + id] = kc[(x * nf + ncart_upto(lc) + ic) * nd + id];
    free(kc);
    free(ab);
}

/* A primitive-pair record (eri/fock/kernel_f.py:529-560; walk_f.py:188-210):
 * P, X = P - A, zeta, U = c_a c_b (pi/zeta)^{3/2} exp(-a b/zeta |AB|^2).    */
static prim_t make_prim(const shell_t *a, int i, const shell_t *b, int j) {
    prim_t r;
    double al = a->ex[i], be = b->ex[j], p = al + be, ab2 = 0.0;
    for (int x = 0; x < 3; ++x) {
        double d = a->A[x] - b->A[x];
        ab2 += d * d;
        r.P[x] = (al * a->A[x] + be * b->A[x]) / p;
        r.PA[x] = r.P[x] - a->A[x];
    }
    r.p = p;
    r.U = a->co[i] * b->co[j] * pow(PI / p, 1.5) * exp(-al * be / p * ab2);
    return r;
}

void oracle_eri_hgp(int sa, int sb, int sc, int sd, const int *shell_l,
                    const double *centre, const int *nprim, const int *pstart,
                    const double *exps, const double *coefs, double *out) {
    shell_t a = mkshell(sa, shell_l, centre, nprim, pstart, exps, coefs);
    shell_t b = mkshell(sb, shell_l, centre, nprim, pstart, exps, coefs);
    shell_t c = mkshell(sc, shell_l, centre, nprim, pstart, exps, coefs);
    shell_t d = mkshell(sd, shell_l, centre, nprim, pstart, exps, coefs);
    int E = a.l + b.l, Fk = c.l + d.l, L = E + Fk;
    int ne = ncart_upto(E + 1), nf = ncart_upto(Fk + 1);
    double *V = malloc(sizeof(double) * ne * nf * (L + 1));
    double *C = calloc((size_t)ne * nf, sizeof(double));
    for (int i = 0; i < a.np; ++i)
        for (int j = 0; j < b.np; ++j) {
            prim_t bp = make_prim(&a, i, &b, j);
            for (int k = 0; k < c.np; ++k)
                for (int m = 0; m < d.np; ++m) {
                    prim_t kp = make_prim(&c, k, &d, m);
                    hgp_vrr_add(a.l, E, c.l, Fk, &bp, &kp, V, C);
                }
        }
    double AB[3], CD[3];
    for (int x = 0; x < 3; ++x) { AB[x] = a.A[x] - b.A[x]; CD[x] = c.A[x] - d.A[x]; }
    hgp_hrr(a.l, b.l, c.l, d.l, AB, CD, C, out);
    free(V);
    free(C);
}

/* ---------------------------------------------------------------------- */
/* One fused-F launch, on the CPU, with the device entry's arguments
 * (levers/seams.py:228-235,739-770).  Semantics (kernel_f.py:29-165,
 * seams.py:424-737): per bra pair and per ket pair k with k >= diag*bra
 * (the canonical half), dmax = max(2 Dsh_ab, 2 Dsh_cd, 1/2 Dsh_ac,
 * 1/2 Dsh_ad, 1/2 Dsh_bc, 1/2 Dsh_bd); skip when G_ab G_cd dmax < tau;
 * with `pre` set, also skip when G_ab G_cd < tau (the density-free
 * pre-launch screen of launch_f.py:448-466, applied per quartet, as the
 * builder's plan applies it);
 * otherwise the K_CD x K_AB nest of HGP VRR, the HRR on the contracted
 * boundary, and the six-block digestion with weights (2, -1/2) times
 * deg = (1/2)^{[a=b]+[c=d]+[P=Q]} (kernel_f.py:439-487), ADDED into F
 * (the caller closes F <- F + F^T).  AB / CD may be NULL when lb / ld is 0.
 * Rows of bra_prim / ket_prim: (Px,Py,Pz, Xx,Xy,Xz, zeta, U).               */
void oracle_fx_launch(int la, int lb, int lc, int ld, int n_bra, int n_ket,
                      int k_ab, int k_cd, const int *bra_off, const int *ket_off,
                      int n_sh, int n_ao, int diag, double tau, int pre,
                      const double *bra_prim, const double *AB, const double *G_ab,
                      const int *bra_sh, const int *bra_ao, const double *ket_prim,
                      const double *CD, const double *G_cd, const int *ket_sh,
                      const int *ket_ao, const double *Dsh, const double *D, double *F) {
    int E = la + lb, Fk = lc + ld, L = E + Fk;
    int ne = ncart_upto(E + 1), nf = ncart_upto(Fk + 1);
    int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    double *Fp = calloc((size_t)nthr * n_ao * n_ao, sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *Ft = Fp + (size_t)tid * n_ao * n_ao;
        double *V = malloc(sizeof(double) * ne * nf * (L + 1));
        double *C = malloc(sizeof(double) * ne * nf);
        double *x = malloc(sizeof(double) * na * nb * nc * nd);
#pragma omp for schedule(dynamic, 1)
        for (int bra = 0; bra < n_bra; ++bra) {
            int sa = bra_sh[2 * bra], sb = bra_sh[2 * bra + 1];
            int oa = bra_ao[2 * bra], ob = bra_ao[2 * bra + 1];
            double ab0[3] = {0, 0, 0};
            const double *ABp = AB ? AB + 3 * bra : ab0;
            for (int k = diag * bra; k < n_ket; ++k) {
                int sc = ket_sh[2 * k], sd = ket_sh[2 * k + 1];
                int oc = ket_ao[2 * k], od = ket_ao[2 * k + 1];
                double dm = 2.0 * Dsh[sa * n_sh + sb], t;
                if ((t = 2.0 * Dsh[sc * n_sh + sd]) > dm) dm = t;
                if ((t = 0.5 * Dsh[sa * n_sh + sc]) > dm) dm = t;
                if ((t = 0.5 * Dsh[sa * n_sh + sd]) > dm) dm = t;
                if ((t = 0.5 * Dsh[sb * n_sh + sc]) > dm) dm = t;
                if ((t = 0.5 * Dsh[sb * n_sh + sd]) > dm) dm = t;
                if (pre && G_ab[bra] * G_cd[k] < tau) continue;
                if (G_ab[bra] * (G_cd[k] * dm) < tau) continue;
                memset(C, 0, sizeof(double) * ne * nf);
                int kc0 = ket_off ? ket_off[k] : k * k_cd;
                int nkc = ket_off ? ket_off[k + 1] - kc0 : k_cd;
                int kb0 = bra_off ? bra_off[bra] : bra * k_ab;
                int nkb = bra_off ? bra_off[bra + 1] - kb0 : k_ab;
                for (int kc = 0; kc < nkc; ++kc) {
                    const double *r = ket_prim + ((size_t)kc0 + kc) * 8;
                    prim_t kp = {{r[0], r[1], r[2]}, {r[3], r[4], r[5]}, r[6], r[7]};
                    for (int kb = 0; kb < nkb; ++kb) {
                        const double *s = bra_prim + ((size_t)kb0 + kb) * 8;
                        prim_t bp = {{s[0], s[1], s[2]}, {s[3], s[4], s[5]}, s[6], s[7]};
                        hgp_vrr_add(la, E, lc, Fk, &bp, &kp, V, C);
                    }
                }
                double cd0[3] = {0, 0, 0};
                hgp_hrr(la, lb, lc, ld, ABp, CD ? CD + 3 * k : cd0, C, x);
                double deg = 1.0;
                if (sa == sb) deg *= 0.5;
                if (sc == sd) deg *= 0.5;
                if (diag && k == bra) deg *= 0.5;
                double sj = 2.0 * deg, sk = -0.5 * deg;
                for (int i = 0; i < na; ++i)
                    for (int j = 0; j < nb; ++j)
                        for (int c = 0; c < nc; ++c)
                            for (int d = 0; d < nd; ++d) {
                                double v = x[((i * nb + j) * nc + c) * nd + d];
                                int A = oa + i, B = ob + j, Cc = oc + c, Dd = od + d;
                                Ft[A * n_ao + B] += sj * D[Cc * n_ao + Dd] * v;
                                Ft[Cc * n_ao + Dd] += sj * D[A * n_ao + B] * v;
                                Ft[A * n_ao + Cc] += sk * D[B * n_ao + Dd] * v;
                                Ft[A * n_ao + Dd] += sk * D[B * n_ao + Cc] * v;
                                Ft[B * n_ao + Cc] += sk * D[A * n_ao + Dd] * v;
                                Ft[B * n_ao + Dd] += sk * D[A * n_ao + Cc] * v;
                            }
            }
        }
        free(V);
        free(C);
        free(x);
    }
    for (int t = 0; t < nthr; ++t)
        for (size_t i = 0; i < (size_t)n_ao * n_ao; ++i) F[i] += Fp[(size_t)t * n_ao * n_ao + i];
    free(Fp);
}

/* The same launch (screen, canonical half, weights, six-block digestion)
 * with every quartet's integrals from the MMD path over the shells' full
 * primitive lists instead of the HGP VRR/HRR over the launch's records:
 * MMD has no HRR, so it keeps full FP64 accuracy where the HGP transfer
 * cancels (a very tight primitive paired with a distant diffuse shell). */
void oracle_fx_launch_mmd(int la, int lb, int lc, int ld, int n_bra, int n_ket, int n_sh,
                          int n_ao, int diag, double tau, int pre, const double *G_ab,
                          const int *bra_sh,
                          const int *bra_ao, const double *G_cd, const int *ket_sh,
                          const int *ket_ao, const double *Dsh, const double *D,
                          const int *shell_l, const double *centre, const int *nprim,
                          const int *pstart, const double *exps, const double *coefs,
                          double *F) {
    int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    double *Fp = calloc((size_t)nthr * n_ao * n_ao, sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *Ft = Fp + (size_t)tid * n_ao * n_ao;
        double *x = malloc(sizeof(double) * na * nb * nc * nd);
#pragma omp for schedule(dynamic, 1)
        for (int bra = 0; bra < n_bra; ++bra) {
            int sa = bra_sh[2 * bra], sb = bra_sh[2 * bra + 1];
            int oa = bra_ao[2 * bra], ob = bra_ao[2 * bra + 1];
            shell_t A = mkshell(sa, shell_l, centre, nprim, pstart, exps, coefs);
            shell_t B = mkshell(sb, shell_l, centre, nprim, pstart, exps, coefs);
            for (int k = diag * bra; k < n_ket; ++k) {
                int sc = ket_sh[2 * k], sd = ket_sh[2 * k + 1];
                int oc = ket_ao[2 * k], od = ket_ao[2 * k + 1];
                double dm = 2.0 * Dsh[sa * n_sh + sb], t;
                if ((t = 2.0 * Dsh[sc * n_sh + sd]) > dm) dm = t;
                if ((t = 0.5 * Dsh[sa * n_sh + sc]) > dm) dm = t;
                if ((t = 0.5 * Dsh[sa * n_sh + sd]) > dm) dm = t;
                if ((t = 0.5 * Dsh[sb * n_sh + sc]) > dm) dm = t;
                if ((t = 0.5 * Dsh[sb * n_sh + sd]) > dm) dm = t;
                if (pre && G_ab[bra] * G_cd[k] < tau) continue;
                if (G_ab[bra] * (G_cd[k] * dm) < tau) continue;
                shell_t Cs = mkshell(sc, shell_l, centre, nprim, pstart, exps, coefs);
                shell_t Ds = mkshell(sd, shell_l, centre, nprim, pstart, exps, coefs);
                eri_mmd_sh(&A, &B, &Cs, &Ds, x);
                double deg = 1.0;
                if (sa == sb) deg *= 0.5;
                if (sc == sd) deg *= 0.5;
                if (diag && k == bra) deg *= 0.5;
                double sj = 2.0 * deg, sk = -0.5 * deg;
                for (int i = 0; i < na; ++i)
                    for (int j = 0; j < nb; ++j)
                        for (int c = 0; c < nc; ++c)
                            for (int d = 0; d < nd; ++d) {
                                double v = x[((i * nb + j) * nc + c) * nd + d];
                                int Ai = oa + i, Bj = ob + j, Cc = oc + c, Dd = od + d;
                                Ft[Ai * n_ao + Bj] += sj * D[Cc * n_ao + Dd] * v;
                                Ft[Cc * n_ao + Dd] += sj * D[Ai * n_ao + Bj] * v;
                                Ft[Ai * n_ao + Cc] += sk * D[Bj * n_ao + Dd] * v;
                                Ft[Ai * n_ao + Dd] += sk * D[Bj * n_ao + Cc] * v;
                                Ft[Bj * n_ao + Cc] += sk * D[Ai * n_ao + Dd] * v;
                                Ft[Bj * n_ao + Dd] += sk * D[Ai * n_ao + Cc] * v;
                            }
            }
        }
        free(x);
    }
    for (int t = 0; t < nthr; ++t)
        for (size_t i = 0; i < (size_t)n_ao * n_ao; ++i) F[i] += Fp[(size_t)t * n_ao * n_ao + i];
    free(Fp);
}

/* The upstream per-class argument list: uniform degrees k_ab, k_cd. */
void oracle_f_launch(int la, int lb, int lc, int ld, int n_bra, int n_ket,
                     int k_ab, int k_cd, int n_sh, int n_ao, int diag, double tau,
                     const double *bra_prim, const double *AB, const double *G_ab,
                     const int *bra_sh, const int *bra_ao, const double *ket_prim,
                     const double *CD, const double *G_cd, const int *ket_sh,
                     const int *ket_ao, const double *Dsh, const double *D, double *F) {
    oracle_fx_launch(la, lb, lc, ld, n_bra, n_ket, k_ab, k_cd, NULL, NULL, n_sh, n_ao, diag,
                     tau, 0, bra_prim, AB, G_ab, bra_sh, bra_ao, ket_prim, CD, G_cd, ket_sh,
                     ket_ao, Dsh, D, F);
}

/* ---------------------------------------------------------------------- */
/* J and K by every ORDERED shell quartet at weight 1 (eri/fock/fock.py:
 * 126-163): J[a_i][b_j] += D[c_k][d_l] x,  K[a_i][c_k] += D[b_j][d_l] x.   */
void oracle_fock_ordered(int n_sh, const int *shell_l, const double *centre,
                         const int *nprim, const int *pstart, const double *exps,
                         const double *coefs, const int *shell_ao, int n_ao,
                         const double *D, double *J, double *K) {
    memset(J, 0, sizeof(double) * n_ao * n_ao);
    memset(K, 0, sizeof(double) * n_ao * n_ao);
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    double *Jp = calloc((size_t)nthr * n_ao * n_ao, sizeof(double));
    double *Kp = calloc((size_t)nthr * n_ao * n_ao, sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *Jt = Jp + (size_t)tid * n_ao * n_ao, *Kt = Kp + (size_t)tid * n_ao * n_ao;
        double *x = malloc(sizeof(double) * 10 * 10 * 10 * 10);
#pragma omp for schedule(dynamic, 1) collapse(2)
        for (int a = 0; a < n_sh; ++a)
            for (int b = 0; b < n_sh; ++b)
                for (int c = 0; c < n_sh; ++c)
                    for (int d = 0; d < n_sh; ++d) {
                        shell_t A = mkshell(a, shell_l, centre, nprim, pstart, exps, coefs);
                        shell_t B = mkshell(b, shell_l, centre, nprim, pstart, exps, coefs);
                        shell_t Cc = mkshell(c, shell_l, centre, nprim, pstart, exps, coefs);
                        shell_t Dd = mkshell(d, shell_l, centre, nprim, pstart, exps, coefs);
                        eri_mmd_sh(&A, &B, &Cc, &Dd, x);
                        int na = ncart(A.l), nb = ncart(B.l), nc = ncart(Cc.l), nd = ncart(Dd.l);
                        int oa = shell_ao[a], ob = shell_ao[b], oc = shell_ao[c], od = shell_ao[d];
                        for (int i = 0; i < na; ++i)
                            for (int j = 0; j < nb; ++j)
                                for (int k = 0; k < nc; ++k)
                                    for (int l = 0; l < nd; ++l) {
                                        double v = x[((i * nb + j) * nc + k) * nd + l];
                                        Jt[(oa + i) * n_ao + ob + j] += D[(oc + k) * n_ao + od + l] * v;
                                        Kt[(oa + i) * n_ao + oc + k] += D[(ob + j) * n_ao + od + l] * v;
                                    }
                    }
        free(x);
    }
    for (int t = 0; t < nthr; ++t)
        for (size_t i = 0; i < (size_t)n_ao * n_ao; ++i) {
            J[i] += Jp[(size_t)t * n_ao * n_ao + i];
            K[i] += Kp[(size_t)t * n_ao * n_ao + i];
        }
    free(Jp);
    free(Kp);
}

/* Q_mn = sqrt(max((mn|mn), 0)) for every AO pair (eri/schwarz.py:64-106). */
void oracle_schwarz(int n_sh, const int *shell_l, const double *centre,
                    const int *nprim, const int *pstart, const double *exps,
                    const double *coefs, const int *shell_ao, int n_ao, double *Q) {
#pragma omp parallel
    {
        double *x = malloc(sizeof(double) * 10 * 10 * 10 * 10);
#pragma omp for schedule(dynamic, 1)
        for (int a = 0; a < n_sh; ++a)
            for (int b = 0; b < n_sh; ++b) {
                shell_t A = mkshell(a, shell_l, centre, nprim, pstart, exps, coefs);
                shell_t B = mkshell(b, shell_l, centre, nprim, pstart, exps, coefs);
                eri_mmd_sh(&A, &B, &A, &B, x);
                int na = ncart(A.l), nb = ncart(B.l);
                for (int i = 0; i < na; ++i)
                    for (int j = 0; j < nb; ++j) {
                        double v = x[((i * nb + j) * na + i) * nb + j];
                        Q[(shell_ao[a] + i) * n_ao + shell_ao[b] + j] = sqrt(v > 0.0 ? v : 0.0);
                    }
            }
        free(x);
    }
}

/* dE_2e/dR, E_2e = 1/2 tr(D J) - 1/4 tr(D K) = sum_ijkl Gam_ijkl (ij|kl)
 * (gradient/kernel_gf.py:12-24), by every ORDERED shell quartet with only
 * its first function's centre differentiated: the four positions of an
 * ordered sweep are equivalent under the ERI symmetry, so with Gam
 * symmetrised over it,
 *
 *   dE/dR_{A,x} = sum_{ijkl, i on A} g_ijkl d_i,x (ij|kl),
 *   g_ijkl = 2 D_ij D_kl - 1/2 (D_ik D_jl + D_il D_jk)
 *
 * and d_i,x (ij|kl) = (i+1_x j|kl)^{2 alpha} - i_x (i-1_x j|kl) (the
 * derivative of a Cartesian Gaussian, gradient/kernel_gf.py:44-50): the
 * raised block from the shell at l + 1 with coefficients c 2 alpha, the
 * lowered one from the shell at l - 1. G [n_atom][3] is overwritten. This is
 * another formulation than the unique-quartet digestion (no weights w, no
 * degeneracy, one centre), pinned to PySCF's gradient terms and to finite
 * differences of oracle_fock_ordered's energy in the tests. */
void oracle_grad_ordered(int n_sh, const int *shell_l, const double *centre,
                         const int *nprim, const int *pstart, const double *exps,
                         const double *coefs, const int *shell_ao, const int *shell_atom,
                         int n_ao, const double *D, int n_atom, double *G) {
    memset(G, 0, sizeof(double) * n_atom * 3);
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    double *Gp = calloc((size_t)nthr * n_atom * 3, sizeof(double));
    int np_max = 1;
    for (int s = 0; s < n_sh; ++s)
        if (nprim[s] > np_max) np_max = nprim[s];
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *Gt = Gp + (size_t)tid * n_atom * 3;
        double *up = malloc(sizeof(double) * 15 * 15 * 15 * 15);
        double *dn = malloc(sizeof(double) * 15 * 15 * 15 * 15);
        double *cup = malloc(sizeof(double) * np_max);
#pragma omp for schedule(dynamic, 1) collapse(2)
        for (int a = 0; a < n_sh; ++a)
            for (int b = 0; b < n_sh; ++b) {
                shell_t A = mkshell(a, shell_l, centre, nprim, pstart, exps, coefs);
                shell_t B = mkshell(b, shell_l, centre, nprim, pstart, exps, coefs);
                shell_t Au = A, Ad = A;
                for (int i = 0; i < A.np; ++i) cup[i] = 2.0 * A.ex[i] * A.co[i];
                Au.l = A.l + 1;
                Au.co = cup;
                Ad.l = A.l - 1;
                int la = A.l, na = ncart(la), nb = ncart(B.l);
                int oa = shell_ao[a], ob = shell_ao[b], at = shell_atom[a];
                for (int c = 0; c < n_sh; ++c)
                    for (int d = 0; d < n_sh; ++d) {
                        shell_t Cc = mkshell(c, shell_l, centre, nprim, pstart, exps, coefs);
                        shell_t Dd = mkshell(d, shell_l, centre, nprim, pstart, exps, coefs);
                        int nc = ncart(Cc.l), nd = ncart(Dd.l);
                        int oc = shell_ao[c], od = shell_ao[d];
                        eri_mmd_sh(&Au, &B, &Cc, &Dd, up);
                        if (la > 0) eri_mmd_sh(&Ad, &B, &Cc, &Dd, dn);
                        for (int i = 0; i < na; ++i) {
                            cart3 ci = cart_of(la, i);
                            for (int j = 0; j < nb; ++j)
                                for (int k = 0; k < nc; ++k)
                                    for (int l = 0; l < nd; ++l) {
                                        double g = 2.0 * D[(oa + i) * n_ao + ob + j] * D[(oc + k) * n_ao + od + l]
                                                 - 0.5 * (D[(oa + i) * n_ao + oc + k] * D[(ob + j) * n_ao + od + l]
                                                        + D[(oa + i) * n_ao + od + l] * D[(ob + j) * n_ao + oc + k]);
                                        for (int x = 0; x < 3; ++x) {
                                            int e[3] = {ci.c[0], ci.c[1], ci.c[2]};
                                            e[x] += 1;
                                            int iu = cidx(e[0], e[1], e[2]);
                                            double v = up[((iu * nb + j) * nc + k) * nd + l];
                                            if (ci.c[x] > 0) {
                                                e[x] -= 2;
                                                int id = cidx(e[0], e[1], e[2]);
                                                v -= (double)ci.c[x] * dn[((id * nb + j) * nc + k) * nd + l];
                                            }
                                            Gt[at * 3 + x] += g * v;
                                        }
                                    }
                        }
                    }
            }
        free(up);
        free(dn);
        free(cup);
    }
    for (int t = 0; t < nthr; ++t)
        for (int i = 0; i < n_atom * 3; ++i) G[i] += Gp[(size_t)t * n_atom * 3 + i];
    free(Gp);
}

int oracle_version(void) { return 1; }

/* The OpenMP threads the oracle's parallel loops use. */
int oracle_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
