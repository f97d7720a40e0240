// fj_kloop.h -- the MMD kernels in loop form, every class through f: K
// (below) and the Schwarz diagonal (fj_schwarz_loop_k, at the end).
//
// K:
// K_ab += sum_cd D_cd (ac|bd) per bra record (upstream
// eri/k_exchange/kernel_k.py:314, host k_quartet k.py:395), with the class
// (la lc | lb ld) as run-time arguments instead of straight-line code.
//
// The straight-line K kernels (gen_mmd.emit_k) grow as L^6 in source; this
// one is a single kernel whose loops read small index tables. One workgroup
// of two wavefronts (FJ_KL_BLOCK lanes) per bra record and slice of the
// kets (gridDim.y slices, so that a launch of few bras fills the chip); the
// kets of a slice are taken one at a time by the whole workgroup,
// cooperating through LDS:
//
//   per-axis E tables   E^{ij}_t of the bra (once) and of the ket, one lane
//                       per axis (upstream mmd/hermite.py:45-79)
//   Hermite tower       [t]^(m), m = L .. 0, one level per step, lanes over t
//                       (upstream reference/pq.py:54-133), two LDS rows
//   X[c][q]             s_bd (-1)^|q| sum_d D_cd E^{bd}_q   (per b)
//   Y[c][p]             sum_q [p + q] X[c][q]   (4 x 4 register tiles of (c, p))
//   kac[a][c][b]       += sum_p E^{ac}_p Y[c][p]
//
// and K[bra][a][b] = s_ac sum_c kac[a][c][b] at the end. Every loop bound is
// uniform across the wavefront, so each barrier is reached by all lanes.
//
// The includer provides fj_device.h (the device one, or the CPU emulation).
#pragma once

#define FJ_KL_LMAX 3                       // through f
#define FJ_KL_MMAX (4 * FJ_KL_LMAX)        // the tower's top order, L <= 12
#define FJ_KL_NH 455                       // Hermite functions with |t| <= 12
#define FJ_KL_NP 84                        // |p| <= 6 (one side)
#define FJ_KL_NC 10                        // Cartesians of an f shell
#define FJ_KL_AX 112                       // one axis's E table: 4 x 4 x 7
#define FJ_KL_BLOCK 128                    // lanes per workgroup of the K kernel
#define FJ_KL_WORKGROUPS 8192              // the K launch's bras x ket slices, at most

// Hermite functions with |t| < n.
__device__ __forceinline__ int fj_herm_base(int n) { return n * (n + 1) * (n + 2) / 6; }

// Index of t = (x, y, z) in the Hermite order: |t| ascending, then x
// descending, then y descending (the order of the Cartesians of one shell).
__device__ __forceinline__ int fj_herm_index(int x, int y, int z) {
    const int u = y + z;
    return fj_herm_base(x + u) + u * (u + 1) / 2 + z;
}

// The components of Hermite index j, packed x | y << 4 | z << 8.
__device__ __forceinline__ int fj_herm_decode(int j) {
    int n = 0;
    while (fj_herm_base(n + 1) <= j) ++n;
    const int pos = j - fj_herm_base(n);
    int u = 0;
    while ((u + 1) * (u + 2) / 2 <= pos) ++u;
    const int z = pos - u * (u + 1) / 2;
    return (n - u) | ((u - z) << 4) | (z << 8);
}

// One axis's E^{ij}_t, i <= l1, j <= l2, at e[(i (l2 + 1) + j) (l1 + l2 + 1) + t],
// zero where t > i + j:  E^{00}_0 = exp(zeta X_PA X_PB),
// E^{i+1,j}_t = E^{ij}_{t-1} / (2 zeta) + X_PA E^{ij}_t + (t + 1) E^{ij}_{t+1},
// the j raise with X_PB.
// The seed E^{00}_0 is e00 (fj_kl_eaxis: exp(zeta X_PA X_PB); the F records
// carry that factor in U, so their tables start from 1).
__device__ __forceinline__ void fj_kl_eaxis_from(double* e, int l1, int l2, double zeta, double pa,
                                                 double pb, double e00) {
    const int nj = l2 + 1, nt = l1 + l2 + 1;
    for (int k = 0; k < (l1 + 1) * nj * nt; ++k) e[k] = 0.0;
    const double h = 0.5 / zeta;
    e[0] = e00;
    for (int i = 0; i <= l1; ++i) {
        for (int j = (i == 0 ? 1 : 0); j <= l2; ++j) {
            // (i, 0) from (i - 1, 0); (i, j > 0) from (i, j - 1)
            const double* s = j == 0 ? e + (i - 1) * nj * nt : e + (i * nj + j - 1) * nt;
            const double x = j == 0 ? pa : pb;
            const int top = i + j;
            double* d = e + (i * nj + j) * nt;
            for (int t = 0; t <= top; ++t) {
                double v = 0.0;
                if (t >= 1) v += h * s[t - 1];
                if (t <= top - 1) v += x * s[t];
                if (t + 1 <= top - 1) v += (double)(t + 1) * s[t + 1];
                d[t] = v;
            }
        }
    }
}

__device__ __forceinline__ void fj_kl_eaxis(double* e, int l1, int l2, double zeta, double pa,
                                            double pb) {
    fj_kl_eaxis_from(e, l1, l2, zeta, pa, pb, exp(zeta * pa * pb));
}

// The tower's step table for every Hermite index j < nh: for t != 0 with
// first nonzero axis d, [t] = PQ_d [t - 1_d] + (t_d - 1) [t - 2_d], packed as
// index(t - 1_d) | index(t - 2_d) << 10 | d << 20 | (t_d - 1) << 22; with hc,
// also the packed components of j.
__device__ __forceinline__ void fj_kl_steps_n(int* tw, int* hc, int nh, int lane, int nthr) {
    for (int j = lane; j < nh; j += nthr) {
        const int c = fj_herm_decode(j);
        const int x = c & 15, y = (c >> 4) & 15, z = c >> 8;
        if (hc) hc[j] = c;
        int w = 0;
        if (j > 0) {
            const int d = x ? 0 : (y ? 1 : 2);
            const int td = d == 0 ? x : (d == 1 ? y : z);
            const int i1 = fj_herm_index(x - (d == 0), y - (d == 1), z - (d == 2));
            const int i2 = td >= 2 ? fj_herm_index(x - 2 * (d == 0), y - 2 * (d == 1), z - 2 * (d == 2))
                                   : 0;
            w = i1 | (i2 << 10) | (d << 20) | ((td - 1) << 22);
        }
        tw[j] = w;
    }
}

__device__ __forceinline__ void fj_kl_steps(int* tw, int* hc, int nh, int lane) {
    fj_kl_steps_n(tw, hc, nh, lane, 64);
}

// wf[m] = [0]^(m) = (2 / sqrt(pi)) sqrt(rho) (-2 rho)^m F_m(T), m <= M.
template <int M>
__device__ __forceinline__ void fj_kl_seed_m(double* wf, double rho, double T) {
    double Fm[M + 1];
    fj_boys<M>(T, Fm);
    const double m2r = -2.0 * rho;
    double wm = FJ_TWO_OVER_SQRT_PI * sqrt(rho);
#pragma unroll
    for (int m = 0; m <= M; ++m) {
        wf[m] = wm * Fm[m];
        wm *= m2r;
    }
}

// The same for m <= L (the tower's top order, 0 .. 12): the Boys function
// evaluated at that order, not at the largest one.
__device__ __forceinline__ void fj_kl_seed(double* wf, double rho, double T, int L) {
    switch (L) {
    case 0: fj_kl_seed_m<0>(wf, rho, T); break;
    case 1: fj_kl_seed_m<1>(wf, rho, T); break;
    case 2: fj_kl_seed_m<2>(wf, rho, T); break;
    case 3: fj_kl_seed_m<3>(wf, rho, T); break;
    case 4: fj_kl_seed_m<4>(wf, rho, T); break;
    case 5: fj_kl_seed_m<5>(wf, rho, T); break;
    case 6: fj_kl_seed_m<6>(wf, rho, T); break;
    case 7: fj_kl_seed_m<7>(wf, rho, T); break;
    case 8: fj_kl_seed_m<8>(wf, rho, T); break;
    case 9: fj_kl_seed_m<9>(wf, rho, T); break;
    case 10: fj_kl_seed_m<10>(wf, rho, T); break;
    case 11: fj_kl_seed_m<11>(wf, rho, T); break;
    default: fj_kl_seed_m<FJ_KL_MMAX>(wf, rho, T); break;
    }
}

// Levels m = L - 1 .. 0 of the tower, lanes over t (lane of nthr; the _n
// forms take the workgroup's width, the plain ones one wavefront): level m holds [t]^(m)
// for |t| <= L - m in row m & 1 (R0 or R1), so [t]^(0) ends in R0. The
// caller has written wf and row L & 1's [0] = wf[L] and passed a barrier;
// every lane of the workgroup calls this (a barrier after each level).
__device__ __forceinline__ void fj_kl_tower_n(double* R0, double* R1, const double* wf, const int* tw,
                                              int L, double PQx, double PQy, double PQz, int lane,
                                              int nthr) {
    for (int m = L - 1; m >= 0; --m) {
        double* __restrict__ cur = (m & 1) ? R1 : R0;
        const double* __restrict__ prv = (m & 1) ? R0 : R1;
        const int n_m = fj_herm_base(L - m + 1);
        for (int j = lane; j < n_m; j += nthr) {
            double v;
            if (j == 0) {
                v = wf[m];
            } else {
                const int w = tw[j];
                const int d = (w >> 20) & 3, c1 = w >> 22;
                const double pq = d == 0 ? PQx : (d == 1 ? PQy : PQz);
                v = pq * prv[w & 1023];
                if (c1 > 0) v += (double)c1 * prv[(w >> 10) & 1023];
            }
            cur[j] = v;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void fj_kl_tower(double* R0, double* R1, const double* wf, const int* tw,
                                            int L, double PQx, double PQy, double PQz, int lane) {
    fj_kl_tower_n(R0, R1, wf, tw, L, PQx, PQy, PQz, lane, 64);
}

// An includer that needs only the helpers above (the E tables, the tower)
// defines FJ_KLOOP_HELPERS_ONLY first; the kernels below stay in one unit.
#ifndef FJ_KLOOP_HELPERS_ONLY

// The K kernel's LDS, sized for its class and carved from one dynamic block:
// byte offsets of each array (doubles, then ints, then the u16 index table)
// and the total. The launcher and the kernel both use this layout.
struct FjKlLds {
    int R0, R1, ea, eb, X, Y, kac, dsh, wf, hc, tw, pqi, bytes, axa, axb;
};

__host__ __device__ inline FjKlLds fj_kloop_lds(int la, int lc, int lb, int ld) {
    auto herm = [](int n) { return n * (n + 1) * (n + 2) / 6; };
    const int na = (la + 1) * (la + 2) / 2, nc = (lc + 1) * (lc + 2) / 2;
    const int nb = (lb + 1) * (lb + 2) / 2, nd = (ld + 1) * (ld + 2) / 2;
    const int np = herm(la + lc + 1), nq = herm(lb + ld + 1), nh = herm(la + lc + lb + ld + 1);
    FjKlLds o;
    o.axa = (la + 1) * (lc + 1) * (la + lc + 1);
    o.axb = (lb + 1) * (ld + 1) * (lb + ld + 1);
    int at = 0;
    auto take = [&](int n, int size) { const int r = at; at += ((n * size + 7) / 8) * 8; return r; };
    o.R0 = take(nh, 8);
    o.R1 = take(nh, 8);
    o.ea = take(3 * o.axa, 8);
    o.eb = take(3 * o.axb, 8);
    o.X = take(nc * nq, 8);
    o.Y = take(nc * np, 8);
    o.kac = take(na * nc * nb, 8);
    o.dsh = take(nc * nd, 8);
    o.wf = take(FJ_KL_MMAX + 1, 8);
    o.hc = take(nh, 4);
    o.tw = take(nh, 4);
    o.pqi = take(np * nq, 2);
    o.bytes = at;
    return o;
}

__global__ void __launch_bounds__(256) fj_kloop_k(
    int la, int lc, int lb, int ld, int n_bra, int n_ket, int n_dc, double tau, double maxD,
    const double* __restrict__ P, const double* __restrict__ zeta_ac,
    const double* __restrict__ XPA, const double* __restrict__ XPC,
    const double* __restrict__ s_ac, const double* __restrict__ Ip,
    const int* __restrict__ dc, const double* __restrict__ Q,
    const double* __restrict__ eta, const double* __restrict__ XQB,
    const double* __restrict__ XQD, const double* __restrict__ s_bd,
    const double* __restrict__ Iq, const int* __restrict__ dd,
    const double* __restrict__ Dcd, double* __restrict__ K) {
    FJ_DYN_SHARED(fj_kloop_lds_block);

    const int bra = (int)blockIdx.x, lane = (int)threadIdx.x, nthr = (int)blockDim.x;
    if (bra >= n_bra) return;   // uniform per workgroup
    const int lp = la + lc, lq = lb + ld, L = lp + lq;
    const int na = (la + 1) * (la + 2) / 2, nc = (lc + 1) * (lc + 2) / 2;
    const int nb = (lb + 1) * (lb + 2) / 2, nd = (ld + 1) * (ld + 2) / 2;
    const int np = fj_herm_base(lp + 1), nq = fj_herm_base(lq + 1), nh = fj_herm_base(L + 1);
    const int oa = fj_herm_base(la), ob = fj_herm_base(lb), oc = fj_herm_base(lc),
              od = fj_herm_base(ld);
    const int ta = lp + 1, tb = lq + 1;   // t extents of the two E tables
    // the class's LDS arrays
    const FjKlLds lay = fj_kloop_lds(la, lc, lb, ld);
    char* const lds = reinterpret_cast<char*>(fj_kloop_lds_block);
    double* const Rr0 = reinterpret_cast<double*>(lds + lay.R0);
    double* const Rr1 = reinterpret_cast<double*>(lds + lay.R1);
    double* const ea = reinterpret_cast<double*>(lds + lay.ea);
    double* const eb = reinterpret_cast<double*>(lds + lay.eb);
    double* const X = reinterpret_cast<double*>(lds + lay.X);
    double* const Y = reinterpret_cast<double*>(lds + lay.Y);
    double* const kac = reinterpret_cast<double*>(lds + lay.kac);
    double* const dsh = reinterpret_cast<double*>(lds + lay.dsh);
    double* const wf = reinterpret_cast<double*>(lds + lay.wf);
    int* const hc = reinterpret_cast<int*>(lds + lay.hc);
    int* const tw = reinterpret_cast<int*>(lds + lay.tw);
    unsigned short* const pqi = reinterpret_cast<unsigned short*>(lds + lay.pqi);
    const int axa = lay.axa, axb = lay.axb;

    // index tables: components of every Hermite index, the tower's step of
    // each t (first nonzero axis d: [t] = PQ_d [t - 1_d] + (t_d - 1) [t - 2_d]),
    // and the index of p + q
    fj_kl_steps_n(tw, hc, nh, lane, nthr);
    for (int j = lane; j < np * nq; j += nthr) {
        const int p = j / nq, q = j - p * nq;
        const int cp = fj_herm_decode(p), cq = fj_herm_decode(q);
        pqi[j] = (unsigned short)fj_herm_index((cp & 15) + (cq & 15),
                                               ((cp >> 4) & 15) + ((cq >> 4) & 15),
                                               (cp >> 8) + (cq >> 8));
    }
    for (int j = lane; j < na * nc * nb; j += nthr) kac[j] = 0.0;
    const double z = zeta_ac[bra];
    const double Px = P[3 * bra], Py = P[3 * bra + 1], Pz = P[3 * bra + 2];
    if (lane < 3) fj_kl_eaxis(ea + lane * axa, la, lc, z, XPA[3 * bra + lane], XPC[3 * bra + lane]);
    const double ip = Ip[bra];
    const int dcb = dc[bra];
    __syncthreads();

    // the kets: the y-th of gridDim.y equal slices (one slice: all of them)
    const int ksub = (n_ket + (int)gridDim.y - 1) / (int)gridDim.y;
    const int k_lo = (int)blockIdx.y * ksub, k_hi = min(n_ket, k_lo + ksub);
    for (int k = k_lo; k < k_hi; ++k) {
        if (ip * Iq[k] * maxD < tau) break;   // kets sorted by Iq: the rest are smaller
        const int blk = dcb + dd[k];
        if (blk < 0 || blk >= n_dc) continue;   // the host never names such a block
        const double e = eta[k], sb = s_bd[k];
        const double PQx = Px - Q[3 * k], PQy = Py - Q[3 * k + 1], PQz = Pz - Q[3 * k + 2];
        const double rho = z * e / (z + e);
        if (lane < 3) fj_kl_eaxis(eb + lane * axb, lb, ld, e, XQB[3 * k + lane], XQD[3 * k + lane]);
        const double* __restrict__ dm = Dcd + (size_t)blk * (nc * nd);
        for (int j = lane; j < nc * nd; j += nthr) dsh[j] = dm[j];
        if (lane == 0) {
            fj_kl_seed(wf, rho, rho * (PQx * PQx + PQy * PQy + PQz * PQz), L);
            ((L & 1) ? Rr1 : Rr0)[0] = wf[L];
        }
        __syncthreads();
        fj_kl_tower_n(Rr0, Rr1, wf, tw, L, PQx, PQy, PQz, lane, nthr);
        const double* __restrict__ R0 = Rr0;
        for (int ib = 0; ib < nb; ++ib) {
            const int cb = hc[ob + ib];
            const int bx = cb & 15, by = (cb >> 4) & 15, bz = cb >> 8;
            for (int j = lane; j < nc * nq; j += nthr) {
                const int c = j / nq, q = j - c * nq;
                const int cq = hc[q];
                const int qx = cq & 15, qy = (cq >> 4) & 15, qz = cq >> 8;
                double s = 0.0;
                for (int d = 0; d < nd; ++d) {
                    const int cd = hc[od + d];
                    const double ex = eb[(bx * (ld + 1) + (cd & 15)) * tb + qx];
                    const double ey = eb[axb + (by * (ld + 1) + ((cd >> 4) & 15)) * tb + qy];
                    const double ez = eb[2 * axb + (bz * (ld + 1) + (cd >> 8)) * tb + qz];
                    s += dsh[c * nd + d] * (ex * ey * ez);
                }
                X[j] = ((qx + qy + qz) & 1 ? -sb : sb) * s;
            }
            __syncthreads();
            // Y[c][p] = sum_q [p + q] X[c][q] in 4 x 4 register tiles of (c, p):
            // per q four X and four tower entries feed sixteen multiply-adds
            {
                const int tp = (np + 3) >> 2, nt = ((nc + 3) >> 2) * tp;
                for (int t = lane; t < nt; t += nthr) {
                    const int c0 = (t / tp) << 2, p0 = (t - (t / tp) * tp) << 2;
                    const double* xr[4];
                    const unsigned short* ir[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        xr[i] = X + min(c0 + i, nc - 1) * nq;
                        ir[i] = pqi + min(p0 + i, np - 1) * nq;
                    }
                    double acc[4][4];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j2 = 0; j2 < 4; ++j2) acc[i][j2] = 0.0;
                    for (int q = 0; q < nq; ++q) {
                        double xv[4], rv[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i) xv[i] = xr[i][q];
#pragma unroll
                        for (int j2 = 0; j2 < 4; ++j2) rv[j2] = R0[ir[j2][q]];
#pragma unroll
                        for (int i = 0; i < 4; ++i)
#pragma unroll
                            for (int j2 = 0; j2 < 4; ++j2) acc[i][j2] += rv[j2] * xv[i];
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j2 = 0; j2 < 4; ++j2)
                            if (c0 + i < nc && p0 + j2 < np) Y[(c0 + i) * np + p0 + j2] = acc[i][j2];
                }
            }
            __syncthreads();
            // X of the next b is written before the next barrier, Y after it
            for (int j = lane; j < na * nc; j += nthr) {
                const int a = j / nc, c = j - a * nc;
                const int ca = hc[oa + a], cc = hc[oc + c];
                const double* __restrict__ ex = ea + ((ca & 15) * (lc + 1) + (cc & 15)) * ta;
                const double* __restrict__ ey = ea + axa + (((ca >> 4) & 15) * (lc + 1) + ((cc >> 4) & 15)) * ta;
                const double* __restrict__ ez = ea + 2 * axa + ((ca >> 8) * (lc + 1) + (cc >> 8)) * ta;
                const double* __restrict__ yr = Y + c * np;
                double s = 0.0;
                for (int p = 0; p < np; ++p) {
                    const int cp = hc[p];
                    s += (ex[cp & 15] * ey[(cp >> 4) & 15] * ez[cp >> 8]) * yr[p];
                }
                kac[j * nb + ib] += s;
            }
        }
    }
    __syncthreads();
    const double sa = s_ac[bra];
    for (int j = lane; j < na * nb; j += nthr) {
        const int a = j / nb, b = j - a * nb;
        double s = 0.0;
        for (int c = 0; c < nc; ++c) s += kac[(a * nc + c) * nb + b];
        // one slice writes its bra's block; several add their shares into
        // the block the launcher cleared
        if (gridDim.y == 1)
            K[(size_t)bra * (na * nb) + j] = sa * s;
        else
            unsafeAtomicAdd(K + (size_t)bra * (na * nb) + j, sa * s);
    }
}

// Schwarz: Q_mn = sqrt(max((mn|mn), 0)) for every AO pair of one shell pair
// (upstream eri/kernel_schwarz.py:249-412): (ab|ab) = sum_{i,k records}
// s_i s_k sum_p E^i_p sum_q (-1)^|q| [p + q] E^k_q, p and q in the box
// t_d <= a_d + b_d. One wavefront per shell pair; per record pair the E
// tables and the tower in LDS, then lanes over the component rows (a, b).
__global__ void __launch_bounds__(64) fj_schwarz_loop_k(
    int la, int lb, int n_pair, int n_ao, const double* __restrict__ zeta_ab,
    const double* __restrict__ P, const double* __restrict__ XPA, const double* __restrict__ XPB,
    const double* __restrict__ s_ab, const int* __restrict__ rec0, const int* __restrict__ rec1,
    const int* __restrict__ off_a, const int* __restrict__ off_b, double* __restrict__ Q) {
    __shared__ double R[2][FJ_KL_NH];
    __shared__ double ei[3 * FJ_KL_AX], ek[3 * FJ_KL_AX];
    __shared__ double wf[FJ_KL_MMAX + 1];
    __shared__ int tw[FJ_KL_NH];

    const int pair = (int)blockIdx.x, lane = (int)threadIdx.x;
    if (pair >= n_pair) return;   // uniform per workgroup
    const int L = 2 * (la + lb), nh = fj_herm_base(L + 1);
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2, nr = na * nb;
    const int nj = lb + 1, tt = la + lb + 1;   // E table strides
    fj_kl_steps(tw, nullptr, nh, lane);
    double acc[2] = {0.0, 0.0};   // rows lane and lane + 64 (nr <= 100)
    const int r0 = rec0[pair], r1 = rec1[pair];
    __syncthreads();
    for (int i = r0; i < r1; ++i) {
        for (int k = r0; k < r1; ++k) {
            const double zi = zeta_ab[i], zk = zeta_ab[k];
            if (lane < 3)
                fj_kl_eaxis(ei + lane * FJ_KL_AX, la, lb, zi, XPA[3 * i + lane], XPB[3 * i + lane]);
            else if (lane < 6)
                fj_kl_eaxis(ek + (lane - 3) * FJ_KL_AX, la, lb, zk, XPA[3 * k + lane - 3],
                            XPB[3 * k + lane - 3]);
            const double PQx = P[3 * i] - P[3 * k], PQy = P[3 * i + 1] - P[3 * k + 1],
                         PQz = P[3 * i + 2] - P[3 * k + 2];
            const double rho = zi * zk / (zi + zk);
            if (lane == 0) {
                fj_kl_seed(wf, rho, rho * (PQx * PQx + PQy * PQy + PQz * PQz), L);
                R[L & 1][0] = wf[L];
            }
            __syncthreads();
            fj_kl_tower(R[0], R[1], wf, tw, L, PQx, PQy, PQz, lane);
            const double sik = s_ab[i] * s_ab[k];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int r = lane + 64 * h;
                if (r >= nr) continue;
                const int a = r / nb, b = r - a * nb;
                const int ca = fj_herm_decode(fj_herm_base(la) + a), cb = fj_herm_decode(fj_herm_base(lb) + b);
                const int ax = ca & 15, ay = (ca >> 4) & 15, az = ca >> 8;
                const int bx = cb & 15, by = (cb >> 4) & 15, bz = cb >> 8;
                const double* __restrict__ ix = ei + (ax * nj + bx) * tt;
                const double* __restrict__ iy = ei + FJ_KL_AX + (ay * nj + by) * tt;
                const double* __restrict__ iz = ei + 2 * FJ_KL_AX + (az * nj + bz) * tt;
                const double* __restrict__ kx = ek + (ax * nj + bx) * tt;
                const double* __restrict__ ky = ek + FJ_KL_AX + (ay * nj + by) * tt;
                const double* __restrict__ kz = ek + 2 * FJ_KL_AX + (az * nj + bz) * tt;
                double v = 0.0;
                for (int px = 0; px <= ax + bx; ++px)
                    for (int py = 0; py <= ay + by; ++py)
                        for (int pz = 0; pz <= az + bz; ++pz) {
                            double w = 0.0;
                            for (int qx = 0; qx <= ax + bx; ++qx)
                                for (int qy = 0; qy <= ay + by; ++qy)
                                    for (int qz = 0; qz <= az + bz; ++qz) {
                                        const double e = kx[qx] * ky[qy] * kz[qz];
                                        const double rv = R[0][fj_herm_index(px + qx, py + qy, pz + qz)];
                                        w += ((qx + qy + qz) & 1 ? -e : e) * rv;
                                    }
                            v += (ix[px] * iy[py] * iz[pz]) * w;
                        }
                acc[h] += sik * v;
            }
            __syncthreads();   // before the next record pair rewrites ei, ek, wf and R
        }
    }
    const int oa = off_a[pair], ob = off_b[pair];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int r = lane + 64 * h;
        if (r >= nr) continue;
        const int a = r / nb, b = r - a * nb;
        const double q = sqrt(fmax(acc[h], 0.0));
        if (la != lb || oa != ob || a >= b) {
            Q[(size_t)(oa + a) * n_ao + ob + b] = q;
            Q[(size_t)(ob + b) * n_ao + oa + a] = q;
        }
    }
}

#endif  // FJ_KLOOP_HELPERS_ONLY
