// fj_floop.h -- the fused F kernel in loop form for the classes with an f
// shell: the same work items, screen, degeneracy and six-block digestion as
// the generated HGP kernels (upstream eri/fock/kernel_f.py:439-487,529-628),
// with each primitive quartet's integrals by McMurchie-Davidson instead of
// the HGP VRR/HRR:
//
//   (ab|cd) = U_ab U_cd sum_p E^{ab}_p sum_q (-1)^|q| E^{cd}_q [p + q]
//
// (upstream mmd/kernel.py:232-487 for the Hermite form; the records' U
// carries exp(-mu |AB|^2), so the E tables start from 1).
//
// One workgroup (FJ_FL_BLOCK lanes) per work item. The item's kets are taken one at a time; per
// live ket, the weighted density blocks and the six F accumulators sit in
// LDS, and per primitive quartet the lanes build the E tables and the
// Hermite tower, then, for the bra rows ab in chunks of FJ_FL_CH,
//   W[ab][q] = sum_p E^{ab}_p [p + q]           (lanes over 4 x 4 tiles of (ab, q))
//   x(ab,cd) = U U sum_q W[ab][q] (-1)^|q| E^{cd}_q   (lanes over 4 x 4 tiles of (ab, cd))
// and digest x at once (LDS atomics into the six blocks). In a tile the
// E products come from the per-axis tables and are shared across the tile's
// row or column: eight operands feed sixteen multiply-adds. After the ket's
// primitives the blocks go to F with one global atomic per element.
// Every barrier sits in control flow that is uniform over the wavefront.
//
// The includer provides fj_device.h and fj_kloop.h.
#pragma once

#define FJ_FL_CH 32   // bra rows per W chunk
#define FJ_FL_BLOCK 256   // lanes per workgroup
#define FJ_FL_WORKGROUPS 16384   // workgroups a work-list launch is cut into, at most

// The kernel's LDS, sized for its class and carved from one dynamic block:
// byte offsets of each array (doubles, then ints, then the u16 index table)
// and the total. The launcher and the kernel both use this layout.
struct FjFlLds {
    int R0, R1, eab, ecd, W, w[6], f[6], wf, hc, tw, pqi, bytes, axa, axc;
};

__host__ __device__ inline FjFlLds fj_floop_lds(int la, int lb, int lc, int ld) {
    auto herm = [](int n) { return n * (n + 1) * (n + 2) / 6; };
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
    const int nc = (lc + 1) * (lc + 2) / 2, nd = (ld + 1) * (ld + 2) / 2;
    const int np = herm(la + lb + 1), nq = herm(lc + ld + 1), nh = herm(la + lb + lc + ld + 1);
    FjFlLds o;
    o.axa = (la + 1) * (lb + 1) * (la + lb + 1);
    o.axc = (lc + 1) * (ld + 1) * (lc + ld + 1);
    const int blk[6] = {na * nb, nc * nd, na * nc, na * nd, nb * nc, nb * nd};
    int at = 0;
    auto take = [&](int n, int size) { const int r = at; at += ((n * size + 7) / 8) * 8; return r; };
    o.R0 = take(nh, 8);
    o.R1 = take(nh, 8);
    o.eab = take(3 * o.axa, 8);
    o.ecd = take(3 * o.axc, 8);
    o.W = take(FJ_FL_CH * nq, 8);
    for (int i = 0; i < 6; ++i) o.w[i] = take(blk[i], 8);
    for (int i = 0; i < 6; ++i) o.f[i] = take(blk[i], 8);
    o.wf = take(FJ_KL_MMAX + 1, 8);
    o.hc = take(nh, 4);
    o.tw = take(nh, 4);
    o.pqi = take(np * nq, 2);
    o.bytes = at;
    return o;
}

__global__ void __launch_bounds__(256) fj_floop_k(int la, int lb, int lc, int ld, int span, FjFArgs a) {
    FJ_DYN_SHARED(fj_floop_lds_block);
    const int lane = (int)threadIdx.x, nthr = (int)blockDim.x;
    // the work item: (bra pair, first ket, system | bra chunk, ket end), or
    // (x, y * span, the single system, n_ket) without a work list
    int bra, k0, k1, kb_lo = 0, kb_n = 0;
    FjSys S;
    if (a.work) {
        // an item covers at most `span` kets; workgroup (x, y) takes the
        // y-th of gridDim.y equal slices of item x's kets
        const int4 w = a.work[blockIdx.x];
        const int end = min(w.w, w.y + span), sub = (span + (int)gridDim.y - 1) / (int)gridDim.y;
        bra = w.x;
        k0 = w.y + (int)blockIdx.y * sub;
        k1 = min(end, k0 + sub);
        S = a.sys[w.z & 0xffff];
        kb_lo = (w.z >> 16) & 0xff;
        kb_n = (w.z >> 24) & 0xff;
    } else {
        bra = (int)blockIdx.x;
        k0 = (int)blockIdx.y * span;
        S = a.one;
        k1 = min(a.ket.n, k0 + span);
    }
    const int kmin = a.diag * bra;
    if (k0 >= k1 || (!a.ket_list && k0 + span <= kmin)) return;   // uniform
    int kb0, kab;
    fj_side_range(a.bra, bra, kb0, kab);
    if (kab > FJ_MAX_PRIM_PAIRS || kb_lo >= kab) return;
    kb0 += kb_lo;
    kab = kb_n ? min(kab - kb_lo, kb_n) : kab - kb_lo;

    const int lp = la + lb, lq = lc + ld, L = lp + lq;
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
    const int nc = (lc + 1) * (lc + 2) / 2, nd = (ld + 1) * (ld + 2) / 2;
    const int nab = na * nb, ncd = nc * nd;
    const int np = fj_herm_base(lp + 1), nq = fj_herm_base(lq + 1), nh = fj_herm_base(L + 1);
    const int ha = fj_herm_base(la), hb = fj_herm_base(lb), hcc = fj_herm_base(lc),
              hd = fj_herm_base(ld);
    const int ta = lp + 1, tc = lq + 1;   // t extents of the two E tables

    // the class's LDS arrays
    const FjFlLds lay = fj_floop_lds(la, lb, lc, ld);
    char* const lds = reinterpret_cast<char*>(fj_floop_lds_block);
    double* const Rr0 = reinterpret_cast<double*>(lds + lay.R0);
    double* const Rr1 = reinterpret_cast<double*>(lds + lay.R1);
    double* const eab = reinterpret_cast<double*>(lds + lay.eab);
    double* const ecd = reinterpret_cast<double*>(lds + lay.ecd);
    double* const W = reinterpret_cast<double*>(lds + lay.W);
    double* const wab = reinterpret_cast<double*>(lds + lay.w[0]);
    double* const wcd = reinterpret_cast<double*>(lds + lay.w[1]);
    double* const wac = reinterpret_cast<double*>(lds + lay.w[2]);
    double* const wad = reinterpret_cast<double*>(lds + lay.w[3]);
    double* const wbc = reinterpret_cast<double*>(lds + lay.w[4]);
    double* const wbd = reinterpret_cast<double*>(lds + lay.w[5]);
    double* const fab = reinterpret_cast<double*>(lds + lay.f[0]);
    double* const fcd = reinterpret_cast<double*>(lds + lay.f[1]);
    double* const fac = reinterpret_cast<double*>(lds + lay.f[2]);
    double* const fad = reinterpret_cast<double*>(lds + lay.f[3]);
    double* const fbc = reinterpret_cast<double*>(lds + lay.f[4]);
    double* const fbd = reinterpret_cast<double*>(lds + lay.f[5]);
    double* const wf = reinterpret_cast<double*>(lds + lay.wf);
    int* const hc = reinterpret_cast<int*>(lds + lay.hc);
    int* const tw = reinterpret_cast<int*>(lds + lay.tw);
    unsigned short* const pqi = reinterpret_cast<unsigned short*>(lds + lay.pqi);
    const int axa = lay.axa, axc = lay.axc;
    fj_kl_steps_n(tw, hc, nh, lane, nthr);
    for (int j = lane; j < np * nq; j += nthr) {
        const int p = j / nq, q = j - p * nq;
        const int cp = fj_herm_decode(p), cq = fj_herm_decode(q);
        pqi[j] = (unsigned short)fj_herm_index((cp & 15) + (cq & 15),
                                               ((cp >> 4) & 15) + ((cq >> 4) & 15),
                                               (cp >> 8) + (cq >> 8));
    }
    const int sa = a.bra.sh[2 * bra], sb = a.bra.sh[2 * bra + 1];
    const int oa = a.bra.ao[2 * bra], ob = a.bra.ao[2 * bra + 1];
    const double Gb = a.bra.G[bra];
    const double ABx = lb ? a.bra.sep[3 * bra] : 0.0, ABy = lb ? a.bra.sep[3 * bra + 1] : 0.0,
                 ABz = lb ? a.bra.sep[3 * bra + 2] : 0.0;
    const int n_ao = S.n_ao, ns = S.n_sh;
    const double* __restrict__ D = S.D;
    __syncthreads();

    for (int kk = k0; kk < k1; ++kk) {
        int k = kk, kc_lo = 0, kc_n = 0;
        if (a.ket_list) {
            const int v = a.ket_list[kk];
            k = v & 0x1ffff;
            kc_lo = (v >> 17) & 0x7f;
            kc_n = (v >> 24) & 0x7f;
        }
        if (k < kmin) continue;   // uniform: every lane reads the same ket
        const int sc = a.ket.sh[2 * k], sd = a.ket.sh[2 * k + 1];
        const double* __restrict__ Dsh = S.Dsh;
        double dmax = 2.0 * Dsh[sa * ns + sb];
        dmax = fmax(dmax, 2.0 * Dsh[sc * ns + sd]);
        dmax = fmax(dmax, 0.5 * Dsh[sa * ns + sc]);
        dmax = fmax(dmax, 0.5 * Dsh[sa * ns + sd]);
        dmax = fmax(dmax, 0.5 * Dsh[sb * ns + sc]);
        dmax = fmax(dmax, 0.5 * Dsh[sb * ns + sd]);
        if (Gb * (a.ket.G[k] * dmax) < a.tau) continue;
        int kc0, kcd;
        fj_side_range(a.ket, k, kc0, kcd);
        if (kcd > FJ_MAX_PRIM_PAIRS || kcd <= kc_lo) continue;
        kc0 += kc_lo;
        kcd = kc_n ? min(kcd - kc_lo, kc_n) : kcd - kc_lo;
        double deg = 1.0;
        if (sa == sb) deg *= 0.5;
        if (sc == sd) deg *= 0.5;
        if (a.diag && k == bra) deg *= 0.5;
        const double s_j = 2.0 * deg, s_k = -0.5 * deg;
        const int oc = a.ket.ao[2 * k], od = a.ket.ao[2 * k + 1];
        const double CDx = ld ? a.ket.sep[3 * k] : 0.0, CDy = ld ? a.ket.sep[3 * k + 1] : 0.0,
                     CDz = ld ? a.ket.sep[3 * k + 2] : 0.0;
        for (int j = lane; j < nab; j += nthr) {
            wab[j] = s_j * D[(size_t)(oa + j / nb) * n_ao + ob + j % nb];
            fab[j] = 0.0;
        }
        for (int j = lane; j < ncd; j += nthr) {
            wcd[j] = s_j * D[(size_t)(oc + j / nd) * n_ao + od + j % nd];
            fcd[j] = 0.0;
        }
        for (int j = lane; j < na * nc; j += nthr) {
            wac[j] = s_k * D[(size_t)(oa + j / nc) * n_ao + oc + j % nc];
            fac[j] = 0.0;
        }
        for (int j = lane; j < na * nd; j += nthr) {
            wad[j] = s_k * D[(size_t)(oa + j / nd) * n_ao + od + j % nd];
            fad[j] = 0.0;
        }
        for (int j = lane; j < nb * nc; j += nthr) {
            wbc[j] = s_k * D[(size_t)(ob + j / nc) * n_ao + oc + j % nc];
            fbc[j] = 0.0;
        }
        for (int j = lane; j < nb * nd; j += nthr) {
            wbd[j] = s_k * D[(size_t)(ob + j / nd) * n_ao + od + j % nd];
            fbd[j] = 0.0;
        }
        // the side of higher angular momentum in the outer loop: its E tables
        // are built once per outer primitive pair, the other side's per pair
        // of primitive pairs
        const bool ket_outer = lq > lp;
        const int n_out = ket_outer ? kcd : kab, n_in = ket_outer ? kab : kcd;
        for (int io = 0; io < n_out; ++io) {
            for (int ii = 0; ii < n_in; ++ii) {
                const int ib = kb0 + (ket_outer ? ii : io), ik = kc0 + (ket_outer ? io : ii);
                const double* __restrict__ rb = a.bra.prim + 8 * (size_t)ib;
                const double* __restrict__ rk = a.ket.prim + 8 * (size_t)ik;
                const double p = rb[6], q = rk[6];
                const bool bra_new = ket_outer || ii == 0, ket_new = !ket_outer || ii == 0;
                if (lane < 3) {
                    if (bra_new) {
                        const double pa = rb[3 + lane], ab = lane == 0 ? ABx : (lane == 1 ? ABy : ABz);
                        fj_kl_eaxis_from(eab + lane * axa, la, lb, p, pa, pa + ab, 1.0);
                    }
                } else if (lane < 6) {
                    if (ket_new) {
                        const int x = lane - 3;
                        const double qc = rk[3 + x], cd = x == 0 ? CDx : (x == 1 ? CDy : CDz);
                        fj_kl_eaxis_from(ecd + x * axc, lc, ld, q, qc, qc + cd, 1.0);
                    }
                }
                const double PQx = rb[0] - rk[0], PQy = rb[1] - rk[1], PQz = rb[2] - rk[2];
                const double rho = p * q / (p + q);
                if (lane == 0) {
                    fj_kl_seed(wf, rho, rho * (PQx * PQx + PQy * PQy + PQz * PQz), L);
                    ((L & 1) ? Rr1 : Rr0)[0] = wf[L];
                }
                __syncthreads();
                fj_kl_tower_n(Rr0, Rr1, wf, tw, L, PQx, PQy, PQz, lane, nthr);
                const double* __restrict__ R0 = Rr0;
                const double UU = rb[7] * rk[7];
                for (int r0 = 0; r0 < nab; r0 += FJ_FL_CH) {
                    const int nr = min(FJ_FL_CH, nab - r0);
                    // W[r][q] = sum_p E^{ab}_p [p + q] in 4 x 4 register tiles of
                    // (r, q): per p four E products (from the axis tables) and
                    // four tower entries make sixteen multiply-adds
                    {
                        const int tq = (nq + 3) >> 2, nt = ((nr + 3) >> 2) * tq;
                        for (int t = lane; t < nt; t += nthr) {
                            const int m0 = (t / tq) << 2, n0 = (t - (t / tq) * tq) << 2;
                            const double* ex[4];
                            const double* ey[4];
                            const double* ez[4];
                            int qq[4];
#pragma unroll
                            for (int i = 0; i < 4; ++i) {
                                const int ab = r0 + min(m0 + i, nr - 1);
                                const int ca = hc[ha + ab / nb], cb = hc[hb + ab % nb];
                                ex[i] = eab + ((ca & 15) * (lb + 1) + (cb & 15)) * ta;
                                ey[i] = eab + axa + (((ca >> 4) & 15) * (lb + 1) + ((cb >> 4) & 15)) * ta;
                                ez[i] = eab + 2 * axa + ((ca >> 8) * (lb + 1) + (cb >> 8)) * ta;
                                qq[i] = min(n0 + i, nq - 1);
                            }
                            double c[4][4];
#pragma unroll
                            for (int i = 0; i < 4; ++i)
#pragma unroll
                                for (int j = 0; j < 4; ++j) c[i][j] = 0.0;
                            for (int pi = 0; pi < np; ++pi) {
                                const int cp = hc[pi];
                                const int px = cp & 15, py = (cp >> 4) & 15, pz = cp >> 8;
                                const unsigned short* __restrict__ pr = pqi + pi * nq;
                                double ea[4], rv[4];
#pragma unroll
                                for (int i = 0; i < 4; ++i) ea[i] = ex[i][px] * ey[i][py] * ez[i][pz];
#pragma unroll
                                for (int j = 0; j < 4; ++j) rv[j] = R0[pr[qq[j]]];
#pragma unroll
                                for (int i = 0; i < 4; ++i)
#pragma unroll
                                    for (int j = 0; j < 4; ++j) c[i][j] += ea[i] * rv[j];
                            }
#pragma unroll
                            for (int i = 0; i < 4; ++i)
#pragma unroll
                                for (int j = 0; j < 4; ++j)
                                    if (m0 + i < nr && n0 + j < nq) W[(m0 + i) * nq + n0 + j] = c[i][j];
                        }
                    }
                    __syncthreads();
                    // x(ab, cd) = U U sum_q W[r][q] (-1)^|q| E^{cd}_q in 4 x 4 tiles
                    // of (r, cd), each element digested at once into the six blocks
                    {
                        const int tcd = (ncd + 3) >> 2, nt = ((nr + 3) >> 2) * tcd;
                        for (int t = lane; t < nt; t += nthr) {
                            const int m0 = (t / tcd) << 2, n0 = (t - (t / tcd) * tcd) << 2;
                            const double* fx[4];
                            const double* fy[4];
                            const double* fz[4];
                            int wr[4];
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const int cdi = min(n0 + j, ncd - 1);
                                const int cc = hc[hcc + cdi / nd], cdd = hc[hd + cdi % nd];
                                fx[j] = ecd + ((cc & 15) * (ld + 1) + (cdd & 15)) * tc;
                                fy[j] = ecd + axc + (((cc >> 4) & 15) * (ld + 1) + ((cdd >> 4) & 15)) * tc;
                                fz[j] = ecd + 2 * axc + ((cc >> 8) * (ld + 1) + (cdd >> 8)) * tc;
                                wr[j] = min(m0 + j, nr - 1) * nq;
                            }
                            double c[4][4];
#pragma unroll
                            for (int i = 0; i < 4; ++i)
#pragma unroll
                                for (int j = 0; j < 4; ++j) c[i][j] = 0.0;
                            for (int qi = 0; qi < nq; ++qi) {
                                const int cq = hc[qi];
                                const int qx = cq & 15, qy = (cq >> 4) & 15, qz = cq >> 8;
                                const bool neg = (qx + qy + qz) & 1;
                                double w[4], e[4];
#pragma unroll
                                for (int i = 0; i < 4; ++i) w[i] = W[wr[i] + qi];
#pragma unroll
                                for (int j = 0; j < 4; ++j) {
                                    const double v = fx[j][qx] * fy[j][qy] * fz[j][qz];
                                    e[j] = neg ? -v : v;
                                }
#pragma unroll
                                for (int i = 0; i < 4; ++i)
#pragma unroll
                                    for (int j = 0; j < 4; ++j) c[i][j] += w[i] * e[j];
                            }
#pragma unroll
                            for (int i = 0; i < 4; ++i) {
                                if (m0 + i >= nr) continue;
                                const int ab = r0 + m0 + i;
                                const int ia = ab / nb, jb = ab % nb;
#pragma unroll
                                for (int j = 0; j < 4; ++j) {
                                    const int cdi = n0 + j;
                                    if (cdi >= ncd) continue;
                                    const int kc = cdi / nd, ld_ = cdi % nd;
                                    const double x = UU * c[i][j];
                                    atomicAdd(&fab[ab], wcd[cdi] * x);
                                    atomicAdd(&fcd[cdi], wab[ab] * x);
                                    atomicAdd(&fac[ia * nc + kc], wbd[jb * nd + ld_] * x);
                                    atomicAdd(&fad[ia * nd + ld_], wbc[jb * nc + kc] * x);
                                    atomicAdd(&fbc[jb * nc + kc], wad[ia * nd + ld_] * x);
                                    atomicAdd(&fbd[jb * nd + ld_], wac[ia * nc + kc] * x);
                                }
                            }
                        }
                    }
                    __syncthreads();   // W and the E tables are rewritten next
                }
            }
        }
        // the ket's six blocks into F (the caller closes F + F^T)
        for (int j = lane; j < nab; j += nthr)
            if (fab[j] != 0.0) unsafeAtomicAdd(S.F + (size_t)(oa + j / nb) * n_ao + ob + j % nb, fab[j]);
        for (int j = lane; j < ncd; j += nthr)
            if (fcd[j] != 0.0) unsafeAtomicAdd(S.F + (size_t)(oc + j / nd) * n_ao + od + j % nd, fcd[j]);
        for (int j = lane; j < na * nc; j += nthr)
            if (fac[j] != 0.0) unsafeAtomicAdd(S.F + (size_t)(oa + j / nc) * n_ao + oc + j % nc, fac[j]);
        for (int j = lane; j < na * nd; j += nthr)
            if (fad[j] != 0.0) unsafeAtomicAdd(S.F + (size_t)(oa + j / nd) * n_ao + od + j % nd, fad[j]);
        for (int j = lane; j < nb * nc; j += nthr)
            if (fbc[j] != 0.0) unsafeAtomicAdd(S.F + (size_t)(ob + j / nc) * n_ao + oc + j % nc, fbc[j]);
        for (int j = lane; j < nb * nd; j += nthr)
            if (fbd[j] != 0.0) unsafeAtomicAdd(S.F + (size_t)(ob + j / nd) * n_ao + od + j % nd, fbd[j]);
        __syncthreads();   // before the next ket rewrites the blocks
    }
}
