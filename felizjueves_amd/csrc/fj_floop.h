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
//   W[ab][q] = sum_p E^{ab}_p [p + q]                  (16 x 16 tiles of (ab, q))
//   x(ab,cd) = U U sum_q W[ab][q] (-1)^|q| E^{cd}_q     (16 x 16 tiles of (ab, cd))
// on the FP64 matrix cores (a wavefront per tile, four Hermite indices per
// step; each lane forms its A and B elements from the axis tables, the tower
// and W), and digest x at once (LDS atomics into the six blocks). After the
// ket's primitives the blocks go to F with one global atomic per element.
// Every barrier sits in control flow that is uniform over the wavefront.
//
// fj_fbloop_k is the batched form for the classes whose whole W fits LDS
// several times over (fj_fbloop_plan: B >= 2): it takes a ket's primitive
// quartets B at a time, every phase over the batch at once (lanes over the
// E tables and seeds of the batch, over (b, t) in each tower level, the
// matrix cores over (b, tile) in W), so a batch pays the L + 3 barriers and
// the serial Boys and E-table latencies that the one-at-a-time form pays per
// primitive quartet; x is one matrix-core GEMM over the batch (K = (b, q)),
// digested once per batch instead of once per primitive quartet. Classes
// with fewer than FJ_FB_MFMA_ROWS bra functions, whose 16-row tiles would
// leave the matrix cores mostly idle, contract in 4 x 4 register tiles
// instead (W lanes over (b, tile); x summed over the batch in registers,
// lanes over (slice, tile), each slice digesting its share once).
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

// The workgroup's share of a launch: the bra pair, its kets [k0, k1), the
// system and the bra's primitive-pair range [kb0, kb0 + kab). The work item
// is (bra pair, first ket, system | bra chunk, ket end), or (x, y * span,
// the single system, n_ket) without a work list; an item covers at most
// `span` kets and, with a work list, workgroup (x, y) takes the y-th of
// gridDim.y equal slices of item x's kets. False (uniformly) when there is
// nothing to do.
struct FjFlItem {
    int bra, k0, k1, kmin, kb0, kab;
    FjSys S;
};

__device__ __forceinline__ bool fj_fl_item(const FjFArgs& a, int span, FjFlItem& it) {
    int kb_lo = 0, kb_n = 0;
    if (a.work) {
        const int4 w = a.work[blockIdx.x];
        const int end = min(w.w, w.y + span), sub = (span + (int)gridDim.y - 1) / (int)gridDim.y;
        it.bra = w.x;
        it.k0 = w.y + (int)blockIdx.y * sub;
        it.k1 = min(end, it.k0 + sub);
        it.S = a.sys[w.z & 0xffff];
        kb_lo = (w.z >> 16) & 0xff;
        kb_n = (w.z >> 24) & 0xff;
    } else {
        it.bra = (int)blockIdx.x;
        it.k0 = (int)blockIdx.y * span;
        it.S = a.one;
        it.k1 = min(a.ket.n, it.k0 + span);
    }
    it.kmin = a.diag * it.bra;
    if (it.k0 >= it.k1 || (!a.ket_list && it.k0 + span <= it.kmin)) return false;
    fj_side_range(a.bra, it.bra, it.kb0, it.kab);
    if (it.kab > FJ_MAX_PRIM_PAIRS || kb_lo >= it.kab) return false;
    it.kb0 += kb_lo;
    it.kab = kb_n ? min(it.kab - kb_lo, kb_n) : it.kab - kb_lo;
    return true;
}

// One ket of the item: its pair index, primitive-pair range [kc0, kc0 +
// kcd), shells, AO offsets and the J / K digestion scales of the quartet's
// degeneracy. False (uniformly) when it is below the diagonal, screened out
// (G_ab G_cd times the largest density block it digests into, under tau)
// or empty.
struct FjFlKet {
    int k, kc0, kcd, sc, sd, oc, od;
    double s_j, s_k;
};

__device__ __forceinline__ bool fj_fl_ket(const FjFArgs& a, const FjFlItem& it, int kk, int sa, int sb,
                                          double Gb, FjFlKet& q) {
    int kc_lo = 0, kc_n = 0;
    q.k = kk;
    if (a.ket_list) {
        const int v = a.ket_list[kk];
        q.k = v & 0x1ffff;
        kc_lo = (v >> 17) & 0x7f;
        kc_n = (v >> 24) & 0x7f;
    }
    const int k = q.k;
    if (k < it.kmin) return false;
    q.sc = a.ket.sh[2 * k];
    q.sd = a.ket.sh[2 * k + 1];
    const int ns = it.S.n_sh;
    const double* __restrict__ Dsh = it.S.Dsh;
    double dmax = 2.0 * Dsh[sa * ns + sb];
    dmax = fmax(dmax, 2.0 * Dsh[q.sc * ns + q.sd]);
    dmax = fmax(dmax, 0.5 * Dsh[sa * ns + q.sc]);
    dmax = fmax(dmax, 0.5 * Dsh[sa * ns + q.sd]);
    dmax = fmax(dmax, 0.5 * Dsh[sb * ns + q.sc]);
    dmax = fmax(dmax, 0.5 * Dsh[sb * ns + q.sd]);
    if (Gb * (a.ket.G[k] * dmax) < a.tau) return false;
    fj_side_range(a.ket, k, q.kc0, q.kcd);
    if (q.kcd > FJ_MAX_PRIM_PAIRS || q.kcd <= kc_lo) return false;
    q.kc0 += kc_lo;
    q.kcd = kc_n ? min(q.kcd - kc_lo, kc_n) : q.kcd - kc_lo;
    double deg = 1.0;
    if (sa == sb) deg *= 0.5;
    if (q.sc == q.sd) deg *= 0.5;
    if (a.diag && k == it.bra) deg *= 0.5;
    q.s_j = 2.0 * deg;
    q.s_k = -0.5 * deg;
    q.oc = a.ket.ao[2 * k];
    q.od = a.ket.ao[2 * k + 1];
    return true;
}

// The six blocks of a quartet in LDS, in the order ab, cd, ac, ad, bc, bd:
// w the weighted density blocks (J scale on ab and cd, K scale on the
// others), f the F accumulators. fj_fl_load fills w and zeroes f;
// fj_fl_flush adds f into F with one global atomic per nonzero element.
struct FjFlBlocks {
    double* w[6];
    double* f[6];
};

__device__ __forceinline__ void fj_fl_load(const FjFlBlocks& B, const double* __restrict__ D, int n_ao,
                                           int oa, int ob, int oc, int od, int na, int nb, int nc,
                                           int nd, double s_j, double s_k, int lane, int nthr) {
    const int r0[6] = {oa, oc, oa, oa, ob, ob}, c0[6] = {ob, od, oc, od, oc, od};
    const int nr[6] = {na, nc, na, na, nb, nb}, ncol[6] = {nb, nd, nc, nd, nc, nd};
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double sc = i < 2 ? s_j : s_k;
        for (int j = lane; j < nr[i] * ncol[i]; j += nthr) {
            B.w[i][j] = sc * D[(size_t)(r0[i] + j / ncol[i]) * n_ao + c0[i] + j % ncol[i]];
            B.f[i][j] = 0.0;
        }
    }
}

__device__ __forceinline__ void fj_fl_flush(const FjFlBlocks& B, double* F, int n_ao, int oa, int ob,
                                            int oc, int od, int na, int nb, int nc, int nd, int lane,
                                            int nthr) {
    const int r0[6] = {oa, oc, oa, oa, ob, ob}, c0[6] = {ob, od, oc, od, oc, od};
    const int nr[6] = {na, nc, na, na, nb, nb}, ncol[6] = {nb, nd, nc, nd, nc, nd};
#pragma unroll
    for (int i = 0; i < 6; ++i)
        for (int j = lane; j < nr[i] * ncol[i]; j += nthr)
            if (B.f[i][j] != 0.0)
                unsafeAtomicAdd(F + (size_t)(r0[i] + j / ncol[i]) * n_ao + c0[i] + j % ncol[i], B.f[i][j]);
}

// Digest one element x of (ab|cd), ab = (ia, jb), cd = (kc, ld_), into the
// six LDS blocks.
__device__ __forceinline__ void fj_fl_digest(const FjFlBlocks& B, int nb, int nc, int nd, int ab,
                                             int cdi, double x) {
    const int ia = ab / nb, jb = ab - ia * nb, kc = cdi / nd, ld_ = cdi - kc * nd;
    atomicAdd(&B.f[0][ab], B.w[1][cdi] * x);
    atomicAdd(&B.f[1][cdi], B.w[0][ab] * x);
    atomicAdd(&B.f[2][ia * nc + kc], B.w[5][jb * nd + ld_] * x);
    atomicAdd(&B.f[3][ia * nd + ld_], B.w[4][jb * nc + kc] * x);
    atomicAdd(&B.f[4][jb * nc + kc], B.w[3][ia * nd + ld_] * x);
    atomicAdd(&B.f[5][jb * nd + ld_], B.w[2][ia * nc + kc] * x);
}

// Three wavefronts per SIMD asked for (166 VGPRs, no scratch): the kernel is
// latency-bound, and where the class's LDS allows, a CU holds three
// workgroups instead of two.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) fj_floop_k(int la, int lb, int lc, int ld, int span, FjFArgs a) {
    FJ_DYN_SHARED(fj_floop_lds_block);
    const int lane = (int)threadIdx.x, nthr = (int)blockDim.x;
    const int wave = lane >> 6, lw = lane & 63, nw = nthr >> 6;   // wavefronts of the workgroup
    FjFlItem it;
    if (!fj_fl_item(a, span, it)) return;   // uniform
    const int bra = it.bra, kb0 = it.kb0, kab = it.kab;
    const FjSys& S = it.S;

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
    FjFlBlocks blk;
    for (int i = 0; i < 6; ++i) {
        blk.w[i] = reinterpret_cast<double*>(lds + lay.w[i]);
        blk.f[i] = reinterpret_cast<double*>(lds + lay.f[i]);
    }
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
    __syncthreads();

    for (int kk = it.k0; kk < it.k1; ++kk) {
        FjFlKet q;
        if (!fj_fl_ket(a, it, kk, sa, sb, Gb, q)) continue;   // uniform: every lane reads the same ket
        const int k = q.k, kc0 = q.kc0, kcd = q.kcd;
        const double CDx = ld ? a.ket.sep[3 * k] : 0.0, CDy = ld ? a.ket.sep[3 * k + 1] : 0.0,
                     CDz = ld ? a.ket.sep[3 * k + 2] : 0.0;
        fj_fl_load(blk, S.D, S.n_ao, oa, ob, q.oc, q.od, na, nb, nc, nd, q.s_j, q.s_k, lane, nthr);
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
                    // W[r][q] = sum_p E^{ab}_p [p + q] on the matrix cores: 16 x 16
                    // tiles of (r, q), p in steps of 4, wavefront w taking tiles w +
                    // 2 nw i and w + 2 nw i + nw together (two independent chains);
                    // a lane's A element is an E product from the axis tables, its
                    // B element a tower entry
                    {
                        const int tqw = (nq + 15) >> 4, nt = ((nr + 15) >> 4) * tqw;
                        for (int t0 = wave; t0 < nt; t0 += 2 * nw) {   // uniform over the wavefront
                            const int t1 = t0 + nw < nt ? t0 + nw : t0;
                            const int rt0 = t0 / tqw, qt0 = t0 - rt0 * tqw, rt1 = t1 / tqw, qt1 = t1 - rt1 * tqw;
                            const int ra0 = (rt0 << 4) + (lw & 15), qb0 = (qt0 << 4) + (lw & 15);
                            const int ra1 = (rt1 << 4) + (lw & 15), qb1 = (qt1 << 4) + (lw & 15);
                            // a lane's A element needs its row valid, its B element its column
                            const bool two = t1 != t0;
                            const bool rok0 = ra0 < nr, qok0 = qb0 < nq, rok1 = two && ra1 < nr, qok1 = two && qb1 < nq;
                            const double* ex[2];
                            const double* ey[2];
                            const double* ez[2];
                            int qc[2];
#pragma unroll
                            for (int u = 0; u < 2; ++u) {
                                const int ab = r0 + min(u ? ra1 : ra0, nr - 1);
                                const int ca = hc[ha + ab / nb], cb = hc[hb + ab % nb];
                                ex[u] = eab + ((ca & 15) * (lb + 1) + (cb & 15)) * ta;
                                ey[u] = eab + axa + (((ca >> 4) & 15) * (lb + 1) + ((cb >> 4) & 15)) * ta;
                                ez[u] = eab + 2 * axa + ((ca >> 8) * (lb + 1) + (cb >> 8)) * ta;
                                qc[u] = min(u ? qb1 : qb0, nq - 1);
                            }
                            fj_v4d c0 = {0.0, 0.0, 0.0, 0.0}, c1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
                            for (int p0 = 0; p0 < np; p0 += 4) {
                                const int p = p0 + (lw >> 4), pc = min(p, np - 1);
                                const bool pok = p < np;
                                const int cp = hc[pc];
                                const int px = cp & 15, py = (cp >> 4) & 15, pz = cp >> 8;
                                const unsigned short* __restrict__ pr = pqi + pc * nq;
                                const double a0 = rok0 && pok ? ex[0][px] * ey[0][py] * ez[0][pz] : 0.0;
                                const double b0 = qok0 && pok ? R0[pr[qc[0]]] : 0.0;
                                const double a1 = rok1 && pok ? ex[1][px] * ey[1][py] * ez[1][pz] : 0.0;
                                const double b1 = qok1 && pok ? R0[pr[qc[1]]] : 0.0;
                                c0 = fj_mfma_f64_16x16x4(a0, b0, c0);
                                c1 = fj_mfma_f64_16x16x4(a1, b1, c1);
                            }
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const int r = (rt0 << 4) + (lw >> 4) + 4 * v;
                                if (r < nr && qb0 < nq) W[r * nq + qb0] = c0[v];
                            }
                            if (t1 != t0) {
#pragma unroll
                                for (int v = 0; v < 4; ++v) {
                                    const int r = (rt1 << 4) + (lw >> 4) + 4 * v;
                                    if (r < nr && qb1 < nq) W[r * nq + qb1] = c1[v];
                                }
                            }
                        }
                    }
                    __syncthreads();
                    // x(ab, cd) = U U sum_q W[r][q] (-1)^|q| E^{cd}_q on the matrix
                    // cores, 16 x 16 tiles of (r, cd) two at a time as above, each
                    // element digested at once into the six blocks
                    {
                        const int tcw = (ncd + 15) >> 4, nt = ((nr + 15) >> 4) * tcw;
                        for (int t0 = wave; t0 < nt; t0 += 2 * nw) {   // uniform over the wavefront
                            const int t1 = t0 + nw < nt ? t0 + nw : t0;
                            const int rt0 = t0 / tcw, ct0 = t0 - rt0 * tcw, rt1 = t1 / tcw, ct1 = t1 - rt1 * tcw;
                            const int ra0 = (rt0 << 4) + (lw & 15), cb0 = (ct0 << 4) + (lw & 15);
                            const int ra1 = (rt1 << 4) + (lw & 15), cb1 = (ct1 << 4) + (lw & 15);
                            const bool two = t1 != t0;
                            const bool rok0 = ra0 < nr, cok0 = cb0 < ncd, rok1 = two && ra1 < nr, cok1 = two && cb1 < ncd;
                            const double* fx[2];
                            const double* fy[2];
                            const double* fz[2];
                            int wr[2];
#pragma unroll
                            for (int u = 0; u < 2; ++u) {
                                const int cdi = min(u ? cb1 : cb0, ncd - 1);
                                const int cc = hc[hcc + cdi / nd], cdd = hc[hd + cdi % nd];
                                fx[u] = ecd + ((cc & 15) * (ld + 1) + (cdd & 15)) * tc;
                                fy[u] = ecd + axc + (((cc >> 4) & 15) * (ld + 1) + ((cdd >> 4) & 15)) * tc;
                                fz[u] = ecd + 2 * axc + ((cc >> 8) * (ld + 1) + (cdd >> 8)) * tc;
                                wr[u] = min(u ? ra1 : ra0, nr - 1) * nq;
                            }
                            fj_v4d c0 = {0.0, 0.0, 0.0, 0.0}, c1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
                            for (int q0 = 0; q0 < nq; q0 += 4) {
                                const int qq = q0 + (lw >> 4), qc = min(qq, nq - 1);
                                const bool qok = qq < nq;
                                const int cq = hc[qc];
                                const int qx = cq & 15, qy = (cq >> 4) & 15, qz = cq >> 8;
                                const bool neg = (qx + qy + qz) & 1;
                                const double e0 = fx[0][qx] * fy[0][qy] * fz[0][qz];
                                const double e1 = fx[1][qx] * fy[1][qy] * fz[1][qz];
                                const double a0 = rok0 && qok ? W[wr[0] + qc] : 0.0;
                                const double b0 = cok0 && qok ? (neg ? -e0 : e0) : 0.0;
                                const double a1 = rok1 && qok ? W[wr[1] + qc] : 0.0;
                                const double b1 = cok1 && qok ? (neg ? -e1 : e1) : 0.0;
                                c0 = fj_mfma_f64_16x16x4(a0, b0, c0);
                                c1 = fj_mfma_f64_16x16x4(a1, b1, c1);
                            }
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const int r = (rt0 << 4) + (lw >> 4) + 4 * v;
                                if (r < nr && cb0 < ncd) fj_fl_digest(blk, nb, nc, nd, r0 + r, cb0, UU * c0[v]);
                            }
                            if (t1 != t0) {
#pragma unroll
                                for (int v = 0; v < 4; ++v) {
                                    const int r = (rt1 << 4) + (lw >> 4) + 4 * v;
                                    if (r < nr && cb1 < ncd) fj_fl_digest(blk, nb, nc, nd, r0 + r, cb1, UU * c1[v]);
                                }
                            }
                        }
                    }
                    __syncthreads();   // W and the E tables are rewritten next
                }
            }
        }
        // the ket's six blocks into F (the caller closes F + F^T)
        fj_fl_flush(blk, S.F, S.n_ao, oa, ob, q.oc, q.od, na, nb, nc, nd, lane, nthr);
        __syncthreads();   // before the next ket rewrites the blocks
    }
}

// ---------------------------------------------------------------------------
// The batched form (fj_fbloop_k): B primitive quartets of one ket per pass.

#define FJ_FB_BUDGET (52 * 1024)   // LDS bytes per workgroup: three fit a CU's 160 KB
#define FJ_FB_BMAX 64              // primitive quartets per batch
#define FJ_FB_SC 4                 // per primitive quartet: PQ (3), U_ab U_cd
#define FJ_FB_MFMA_ROWS 8          // bra functions from which the contractions use the
                                   // matrix cores (16-row tiles; fewer rows leave them idle)

// The batched kernel's LDS for batch B: B copies of the per-primitive-quartet
// arrays (the two tower rows, the two sides' axis E tables, W [nab][nq], the
// seed and the scalars), then the six density and six F blocks, the index
// tables, and the total.
struct FjFbLds {
    int B, nh, nW, axa, axc;
    int R0, R1, eab, ecd, W, wf, sc, w[6], f[6], hc, tw, pqi, bytes;
};

__host__ __device__ inline FjFbLds fj_fbloop_layout(int la, int lb, int lc, int ld, int B) {
    auto herm = [](int n) { return n * (n + 1) * (n + 2) / 6; };
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
    const int nc = (lc + 1) * (lc + 2) / 2, nd = (ld + 1) * (ld + 2) / 2;
    const int L = la + lb + lc + ld;
    const int np = herm(la + lb + 1), nq = herm(lc + ld + 1);
    FjFbLds o;
    o.B = B;
    o.nh = herm(L + 1);
    o.nW = na * nb * nq;
    o.axa = (la + 1) * (lb + 1) * (la + lb + 1);
    o.axc = (lc + 1) * (ld + 1) * (lc + ld + 1);
    const int blk[6] = {na * nb, nc * nd, na * nc, na * nd, nb * nc, nb * nd};
    int at = 0;
    auto take = [&](int n, int size) { const int r = at; at += ((n * size + 7) / 8) * 8; return r; };
    o.R0 = take(B * o.nh, 8);
    o.R1 = take(B * o.nh, 8);
    o.eab = take(B * 3 * o.axa, 8);
    o.ecd = take(B * 3 * o.axc, 8);
    o.W = take(B * o.nW, 8);
    o.wf = take(B * (L + 1), 8);
    o.sc = take(B * FJ_FB_SC, 8);
    for (int i = 0; i < 6; ++i) o.w[i] = take(blk[i], 8);
    for (int i = 0; i < 6; ++i) o.f[i] = take(blk[i], 8);
    o.hc = take(o.nh, 4);
    o.tw = take(o.nh, 4);
    o.pqi = take(np * nq, 2);
    o.bytes = at;
    return o;
}

// B for a class: the largest that fits FJ_FB_BUDGET, at most bmax (a
// larger batch pays the per-batch barriers and latencies over more primitive
// quartets). B < 2 means the class runs on the one-at-a-time kernel.
__host__ __device__ inline FjFbLds fj_fbloop_plan(int la, int lb, int lc, int ld, int bmax = FJ_FB_BMAX) {
    const FjFbLds one = fj_fbloop_layout(la, lb, lc, ld, 1), two = fj_fbloop_layout(la, lb, lc, ld, 2);
    const int per_b = two.bytes - one.bytes, fixed = one.bytes - per_b;
    int B = (FJ_FB_BUDGET - fixed) / per_b;
    if (B > bmax) B = bmax;
    if (B < 1) B = 1;
    return fj_fbloop_layout(la, lb, lc, ld, B);
}

// Three wavefronts per SIMD asked for (168 VGPRs, 48 bytes of scratch per
// lane): with the 52 KB budget three workgroups share a CU.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) fj_fbloop_k(int la, int lb, int lc, int ld, int span, int B,
                                                   FjFArgs a) {
    FJ_DYN_SHARED(fj_fbloop_lds_block);
    const int lane = (int)threadIdx.x, nthr = (int)blockDim.x;
    const int wave = lane >> 6, lw = lane & 63, nw = nthr >> 6;   // wavefronts of the workgroup
    FjFlItem it;
    if (!fj_fl_item(a, span, it)) return;   // uniform
    const int bra = it.bra, kb0 = it.kb0, kab = it.kab;
    const FjSys& S = it.S;

    const int lp = la + lb, lq = lc + ld, L = lp + lq;
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
    const int nc = (lc + 1) * (lc + 2) / 2, nd = (ld + 1) * (ld + 2) / 2;
    const int nab = na * nb, ncd = nc * nd;
    const int np = fj_herm_base(lp + 1), nq = fj_herm_base(lq + 1);
    const int ha = fj_herm_base(la), hb = fj_herm_base(lb), hcc = fj_herm_base(lc),
              hd = fj_herm_base(ld);
    const int ta = lp + 1, tc = lq + 1;   // t extents of the two E tables

    const FjFbLds lay = fj_fbloop_layout(la, lb, lc, ld, B);
    const int nh = lay.nh, nW = lay.nW, axa = lay.axa, axc = lay.axc;
    char* const lds = reinterpret_cast<char*>(fj_fbloop_lds_block);
    double* const Rr0 = reinterpret_cast<double*>(lds + lay.R0);
    double* const Rr1 = reinterpret_cast<double*>(lds + lay.R1);
    double* const eab = reinterpret_cast<double*>(lds + lay.eab);
    double* const ecd = reinterpret_cast<double*>(lds + lay.ecd);
    double* const W = reinterpret_cast<double*>(lds + lay.W);
    double* const wf = reinterpret_cast<double*>(lds + lay.wf);
    double* const scl = reinterpret_cast<double*>(lds + lay.sc);
    FjFlBlocks blk;
    for (int i = 0; i < 6; ++i) {
        blk.w[i] = reinterpret_cast<double*>(lds + lay.w[i]);
        blk.f[i] = reinterpret_cast<double*>(lds + lay.f[i]);
    }
    int* const hc = reinterpret_cast<int*>(lds + lay.hc);
    int* const tw = reinterpret_cast<int*>(lds + lay.tw);
    unsigned short* const pqi = reinterpret_cast<unsigned short*>(lds + lay.pqi);
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
    __syncthreads();

    for (int kk = it.k0; kk < it.k1; ++kk) {
        FjFlKet q;
        if (!fj_fl_ket(a, it, kk, sa, sb, Gb, q)) continue;   // uniform: every lane reads the same ket
        const int k = q.k, kc0 = q.kc0, kcd = q.kcd;
        const double CDx = ld ? a.ket.sep[3 * k] : 0.0, CDy = ld ? a.ket.sep[3 * k + 1] : 0.0,
                     CDz = ld ? a.ket.sep[3 * k + 2] : 0.0;
        fj_fl_load(blk, S.D, S.n_ao, oa, ob, q.oc, q.od, na, nb, nc, nd, q.s_j, q.s_k, lane, nthr);
        const int npq = kab * kcd;   // the ket's primitive quartets, bra-major
        for (int i0 = 0; i0 < npq; i0 += B) {
            const int nbe = min(B, npq - i0);
            // the batch's axis E tables, lanes over (b, side and axis)
            for (int j = lane; j < 6 * nbe; j += nthr) {
                const int b = j / 6, x = j - 6 * b, i = i0 + b;
                if (x < 3) {
                    const double* __restrict__ rb = a.bra.prim + 8 * (size_t)(kb0 + i / kcd);
                    const double sep = x == 0 ? ABx : (x == 1 ? ABy : ABz);
                    fj_kl_eaxis_from(eab + (b * 3 + x) * axa, la, lb, rb[6], rb[3 + x], rb[3 + x] + sep, 1.0);
                } else {
                    const double* __restrict__ rk = a.ket.prim + 8 * (size_t)(kc0 + i % kcd);
                    const double sep = x == 3 ? CDx : (x == 4 ? CDy : CDz);
                    fj_kl_eaxis_from(ecd + (b * 3 + x - 3) * axc, lc, ld, rk[6], rk[x], rk[x] + sep, 1.0);
                }
            }
            // the seeds [0]^(m) and the scalars, lanes over b
            for (int b = lane; b < nbe; b += nthr) {
                const int i = i0 + b;
                const double* __restrict__ rb = a.bra.prim + 8 * (size_t)(kb0 + i / kcd);
                const double* __restrict__ rk = a.ket.prim + 8 * (size_t)(kc0 + i % kcd);
                const double p = rb[6], qe = rk[6];
                const double PQx = rb[0] - rk[0], PQy = rb[1] - rk[1], PQz = rb[2] - rk[2];
                const double rho = p * qe / (p + qe);
                double* const w = wf + b * (L + 1);
                fj_kl_seed(w, rho, rho * (PQx * PQx + PQy * PQy + PQz * PQz), L);
                ((L & 1) ? Rr1 : Rr0)[b * nh] = w[L];
                double* const c = scl + b * FJ_FB_SC;
                c[0] = PQx; c[1] = PQy; c[2] = PQz;
                c[3] = rb[7] * rk[7];
            }
            __syncthreads();
            // the towers, levels m = L - 1 .. 0, lanes over (b, t); [t]^(0) ends in R0
            for (int m = L - 1; m >= 0; --m) {
                double* __restrict__ cur = (m & 1) ? Rr1 : Rr0;
                const double* __restrict__ prv = (m & 1) ? Rr0 : Rr1;
                const int n_m = fj_herm_base(L - m + 1);
                for (int j = lane; j < nbe * n_m; j += nthr) {
                    const int b = j / n_m, t = j - b * n_m;
                    double v;
                    if (t == 0) {
                        v = wf[b * (L + 1) + m];
                    } else {
                        const int w = tw[t];
                        const int d = (w >> 20) & 3, c1 = w >> 22;
                        const double* __restrict__ pv = prv + b * nh;
                        v = scl[b * FJ_FB_SC + d] * pv[w & 1023];
                        if (c1 > 0) v += (double)c1 * pv[(w >> 10) & 1023];
                    }
                    cur[b * nh + t] = v;
                }
                __syncthreads();
            }
            if (nab >= FJ_FB_MFMA_ROWS) {   // uniform: the class decides
                // W[b][ab][q] = U U sum_p E^{ab}_p [p + q] on the matrix cores:
                // 16 x 16 tiles of (ab, q) per b, p in steps of 4, wavefront w
                // taking tiles w + 2 nw i and w + 2 nw i + nw (two chains)
                {
                    const int tqw = (nq + 15) >> 4, tpb = ((nab + 15) >> 4) * tqw, nt = nbe * tpb;
                    for (int t0 = wave; t0 < nt; t0 += 2 * nw) {   // uniform over the wavefront
                        const int t1 = t0 + nw < nt ? t0 + nw : t0;
                        const bool two = t1 != t0;
                        int bq[2], rt[2], qb[2], qc[2];
                        bool rok[2], qok[2];
                        const double* ex[2];
                        const double* ey[2];
                        const double* ez[2];
                        const double* Rb[2];
    #pragma unroll
                        for (int u = 0; u < 2; ++u) {
                            const int t = u ? t1 : t0, b = t / tpb, tt = t - b * tpb, r_t = tt / tqw;
                            const int ra = (r_t << 4) + (lw & 15);
                            bq[u] = b;
                            rt[u] = r_t;
                            qb[u] = ((tt - r_t * tqw) << 4) + (lw & 15);
                            rok[u] = (u == 0 || two) && ra < nab;
                            qok[u] = (u == 0 || two) && qb[u] < nq;
                            qc[u] = min(qb[u], nq - 1);
                            const int ab = min(ra, nab - 1);
                            const int ca = hc[ha + ab / nb], cb = hc[hb + ab % nb];
                            const double* __restrict__ E = eab + b * 3 * axa;
                            ex[u] = E + ((ca & 15) * (lb + 1) + (cb & 15)) * ta;
                            ey[u] = E + axa + (((ca >> 4) & 15) * (lb + 1) + ((cb >> 4) & 15)) * ta;
                            ez[u] = E + 2 * axa + ((ca >> 8) * (lb + 1) + (cb >> 8)) * ta;
                            Rb[u] = Rr0 + b * nh;
                        }
                        fj_v4d c0 = {0.0, 0.0, 0.0, 0.0}, c1 = {0.0, 0.0, 0.0, 0.0};
                        for (int p0 = 0; p0 < np; p0 += 4) {
                            const int p = p0 + (lw >> 4), pc = min(p, np - 1);
                            const bool pok = p < np;
                            const int cp = hc[pc];
                            const int px = cp & 15, py = (cp >> 4) & 15, pz = cp >> 8;
                            const unsigned short* __restrict__ pr = pqi + pc * nq;
                            const double a0 = rok[0] && pok ? ex[0][px] * ey[0][py] * ez[0][pz] : 0.0;
                            const double b0 = qok[0] && pok ? Rb[0][pr[qc[0]]] : 0.0;
                            const double a1 = rok[1] && pok ? ex[1][px] * ey[1][py] * ez[1][pz] : 0.0;
                            const double b1 = qok[1] && pok ? Rb[1][pr[qc[1]]] : 0.0;
                            c0 = fj_mfma_f64_16x16x4(a0, b0, c0);
                            c1 = fj_mfma_f64_16x16x4(a1, b1, c1);
                        }
    #pragma unroll
                        for (int u = 0; u < 2; ++u) {
                            if (u && !two) continue;
                            const double uu = scl[bq[u] * FJ_FB_SC + 3];
                            double* __restrict__ Wb = W + bq[u] * nW;
    #pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const int r = (rt[u] << 4) + (lw >> 4) + 4 * v;
                                if (r < nab && qb[u] < nq) Wb[r * nq + qb[u]] = uu * (u ? c1[v] : c0[v]);
                            }
                        }
                    }
                }
                __syncthreads();
                // x(ab, cd) = sum_b sum_q W[b][ab][q] (-1)^|q| E^{cd}_q(b) on the
                // matrix cores: 16 x 16 tiles of (ab, cd), the batch and q one K
                // dimension, two chains per wavefront; each tile digested once
                {
                    const int tcw = (ncd + 15) >> 4, nt = ((nab + 15) >> 4) * tcw;
                    for (int t0 = wave; t0 < nt; t0 += 2 * nw) {   // uniform over the wavefront
                        const int t1 = t0 + nw < nt ? t0 + nw : t0;
                        const bool two = t1 != t0;
                        int rt[2], cb[2], ox[2], oy[2], oz[2], wr[2];
                        bool rok[2], cok[2];
    #pragma unroll
                        for (int u = 0; u < 2; ++u) {
                            const int t = u ? t1 : t0, r_t = t / tcw;
                            const int ra = (r_t << 4) + (lw & 15);
                            rt[u] = r_t;
                            cb[u] = ((t - r_t * tcw) << 4) + (lw & 15);
                            rok[u] = (u == 0 || two) && ra < nab;
                            cok[u] = (u == 0 || two) && cb[u] < ncd;
                            const int cdi = min(cb[u], ncd - 1);
                            const int cc = hc[hcc + cdi / nd], cdd = hc[hd + cdi % nd];
                            ox[u] = ((cc & 15) * (ld + 1) + (cdd & 15)) * tc;
                            oy[u] = axc + (((cc >> 4) & 15) * (ld + 1) + ((cdd >> 4) & 15)) * tc;
                            oz[u] = 2 * axc + ((cc >> 8) * (ld + 1) + (cdd >> 8)) * tc;
                            wr[u] = min(ra, nab - 1) * nq;
                        }
                        fj_v4d c0 = {0.0, 0.0, 0.0, 0.0}, c1 = {0.0, 0.0, 0.0, 0.0};
                        for (int b = 0; b < nbe; ++b) {
                            const double* __restrict__ Wb = W + b * nW;
                            const double* __restrict__ E = ecd + b * 3 * axc;
                            for (int q0 = 0; q0 < nq; q0 += 4) {
                                const int qq = q0 + (lw >> 4), qc = min(qq, nq - 1);
                                const bool qok = qq < nq;
                                const int cq = hc[qc];
                                const int qx = cq & 15, qy = (cq >> 4) & 15, qz = cq >> 8;
                                const bool neg = (qx + qy + qz) & 1;
                                const double e0 = E[ox[0] + qx] * E[oy[0] + qy] * E[oz[0] + qz];
                                const double e1 = E[ox[1] + qx] * E[oy[1] + qy] * E[oz[1] + qz];
                                const double a0 = rok[0] && qok ? Wb[wr[0] + qc] : 0.0;
                                const double b0 = cok[0] && qok ? (neg ? -e0 : e0) : 0.0;
                                const double a1 = rok[1] && qok ? Wb[wr[1] + qc] : 0.0;
                                const double b1 = cok[1] && qok ? (neg ? -e1 : e1) : 0.0;
                                c0 = fj_mfma_f64_16x16x4(a0, b0, c0);
                                c1 = fj_mfma_f64_16x16x4(a1, b1, c1);
                            }
                        }
    #pragma unroll
                        for (int u = 0; u < 2; ++u) {
                            if (u && !two) continue;
    #pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const int r = (rt[u] << 4) + (lw >> 4) + 4 * v;
                                if (r < nab && cb[u] < ncd) fj_fl_digest(blk, nb, nc, nd, r, cb[u], u ? c1[v] : c0[v]);
                            }
                        }
                    }
                }
            } else {
                // W[b][ab][q] = U U sum_p E^{ab}_p [p + q], lanes over (b, 4 x 4 tile of (ab, q))
                const int tq = (nq + 3) >> 2, tW = ((nab + 3) >> 2) * tq;
                for (int j = lane; j < nbe * tW; j += nthr) {
                    const int b = j / tW, t = j - b * tW;
                    const int m0 = (t / tq) << 2, n0 = (t - (t / tq) * tq) << 2;
                    const double* __restrict__ E = eab + b * 3 * axa;
                    const double* __restrict__ R = Rr0 + b * nh;
                    const double* ex[4];
                    const double* ey[4];
                    const double* ez[4];
                    int qq[4];
    #pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int ab = min(m0 + i, nab - 1);
                        const int ca = hc[ha + ab / nb], cb = hc[hb + ab % nb];
                        ex[i] = E + ((ca & 15) * (lb + 1) + (cb & 15)) * ta;
                        ey[i] = E + axa + (((ca >> 4) & 15) * (lb + 1) + ((cb >> 4) & 15)) * ta;
                        ez[i] = E + 2 * axa + ((ca >> 8) * (lb + 1) + (cb >> 8)) * ta;
                        qq[i] = min(n0 + i, nq - 1);
                    }
                    double c[4][4];
    #pragma unroll
                    for (int i = 0; i < 4; ++i)
    #pragma unroll
                        for (int jj = 0; jj < 4; ++jj) c[i][jj] = 0.0;
                    for (int pi = 0; pi < np; ++pi) {
                        const int cp = hc[pi];
                        const int px = cp & 15, py = (cp >> 4) & 15, pz = cp >> 8;
                        const unsigned short* __restrict__ pr = pqi + pi * nq;
                        double ea[4], rv[4];
    #pragma unroll
                        for (int i = 0; i < 4; ++i) ea[i] = ex[i][px] * ey[i][py] * ez[i][pz];
    #pragma unroll
                        for (int jj = 0; jj < 4; ++jj) rv[jj] = R[pr[qq[jj]]];
    #pragma unroll
                        for (int i = 0; i < 4; ++i)
    #pragma unroll
                            for (int jj = 0; jj < 4; ++jj) c[i][jj] += ea[i] * rv[jj];
                    }
                    const double uu = scl[b * FJ_FB_SC + 3];
                    double* __restrict__ Wb = W + b * nW;
    #pragma unroll
                    for (int i = 0; i < 4; ++i)
    #pragma unroll
                        for (int jj = 0; jj < 4; ++jj)
                            if (m0 + i < nab && n0 + jj < nq) Wb[(m0 + i) * nq + n0 + jj] = uu * c[i][jj];
                }
                __syncthreads();
                // x(ab, cd) = sum_b sum_q W[b][ab][q] (-1)^|q| E^{cd}_q(b), lanes over
                // (slice, 4 x 4 tile of (ab, cd)); slice s sums b = s, s + ns, ... and
                // digests its share once
                {
                    const int tcd = (ncd + 3) >> 2, tX = ((nab + 3) >> 2) * tcd;
                    const int ns = max(1, min(nbe, nthr / tX));
                    for (int j = lane; j < ns * tX; j += nthr) {
                        const int sl = j / tX, t = j - sl * tX;
                        const int m0 = (t / tcd) << 2, n0 = (t - (t / tcd) * tcd) << 2;
                        int ox[4], oy[4], oz[4], wr[4];
    #pragma unroll
                        for (int jj = 0; jj < 4; ++jj) {
                            const int cdi = min(n0 + jj, ncd - 1);
                            const int cc = hc[hcc + cdi / nd], cdd = hc[hd + cdi % nd];
                            ox[jj] = ((cc & 15) * (ld + 1) + (cdd & 15)) * tc;
                            oy[jj] = axc + (((cc >> 4) & 15) * (ld + 1) + ((cdd >> 4) & 15)) * tc;
                            oz[jj] = 2 * axc + ((cc >> 8) * (ld + 1) + (cdd >> 8)) * tc;
                            wr[jj] = min(m0 + jj, nab - 1) * nq;
                        }
                        double c[4][4];
    #pragma unroll
                        for (int i = 0; i < 4; ++i)
    #pragma unroll
                            for (int jj = 0; jj < 4; ++jj) c[i][jj] = 0.0;
                        for (int b = sl; b < nbe; b += ns) {
                            const double* __restrict__ Wb = W + b * nW;
                            const double* __restrict__ E = ecd + b * 3 * axc;
                            for (int qi = 0; qi < nq; ++qi) {
                                const int cq = hc[qi];
                                const int qx = cq & 15, qy = (cq >> 4) & 15, qz = cq >> 8;
                                const bool neg = (qx + qy + qz) & 1;
                                double w[4], e[4];
    #pragma unroll
                                for (int i = 0; i < 4; ++i) w[i] = Wb[wr[i] + qi];
    #pragma unroll
                                for (int jj = 0; jj < 4; ++jj) {
                                    const double v = E[ox[jj] + qx] * E[oy[jj] + qy] * E[oz[jj] + qz];
                                    e[jj] = neg ? -v : v;
                                }
    #pragma unroll
                                for (int i = 0; i < 4; ++i)
    #pragma unroll
                                    for (int jj = 0; jj < 4; ++jj) c[i][jj] += w[i] * e[jj];
                            }
                        }
    #pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            if (m0 + i >= nab) continue;
    #pragma unroll
                            for (int jj = 0; jj < 4; ++jj)
                                if (n0 + jj < ncd) fj_fl_digest(blk, nb, nc, nd, m0 + i, n0 + jj, c[i][jj]);
                        }
                    }
                }
            }
            __syncthreads();   // the batch's tables and W are rewritten next
        }
        // the ket's six blocks into F (the caller closes F + F^T)
        fj_fl_flush(blk, S.F, S.n_ao, oa, ob, q.oc, q.od, na, nb, nc, nd, lane, nthr);
        __syncthreads();   // before the next ket rewrites the blocks
    }
}
