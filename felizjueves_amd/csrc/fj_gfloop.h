// fj_gfloop.h -- the fused two-electron nuclear gradient on the GPU:
// dE_2e/dR for E_2e = 1/2 tr(D G), G = J - K/2 (upstream
// gradient/kernel_gf.py:1-108, the fused F kernel differentiated), over the
// same stream of unique shell quartets as the fused F build, every class
// through (dd|dd) (the upstream emits gf_ through l = 2,
// emit/__main__.py:842).
//
// Per unique quartet (a b | c d) with the derived weights (kernel_gf.py:25-40)
//
//   w_abcd = deg (4 D_ab D_cd - D_ac D_bd - D_ad D_bc)
//
// and per centre X and axis k, from the derivative of a Cartesian Gaussian
// (kernel_gf.py:44-50), d/dA_k (ab|cd) = (a+1_k b|cd)^{2 alpha} - a_k
// (a-1_k b|cd): all four centres differentiated, no translational
// invariance used (so sum_A G[A] = 0 stays a check).
//
// The integrals are McMurchie-Davidson's, per primitive quartet
//   (ab|cd) = U_ab U_cd sum_t sum_u E^{ab}_t (-1)^|u| E^{cd}_u [t + u]
// so a centre's derivative replaces that side's E^{ab}_t by the E of the
// differentiated product, per axis 2 alpha E^{a+1,b}_t - a E^{a-1,b}_t (the
// two other axes unchanged). The weights are contracted into Hermite space
// first, once per primitive quartet:
//
//   Wk[ab][u] = (-1)^|u| sum_cd w_abcd E^{cd}_u,   |u| <= lc + ld
//   Wb[cd][t] =          sum_ab w_abcd E^{ab}_t,   |t| <= la + lb
//   gA_k, gB_k += U_ab U_cd sum_ab sum_{|t| <= la+lb+1} dE^{ab}_t sum_u [t+u] Wk[ab][u]
//   gC_k, gD_k += U_ab U_cd sum_cd sum_{|u| <= lc+ld+1} (-1)^|u| dE^{cd}_u sum_t [t+u] Wb[cd][t]
//
// One workgroup per unique quartet (threads = 64 .. 256 cooperating through
// LDS): the per-axis E tables to l + 1 on every centre, the Hermite tower to
// L + 1, then lanes over (ab, t) and (cd, u). The twelve sums are reduced over
// the workgroup and added into G[atom][k] with twelve atomics per quartet.
// No screen (kernel_gf.py:82-85: the F screen bounds an integral, not its
// derivative): every unique quartet of a launch is visited.
//
// The includer provides fj_device.h (the device one, or the CPU emulation);
// the launcher and the per-class entries are in fj_gfloop.hip.
#pragma once
#define FJ_KLOOP_HELPERS_ONLY
#include "fj_kloop.h"

namespace {

constexpr int GF_LMAX = 2;    // through (dd|dd), the tower to L + 1 <= 9
constexpr int GF_REC = 10;    // P(3), X = P - A (3), zeta, U, e1, e2
constexpr int GF_MAXW = 4;    // waves per workgroup (threads <= 256)

struct FjGfArgs {
    int la, lb, lc, ld;
    int n_bra, n_ket, k_ab, k_cd, n_atom, n_ao, diag;
    const double* bra_prim;
    const double* AB;
    const int* bra_sh;
    const int* bra_ao;
    const int* bra_at;
    const double* ket_prim;
    const double* CD;
    const int* ket_sh;
    const int* ket_ao;
    const int* ket_at;
    const double* D;
    double* G;
};

// Byte offsets of the class's LDS arrays in one dynamic block (doubles, then
// ints, then the u16 index table) and the total.
struct FjGfLds {
    int R0, R1, eb, ek, wf, w, wk, wb, red, hc, tw, pab, pcd, add, bytes;
    int axb, axk, nR, nu, nt, nu1, nt1;
};

__host__ __device__ inline int fj_gf_herm(int n) { return n * (n + 1) * (n + 2) / 6; }

__host__ __device__ inline FjGfLds fj_gf_lds(int la, int lb, int lc, int ld) {
    const int L = la + lb + lc + ld, lab = la + lb, lcd = lc + ld;
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
    const int nc = (lc + 1) * (lc + 2) / 2, nd = (ld + 1) * (ld + 2) / 2;
    FjGfLds o;
    o.nR = fj_gf_herm(L + 2);      // |t| <= L + 1
    o.nu = fj_gf_herm(lcd + 1);    // |u| <= lc + ld
    o.nt = fj_gf_herm(lab + 1);
    o.nu1 = fj_gf_herm(lcd + 2);   // |u| <= lc + ld + 1
    o.nt1 = fj_gf_herm(lab + 2);
    o.axb = (la + 2) * (lb + 2) * (lab + 3);
    o.axk = (lc + 2) * (ld + 2) * (lcd + 3);
    int at = 0;
    auto take = [&](int n, int size) { const int r = at; at += ((n * size + 7) / 8) * 8; return r; };
    o.R0 = take(o.nR, 8);
    o.R1 = take(o.nR, 8);
    o.eb = take(3 * o.axb, 8);
    o.ek = take(3 * o.axk, 8);
    o.wf = take(L + 2, 8);
    o.w = take(na * nb * nc * nd, 8);
    o.wk = take(na * nb * o.nu, 8);
    o.wb = take(nc * nd * o.nt, 8);
    o.red = take(12 * GF_MAXW, 8);
    o.hc = take(o.nR, 4);
    o.tw = take(o.nR, 4);
    o.pab = take(na * nb, 4);
    o.pcd = take(nc * nd, 4);
    o.add = take(o.nt1 * o.nu1, 2);
    o.bytes = at;
    return o;
}

// The Hermite tower (fj_kl_tower's levels, fj_kloop.h) with the
// workgroup's lanes: threads tid, nthr of the block, a barrier per level.
__device__ __forceinline__ void fj_gf_tower(double* R0, double* R1, const double* wf, const int* tw,
                                            int L, double PQx, double PQy, double PQz, int tid,
                                            int nthr) {
    for (int m = L - 1; m >= 0; --m) {
        double* __restrict__ cur = (m & 1) ? R1 : R0;
        const double* __restrict__ prv = (m & 1) ? R0 : R1;
        const int n_m = fj_herm_base(L - m + 1);
        for (int j = tid; j < n_m; j += nthr) {
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

// Components of the i-th Cartesian of an l shell, packed x | y << 4 | z << 8.
__device__ __forceinline__ int fj_gf_cart(int l, int i) { return fj_herm_decode(fj_herm_base(l) + i); }

// The differentiated product's E along one axis at Hermite order t: e0 (the
// product itself), ea (its first centre's derivative, 2 x1 E^{i+1,j} - i
// E^{i-1,j}) and eb (the second's, 2 x2 E^{i,j+1} - j E^{i,j-1}); table e of
// stride (nj, nt) built to (i + 1, j + 1).
__device__ __forceinline__ void fj_gf_axis(const double* __restrict__ e, int nj, int nt, int i,
                                           int j, int t, double x1, double x2, double& e0,
                                           double& ea, double& eb) {
    e0 = e[(i * nj + j) * nt + t];
    ea = x1 * e[((i + 1) * nj + j) * nt + t];
    if (i) ea -= (double)i * e[((i - 1) * nj + j) * nt + t];
    eb = x2 * e[(i * nj + j + 1) * nt + t];
    if (j) eb -= (double)j * e[(i * nj + j - 1) * nt + t];
}

}  // namespace

__global__ void __launch_bounds__(256) fj_gfloop_k(FjGfArgs g, FjGfLds o) {
    FJ_DYN_SHARED(fj_gfloop_lds_block);
    char* const lds = reinterpret_cast<char*>(fj_gfloop_lds_block);
    double* const R0 = reinterpret_cast<double*>(lds + o.R0);
    double* const R1 = reinterpret_cast<double*>(lds + o.R1);
    double* const eb = reinterpret_cast<double*>(lds + o.eb);
    double* const ek = reinterpret_cast<double*>(lds + o.ek);
    double* const wf = reinterpret_cast<double*>(lds + o.wf);
    double* const W = reinterpret_cast<double*>(lds + o.w);
    double* const Wk = reinterpret_cast<double*>(lds + o.wk);
    double* const Wb = reinterpret_cast<double*>(lds + o.wb);
    double* const red = reinterpret_cast<double*>(lds + o.red);
    int* const hc = reinterpret_cast<int*>(lds + o.hc);
    int* const tw = reinterpret_cast<int*>(lds + o.tw);
    int* const pab = reinterpret_cast<int*>(lds + o.pab);
    int* const pcd = reinterpret_cast<int*>(lds + o.pcd);
    unsigned short* const add = reinterpret_cast<unsigned short*>(lds + o.add);

    const int tid = (int)threadIdx.x, nthr = (int)blockDim.x;
    const int la = g.la, lb = g.lb, lc = g.lc, ld = g.ld;
    const int L = la + lb + lc + ld, Lt = L + 1;
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
    const int nc = (lc + 1) * (lc + 2) / 2, nd = (ld + 1) * (ld + 2) / 2;
    const int nab = na * nb, ncd = nc * nd;
    const int nu = o.nu, nt = o.nt, nu1 = o.nu1, nt1 = o.nt1;
    const int njb = lb + 2, ntb = la + lb + 3, njk = ld + 2, ntk = lc + ld + 3;

    // the quartet: (bra, ket), ket >= bra on the diagonal launch (its upper
    // triangle, row-major), every pair otherwise
    const long long q = (long long)blockIdx.x;
    int bra, ket;
    if (g.diag) {
        const long long n = g.n_bra;
        const double b2 = (double)(2 * n + 1);
        long long b = (long long)((b2 - sqrt(b2 * b2 - 8.0 * (double)q)) * 0.5);
        if (b < 0) b = 0;
        if (b > n - 1) b = n - 1;
        while (b > 0 && b * n - b * (b - 1) / 2 > q) --b;
        while (b + 1 < n && (b + 1) * n - (b + 1) * b / 2 <= q) ++b;
        bra = (int)b;
        ket = (int)(b + (q - (b * n - b * (b - 1) / 2)));
    } else {
        bra = (int)(q / g.n_ket);
        ket = (int)(q - (long long)bra * g.n_ket);
    }

    const int oa = g.bra_ao[2 * bra], ob = g.bra_ao[2 * bra + 1];
    const int oc = g.ket_ao[2 * ket], od = g.ket_ao[2 * ket + 1];
    // the degeneracy 1/|Stab| (kernel_gf.py:591-606): 1/2 for a = b, for
    // c = d, and for the pair met with itself on the diagonal launch
    double deg = 1.0;
    if (la == lb && g.bra_sh[2 * bra] == g.bra_sh[2 * bra + 1]) deg *= 0.5;
    if (lc == ld && g.ket_sh[2 * ket] == g.ket_sh[2 * ket + 1]) deg *= 0.5;
    if (g.diag && ket == bra) deg *= 0.5;

    // per quartet: the weights, the tower's step table, the components of
    // every Hermite index, the Cartesian pairs and the index table [t + u]
    {
        const size_t n = (size_t)g.n_ao;
        const double* __restrict__ D = g.D;
        for (int r = tid; r < nab * ncd; r += nthr) {
            const int ab = r / ncd, cd = r - ab * ncd;
            const int a = ab / nb, b = ab - a * nb, c = cd / nd, d = cd - c * nd;
            const double w = 4.0 * D[(oa + a) * n + ob + b] * D[(oc + c) * n + od + d] -
                             D[(oa + a) * n + oc + c] * D[(ob + b) * n + od + d] -
                             D[(oa + a) * n + od + d] * D[(ob + b) * n + oc + c];
            W[r] = deg * w;
        }
        for (int j = tid; j < o.nR; j += nthr) {
            const int c = fj_herm_decode(j);
            const int x = c & 15, y = (c >> 4) & 15, z = c >> 8;
            hc[j] = c;
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
        for (int r = tid; r < nab; r += nthr) {
            const int a = r / nb, b = r - a * nb;
            pab[r] = fj_gf_cart(la, a) | (fj_gf_cart(lb, b) << 12);
        }
        for (int r = tid; r < ncd; r += nthr) {
            const int c = r / nd, d = r - c * nd;
            pcd[r] = fj_gf_cart(lc, c) | (fj_gf_cart(ld, d) << 12);
        }
        for (int r = tid; r < nt1 * nu1; r += nthr) {
            const int t = r / nu1, u = r - t * nu1;
            const int ct = fj_herm_decode(t), cu = fj_herm_decode(u);
            add[r] = (unsigned short)fj_herm_index((ct & 15) + (cu & 15), ((ct >> 4) & 15) + ((cu >> 4) & 15),
                                                   (ct >> 8) + (cu >> 8));
        }
    }

    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
    double gC[3] = {0.0, 0.0, 0.0}, gD[3] = {0.0, 0.0, 0.0};
    const double* __restrict__ AB = g.AB + 3 * bra;
    const double* __restrict__ CD = g.CD + 3 * ket;
    __syncthreads();

    for (int i = 0; i < g.k_ab; ++i) {
        const double* __restrict__ rb = g.bra_prim + ((size_t)bra * g.k_ab + i) * GF_REC;
        for (int k = 0; k < g.k_cd; ++k) {
            const double* __restrict__ rk = g.ket_prim + ((size_t)ket * g.k_cd + k) * GF_REC;
            const double p = rb[6], qz = rk[6];
            const double PQx = rb[0] - rk[0], PQy = rb[1] - rk[1], PQz = rb[2] - rk[2];
            if (tid < 3)
                fj_kl_eaxis_from(eb + tid * o.axb, la + 1, lb + 1, p, rb[3 + tid], rb[3 + tid] + AB[tid], 1.0);
            else if (tid < 6)
                fj_kl_eaxis_from(ek + (tid - 3) * o.axk, lc + 1, ld + 1, qz, rk[tid], rk[tid] + CD[tid - 3],
                                 1.0);
            else if (tid == 6) {
                const double rho = p * qz / (p + qz);
                fj_kl_seed(wf, rho, rho * (PQx * PQx + PQy * PQy + PQz * PQz), Lt);
                ((Lt & 1) ? R1 : R0)[0] = wf[Lt];
            }
            __syncthreads();
            fj_gf_tower(R0, R1, wf, tw, Lt, PQx, PQy, PQz, tid, nthr);

            // the weights in Hermite space
            for (int r = tid; r < nab * nu; r += nthr) {
                const int ab = r / nu, u = r - ab * nu;
                const int cu = hc[u];
                const int ux = cu & 15, uy = (cu >> 4) & 15, uz = cu >> 8;
                double s = 0.0;
                for (int cd = 0; cd < ncd; ++cd) {
                    const int pc = pcd[cd];
                    const int cx = pc & 15, cy = (pc >> 4) & 15, cz = (pc >> 8) & 15;
                    const int dx = (pc >> 12) & 15, dy = (pc >> 16) & 15, dz = pc >> 20;
                    const double e = ek[(cx * njk + dx) * ntk + ux] * ek[o.axk + (cy * njk + dy) * ntk + uy] *
                                     ek[2 * o.axk + (cz * njk + dz) * ntk + uz];
                    s += W[ab * ncd + cd] * e;
                }
                Wk[r] = ((ux + uy + uz) & 1) ? -s : s;
            }
            for (int r = tid; r < ncd * nt; r += nthr) {
                const int cd = r / nt, t = r - cd * nt;
                const int ct = hc[t];
                const int tx = ct & 15, ty = (ct >> 4) & 15, tz = ct >> 8;
                double s = 0.0;
                for (int ab = 0; ab < nab; ++ab) {
                    const int pa = pab[ab];
                    const int ax = pa & 15, ay = (pa >> 4) & 15, az = (pa >> 8) & 15;
                    const int bx = (pa >> 12) & 15, by = (pa >> 16) & 15, bz = pa >> 20;
                    const double e = eb[(ax * njb + bx) * ntb + tx] * eb[o.axb + (ay * njb + by) * ntb + ty] *
                                     eb[2 * o.axb + (az * njb + bz) * ntb + tz];
                    s += W[ab * ncd + cd] * e;
                }
                Wb[r] = s;
            }
            __syncthreads();

            const double pref = rb[7] * rk[7];
            const double a2 = 2.0 * rb[8], b2 = 2.0 * rb[9], c2 = 2.0 * rk[8], d2 = 2.0 * rk[9];
            // the bra's centres: lanes over (ab, t), |t| <= la + lb + 1
            for (int r = tid; r < nab * nt1; r += nthr) {
                const int ab = r / nt1, t = r - ab * nt1;
                const int pa = pab[ab], ct = hc[t];
                const int ax = pa & 15, ay = (pa >> 4) & 15, az = (pa >> 8) & 15;
                const int bx = (pa >> 12) & 15, by = (pa >> 16) & 15, bz = pa >> 20;
                const int tx = ct & 15, ty = (ct >> 4) & 15, tz = ct >> 8;
                // outside the box t_k <= a_k + b_k (+1 on at most one axis) every term is zero
                const int ox = tx - ax - bx, oy = ty - ay - by, oz = tz - az - bz;
                if (ox > 1 || oy > 1 || oz > 1 || (ox == 1) + (oy == 1) + (oz == 1) > 1) continue;
                double y = 0.0;
                const unsigned short* __restrict__ ar = add + t * nu1;
                const double* __restrict__ wr = Wk + ab * nu;
                for (int u = 0; u < nu; ++u) y += R0[ar[u]] * wr[u];
                y *= pref;
                double ex, exa, exb, ey, eya, eyb, ez, eza, ezb;
                fj_gf_axis(eb, njb, ntb, ax, bx, tx, a2, b2, ex, exa, exb);
                fj_gf_axis(eb + o.axb, njb, ntb, ay, by, ty, a2, b2, ey, eya, eyb);
                fj_gf_axis(eb + 2 * o.axb, njb, ntb, az, bz, tz, a2, b2, ez, eza, ezb);
                gA[0] += y * exa * ey * ez;
                gA[1] += y * ex * eya * ez;
                gA[2] += y * ex * ey * eza;
                gB[0] += y * exb * ey * ez;
                gB[1] += y * ex * eyb * ez;
                gB[2] += y * ex * ey * ezb;
            }
            // the ket's centres: lanes over (cd, u), |u| <= lc + ld + 1
            for (int r = tid; r < ncd * nu1; r += nthr) {
                const int cd = r / nu1, u = r - cd * nu1;
                const int pc = pcd[cd], cu = hc[u];
                const int cx = pc & 15, cy = (pc >> 4) & 15, cz = (pc >> 8) & 15;
                const int dx = (pc >> 12) & 15, dy = (pc >> 16) & 15, dz = pc >> 20;
                const int ux = cu & 15, uy = (cu >> 4) & 15, uz = cu >> 8;
                const int ox = ux - cx - dx, oy = uy - cy - dy, oz = uz - cz - dz;
                if (ox > 1 || oy > 1 || oz > 1 || (ox == 1) + (oy == 1) + (oz == 1) > 1) continue;
                double x = 0.0;
                const double* __restrict__ wr = Wb + cd * nt;
                for (int t = 0; t < nt; ++t) x += R0[add[t * nu1 + u]] * wr[t];
                x *= ((ux + uy + uz) & 1) ? -pref : pref;
                double ex, exc, exd, ey, eyc, eyd, ez, ezc, ezd;
                fj_gf_axis(ek, njk, ntk, cx, dx, ux, c2, d2, ex, exc, exd);
                fj_gf_axis(ek + o.axk, njk, ntk, cy, dy, uy, c2, d2, ey, eyc, eyd);
                fj_gf_axis(ek + 2 * o.axk, njk, ntk, cz, dz, uz, c2, d2, ez, ezc, ezd);
                gC[0] += x * exc * ey * ez;
                gC[1] += x * ex * eyc * ez;
                gC[2] += x * ex * ey * ezc;
                gD[0] += x * exd * ey * ez;
                gD[1] += x * ex * eyd * ez;
                gD[2] += x * ex * ey * ezd;
            }
            __syncthreads();   // before the next primitive quartet rewrites the tables
        }
    }

    // the twelve sums over the workgroup, then into G
    double v[12] = {gA[0], gA[1], gA[2], gB[0], gB[1], gB[2], gC[0], gC[1], gC[2], gD[0], gD[1], gD[2]};
    const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int c = 0; c < 12; ++c) {
        const double s = fj_wave_sum(v[c]);
        if (lane == 0) red[wave * 12 + c] = s;
    }
    __syncthreads();
    if (tid < 12) {
        double s = 0.0;
        for (int w = 0; w < (nthr >> 6); ++w) s += red[w * 12 + tid];
        const int pos = tid / 3, k = tid - pos * 3;
        const int atom = pos == 0 ? g.bra_at[2 * bra] : pos == 1 ? g.bra_at[2 * bra + 1]
                       : pos == 2 ? g.ket_at[2 * ket] : g.ket_at[2 * ket + 1];
        atomicAdd(g.G + 3 * atom + k, s);
    }
}

