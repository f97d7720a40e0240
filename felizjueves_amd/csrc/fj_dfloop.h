// fj_dfloop.h -- the density-fitting integrals on the GPU: the two-center
// Coulomb metric (P|Q) of an auxiliary basis (upstream
// eri/density_fit/kernel_metric.py, host metric.py) and the three-center
// (mn|P) (upstream eri/density_fit/kernel_eri3.py, host eri3.py), both by
// McMurchie-Davidson with the aux side's Hermite "pair" a single Gaussian:
//
//   (ab|c) = sum_{i bra records} sum_{k aux records} s_i s_k
//            sum_t E^{ab}_t(i) sum_tau (-1)^|tau| [t + tau]^(0)(P_i - A_k) E^c_tau(k)
//
// with theta^2 = zeta_i alpha_k / (zeta_i + alpha_k). For (mn|P) a bra record
// is a primitive pair (zeta, P, X_PA, X_PB, s = c_m c_n (pi/zeta)^{3/2}, the
// E seed exp(zeta X_PA X_PB) carrying the pair's Gaussian factor); for (P|Q)
// it is an aux primitive (zeta = alpha, P = A, X_PA = 0, lb = 0). An aux
// record is (A, alpha, s = c (pi/alpha)^{3/2}).
//
// One kernel in loop form for every class (the classes as arguments, through
// L = la + lb + lc <= 12: (ff|i) and (i|i)). A workgroup takes cells (an AO
// shell pair and an aux shell, or an aux shell pair) grid-stride, and a
// cell's (bra, aux) record pairs in batches of B: per batch the per-axis E
// tables, the Boys seeds and the Hermite towers of B record pairs in LDS,
// then lanes over the output elements ((a, b), c), each adding the batch's
// record pairs into its accumulator in LDS; one store per element at the
// end, mirrored: plain stores (every element belongs to one cell). L + 2
// barriers per batch, not per record pair. B is the largest that fits the LDS
// budget, at most the cell's record pairs (fj_df_plan).
//
// The includer provides fj_device.h (the device one, or the CPU emulation);
// the launchers and the per-class entries are in fj_dfloop.hip.
#pragma once
#define FJ_KLOOP_HELPERS_ONLY
#include "fj_kloop.h"

namespace {

constexpr int DF_LMAX_AUX = 6;            // aux shells through i
constexpr int DF_ACC_MAX = 100 * 28;      // (ff|i): the largest cell
constexpr int DF_LDS_BUDGET = 48 * 1024;  // bytes per workgroup
constexpr int DF_BMAX = 32;               // record pairs per batch
constexpr int DF_SC = 5;                  // per record pair: s_i s_k, PQ (3), spare

struct FjDfArgs {
    int la, lb, lc;
    int n_cell;
    // bra records (X_PA / X_PB may be NULL: zero, a single-center bra)
    const double* zeta;
    const double* P;
    const double* XPA;
    const double* XPB;
    const double* s_bra;
    // aux records
    const double* A;
    const double* alpha;
    const double* s_aux;
    // per cell: record ranges and the output offsets
    const int* b_rec0;
    const int* b_rec1;
    const int* k_rec0;
    const int* k_rec1;
    const int* off_a;
    const int* off_b;   // eri3 only
    const int* off_c;
    // output: mode 0 eri3[n_ao][n_ao][n_aux], mode 1 metric[n_aux][n_aux]
    int mode;
    int n_ao, n_aux;
    double* out;
};

// The class's LDS: byte offsets of the arrays in one dynamic block and the
// total, for B record pairs per batch.
struct FjDfLds {
    int B, nh, axb, axc;
    int acc, R0, R1, eb, ec, wf, sc, tw, bytes;
};

__host__ __device__ inline FjDfLds fj_df_layout(int la, int lb, int lc, int B) {
    auto herm = [](int n) { return n * (n + 1) * (n + 2) / 6; };
    const int L = la + lb + lc;
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2, nc = (lc + 1) * (lc + 2) / 2;
    FjDfLds o;
    o.B = B;
    o.nh = herm(L + 1);
    o.axb = (la + 1) * (lb + 1) * (la + lb + 1);
    o.axc = (lc + 1) * (lc + 1);
    int at = 0;
    auto take = [&](int n, int size) { const int r = at; at += ((n * size + 7) / 8) * 8; return r; };
    o.acc = take(na * nb * nc, 8);
    o.R0 = take(B * o.nh, 8);
    o.R1 = take(B * o.nh, 8);
    o.eb = take(B * 3 * o.axb, 8);
    o.ec = take(B * 3 * o.axc, 8);
    o.wf = take(B * (L + 1), 8);
    o.sc = take(B * DF_SC, 8);
    o.tw = take(o.nh, 4);
    o.bytes = at;
    return o;
}

// B for a class whose cells have at most max_rp record pairs.
__host__ __device__ inline FjDfLds fj_df_plan(int la, int lb, int lc, int max_rp) {
    const FjDfLds one = fj_df_layout(la, lb, lc, 1), two = fj_df_layout(la, lb, lc, 2);
    const int per_b = two.bytes - one.bytes, fixed = one.bytes - per_b;
    int B = (DF_LDS_BUDGET - fixed) / per_b;
    if (B > DF_BMAX) B = DF_BMAX;
    if (B > max_rp) B = max_rp;
    if (B < 1) B = 1;
    return fj_df_layout(la, lb, lc, B);
}

}  // namespace

__global__ void __launch_bounds__(256) fj_dfloop_k(FjDfArgs g, FjDfLds o) {
    FJ_DYN_SHARED(fj_dfloop_lds_block);
    char* const lds = reinterpret_cast<char*>(fj_dfloop_lds_block);
    double* const acc = reinterpret_cast<double*>(lds + o.acc);
    double* const R0 = reinterpret_cast<double*>(lds + o.R0);
    double* const R1 = reinterpret_cast<double*>(lds + o.R1);
    double* const eb = reinterpret_cast<double*>(lds + o.eb);
    double* const ec = reinterpret_cast<double*>(lds + o.ec);
    double* const wf = reinterpret_cast<double*>(lds + o.wf);
    double* const sc = reinterpret_cast<double*>(lds + o.sc);
    int* const tw = reinterpret_cast<int*>(lds + o.tw);

    const int tid = (int)threadIdx.x, nthr = (int)blockDim.x;
    const int la = g.la, lb = g.lb, lc = g.lc;
    const int L = la + lb + lc, nh = o.nh, B = o.B;
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2, nc = (lc + 1) * (lc + 2) / 2;
    const int n_el = na * nb * nc;
    const int nj = lb + 1, tt = la + lb + 1;   // bra E table strides
    const int tc = lc + 1;                     // aux E table stride (l2 = 0)
    const int axb = o.axb, axc = o.axc;
    fj_kl_steps_n(tw, nullptr, nh, tid, nthr);

    for (int cell = (int)blockIdx.x; cell < g.n_cell; cell += (int)gridDim.x) {
        const int i0 = g.b_rec0[cell], i1 = g.b_rec1[cell];
        const int k0 = g.k_rec0[cell], k1 = g.k_rec1[cell];
        const int nk = k1 - k0, n_rp = (i1 - i0) * nk;
        __syncthreads();   // the previous cell's scatter has read acc
        for (int r = tid; r < n_el; r += nthr) acc[r] = 0.0;
        for (int base = 0; base < n_rp; base += B) {
            const int nbq = n_rp - base < B ? n_rp - base : B;
            // per record pair b: the bra and aux E tables and the Boys seed
            for (int r = tid; r < 7 * nbq; r += nthr) {
                const int b = r / 7, s = r - 7 * b;
                const int rp = base + b, i = i0 + rp / nk, k = k0 + rp % nk;
                if (s < 3) {
                    fj_kl_eaxis(eb + b * 3 * axb + s * axb, la, lb, g.zeta[i], g.XPA ? g.XPA[3 * i + s] : 0.0,
                                g.XPB ? g.XPB[3 * i + s] : 0.0);
                } else if (s < 6) {
                    fj_kl_eaxis_from(ec + b * 3 * axc + (s - 3) * axc, lc, 0, g.alpha[k], 0.0, 0.0, 1.0);
                } else {
                    const double zi = g.zeta[i], ak = g.alpha[k];
                    const double PQx = g.P[3 * i] - g.A[3 * k], PQy = g.P[3 * i + 1] - g.A[3 * k + 1],
                                 PQz = g.P[3 * i + 2] - g.A[3 * k + 2];
                    const double rho = zi * ak / (zi + ak);
                    double* const w = wf + b * (L + 1);
                    fj_kl_seed(w, rho, rho * (PQx * PQx + PQy * PQy + PQz * PQz), L);
                    ((L & 1) ? R1 : R0)[b * nh] = w[L];
                    double* const s_ = sc + b * DF_SC;
                    s_[0] = g.s_bra[i] * g.s_aux[k];
                    s_[1] = PQx;
                    s_[2] = PQy;
                    s_[3] = PQz;
                }
            }
            __syncthreads();
            // the towers, levels m = L - 1 .. 0: lanes over (b, t)
            for (int m = L - 1; m >= 0; --m) {
                double* __restrict__ cur = (m & 1) ? R1 : R0;
                const double* __restrict__ prv = (m & 1) ? R0 : R1;
                const int n_m = fj_herm_base(L - m + 1);
                for (int r = tid; r < nbq * n_m; r += nthr) {
                    const int b = r / n_m, j = r - b * n_m;
                    double v;
                    if (j == 0) {
                        v = wf[b * (L + 1) + m];
                    } else {
                        const int w = tw[j];
                        const int d = (w >> 20) & 3, c1 = w >> 22;
                        v = sc[b * DF_SC + 1 + d] * prv[b * nh + (w & 1023)];
                        if (c1 > 0) v += (double)c1 * prv[b * nh + ((w >> 10) & 1023)];
                    }
                    cur[b * nh + j] = v;
                }
                __syncthreads();
            }
            // lanes over the elements, each summing the batch's record pairs
            // (no two lanes write one element: plain LDS updates)
            for (int e = tid; e < n_el; e += nthr) {
                const int ab = e / nc, c = e - ab * nc;
                const int a = ab / nb, bb = ab - a * nb;
                const int ca = fj_herm_decode(fj_herm_base(la) + a);
                const int cb = fj_herm_decode(fj_herm_base(lb) + bb);
                const int cc = fj_herm_decode(fj_herm_base(lc) + c);
                const int ax = ca & 15, ay = (ca >> 4) & 15, az = ca >> 8;
                const int bx = cb & 15, by = (cb >> 4) & 15, bz = cb >> 8;
                const int cx = cc & 15, cy = (cc >> 4) & 15, cz = cc >> 8;
                double tot = 0.0;
                for (int b = 0; b < nbq; ++b) {
                    const double* __restrict__ e_ = eb + b * 3 * axb;
                    const double* __restrict__ f_ = ec + b * 3 * axc;
                    const double* __restrict__ ix = e_ + (ax * nj + bx) * tt;
                    const double* __restrict__ iy = e_ + axb + (ay * nj + by) * tt;
                    const double* __restrict__ iz = e_ + 2 * axb + (az * nj + bz) * tt;
                    const double* __restrict__ kx = f_ + cx * tc;
                    const double* __restrict__ ky = f_ + axc + cy * tc;
                    const double* __restrict__ kz = f_ + 2 * axc + cz * tc;
                    const double* __restrict__ Rb = R0 + b * nh;
                    double v = 0.0;
                    for (int px = 0; px <= ax + bx; ++px)
                        for (int py = 0; py <= ay + by; ++py)
                            for (int pz = 0; pz <= az + bz; ++pz) {
                                double w = 0.0;
                                // the single-center E^c_tau is zero unless c_d - tau_d is even
                                for (int qx = cx & 1; qx <= cx; qx += 2)
                                    for (int qy = cy & 1; qy <= cy; qy += 2)
                                        for (int qz = cz & 1; qz <= cz; qz += 2) {
                                            const double ev = kx[qx] * ky[qy] * kz[qz];
                                            const double rv = Rb[fj_herm_index(px + qx, py + qy, pz + qz)];
                                            w += ((qx + qy + qz) & 1 ? -ev : ev) * rv;
                                        }
                                v += (ix[px] * iy[py] * iz[pz]) * w;
                            }
                    tot += sc[b * DF_SC] * v;
                }
                acc[e] += tot;
            }
            __syncthreads();   // before the next batch rewrites the tables
        }
        // the scatter, mirrored; a diagonal cell writes its lower triangle
        const int oa = g.off_a[cell], oc = g.off_c[cell];
        if (g.mode == 0) {
            const int ob = g.off_b[cell];
            const size_t na_ = (size_t)g.n_ao, nx = (size_t)g.n_aux;
            for (int r = tid; r < n_el; r += nthr) {
                const int ab = r / nc, c = r - ab * nc;
                const int a = ab / nb, b = ab - a * nb;
                if (oa == ob && a < b) continue;
                const double v = acc[r];
                g.out[((oa + a) * na_ + ob + b) * nx + oc + c] = v;
                g.out[((ob + b) * na_ + oa + a) * nx + oc + c] = v;
            }
        } else {
            const size_t nx = (size_t)g.n_aux;
            for (int r = tid; r < n_el; r += nthr) {
                const int a = r / nc, c = r - a * nc;   // nb = 1
                if (oa == oc && a < c) continue;
                const double v = acc[r];
                g.out[(oa + a) * nx + oc + c] = v;
                g.out[(oc + c) * nx + oa + a] = v;
            }
        }
    }
}
