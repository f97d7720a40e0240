// fj_dfloop.hip -- the density-fitting integrals on the GPU: the two-center
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
// L = la + lb + lc <= 12: (ff|i) and (i|i)): one wavefront per cell (an AO
// shell pair and an aux shell, or an aux shell pair); per (bra, aux) record
// pair the per-axis E tables and the Hermite tower in LDS (the K kernel's
// helpers), then lanes over the cell's output elements ((a, b), c), each
// accumulated in LDS across the record pairs and stored once, mirrored:
// plain stores, no atomics (every element belongs to one cell).
#include "fj_device.h"
#define FJ_KLOOP_HELPERS_ONLY
#include "fj_kloop.h"

namespace {

constexpr int DF_LMAX_AUX = 6;            // aux shells through i
constexpr int DF_ACC_MAX = 100 * 28;      // (ff|i): the largest cell

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

}  // namespace

__global__ void __launch_bounds__(64) fj_dfloop_k(FjDfArgs g) {
    __shared__ double R[2][FJ_KL_NH];
    __shared__ double eb[3 * FJ_KL_AX], ec[3 * FJ_KL_AX];
    __shared__ double wf[FJ_KL_MMAX + 1];
    __shared__ int tw[FJ_KL_NH];
    __shared__ double acc[DF_ACC_MAX];

    const int cell = (int)blockIdx.x, lane = (int)threadIdx.x;
    if (cell >= g.n_cell) return;   // uniform per workgroup
    const int la = g.la, lb = g.lb, lc = g.lc;
    const int L = la + lb + lc, nh = fj_herm_base(L + 1);
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2, nc = (lc + 1) * (lc + 2) / 2;
    const int n_el = na * nb * nc;
    const int nj = lb + 1, tt = la + lb + 1;   // bra E table strides
    const int tc = lc + 1;                     // aux E table stride (l2 = 0)
    fj_kl_steps(tw, nullptr, nh, lane);
    for (int r = lane; r < n_el; r += 64) acc[r] = 0.0;
    const int i0 = g.b_rec0[cell], i1 = g.b_rec1[cell];
    const int k0 = g.k_rec0[cell], k1 = g.k_rec1[cell];
    __syncthreads();
    for (int i = i0; i < i1; ++i) {
        for (int k = k0; k < k1; ++k) {
            const double zi = g.zeta[i], ak = g.alpha[k];
            if (lane < 3)
                fj_kl_eaxis(eb + lane * FJ_KL_AX, la, lb, zi, g.XPA ? g.XPA[3 * i + lane] : 0.0,
                            g.XPB ? g.XPB[3 * i + lane] : 0.0);
            else if (lane < 6)
                fj_kl_eaxis_from(ec + (lane - 3) * FJ_KL_AX, lc, 0, ak, 0.0, 0.0, 1.0);
            const double PQx = g.P[3 * i] - g.A[3 * k], PQy = g.P[3 * i + 1] - g.A[3 * k + 1],
                         PQz = g.P[3 * i + 2] - g.A[3 * k + 2];
            const double rho = zi * ak / (zi + ak);
            if (lane == 0) {
                fj_kl_seed(wf, rho, rho * (PQx * PQx + PQy * PQy + PQz * PQz), L);
                R[L & 1][0] = wf[L];
            }
            __syncthreads();
            fj_kl_tower(R[0], R[1], wf, tw, L, PQx, PQy, PQz, lane);
            const double sik = g.s_bra[i] * g.s_aux[k];
            for (int r = lane; r < n_el; r += 64) {
                const int ab = r / nc, c = r - ab * nc;
                const int a = ab / nb, b = ab - a * nb;
                const int ca = fj_herm_decode(fj_herm_base(la) + a);
                const int cb = fj_herm_decode(fj_herm_base(lb) + b);
                const int cc = fj_herm_decode(fj_herm_base(lc) + c);
                const int ax = ca & 15, ay = (ca >> 4) & 15, az = ca >> 8;
                const int bx = cb & 15, by = (cb >> 4) & 15, bz = cb >> 8;
                const int cx = cc & 15, cy = (cc >> 4) & 15, cz = cc >> 8;
                const double* __restrict__ ix = eb + (ax * nj + bx) * tt;
                const double* __restrict__ iy = eb + FJ_KL_AX + (ay * nj + by) * tt;
                const double* __restrict__ iz = eb + 2 * FJ_KL_AX + (az * nj + bz) * tt;
                const double* __restrict__ kx = ec + cx * tc;
                const double* __restrict__ ky = ec + FJ_KL_AX + cy * tc;
                const double* __restrict__ kz = ec + 2 * FJ_KL_AX + cz * tc;
                double v = 0.0;
                for (int px = 0; px <= ax + bx; ++px)
                    for (int py = 0; py <= ay + by; ++py)
                        for (int pz = 0; pz <= az + bz; ++pz) {
                            double w = 0.0;
                            // the single-center E^c_tau is zero unless c_d - tau_d is even
                            for (int qx = cx & 1; qx <= cx; qx += 2)
                                for (int qy = cy & 1; qy <= cy; qy += 2)
                                    for (int qz = cz & 1; qz <= cz; qz += 2) {
                                        const double e = kx[qx] * ky[qy] * kz[qz];
                                        const double rv = R[0][fj_herm_index(px + qx, py + qy, pz + qz)];
                                        w += ((qx + qy + qz) & 1 ? -e : e) * rv;
                                    }
                            v += (ix[px] * iy[py] * iz[pz]) * w;
                        }
                acc[r] += sik * v;
            }
            __syncthreads();   // before the next record pair rewrites eb, ec, wf and R
        }
    }
    // the scatter, mirrored; a diagonal cell writes its lower triangle
    const int oa = g.off_a[cell], oc = g.off_c[cell];
    if (g.mode == 0) {
        const int ob = g.off_b[cell];
        const size_t na_ = (size_t)g.n_ao, nx = (size_t)g.n_aux;
        for (int r = lane; r < n_el; r += 64) {
            const int ab = r / nc, c = r - ab * nc;
            const int a = ab / nb, b = ab - a * nb;
            if (oa == ob && a < b) continue;
            const double v = acc[r];
            g.out[((oa + a) * na_ + ob + b) * nx + oc + c] = v;
            g.out[((ob + b) * na_ + oa + a) * nx + oc + c] = v;
        }
    } else {
        const size_t nx = (size_t)g.n_aux;
        for (int r = lane; r < n_el; r += 64) {
            const int a = r / nc, c = r - a * nc;   // nb = 1
            if (oa == oc && a < c) continue;
            const double v = acc[r];
            g.out[(oa + a) * nx + oc + c] = v;
            g.out[(oc + c) * nx + oa + a] = v;
        }
    }
}

static int fj_df_launch(const FjDfArgs& g, hipStream_t s, int threads) {
    if (threads <= 0) threads = 64;
    if (threads % 64 || threads > 256) return FJ_ERR_THREADS;
    if (g.n_cell < 0 || g.n_aux < 0 || g.n_ao < 0) return FJ_ERR_SHAPE;
    if (g.n_cell == 0) return 0;
    hipLaunchKernelGGL(fj_dfloop_k, dim3((unsigned)g.n_cell), dim3(64), 0, s, g);
    return (int)hipGetLastError();
}

// (P|Q) of the aux shell pairs of class (lP, lQ), the upstream metric
// kernel's parameter list (kernel_metric.py:460-480).
extern "C" int fj_metric_loop(hipStream_t s, int lP, int lQ, int threads, int n_pair, int n_rec,
                              int n_aux, const double* A, const double* alpha, const double* s_aux,
                              const int* p_rec0, const int* p_rec1, const int* q_rec0,
                              const int* q_rec1, const int* off_p, const int* off_q, double* metric) {
    if (lP < 0 || lQ < 0 || lP > DF_LMAX_AUX || lQ > DF_LMAX_AUX) return FJ_ERR_CLASS;
    if (n_pair < 0 || n_rec < 0 || n_aux < 0) return FJ_ERR_SHAPE;
    if (n_pair && (!A || !alpha || !s_aux || !p_rec0 || !p_rec1 || !q_rec0 || !q_rec1 || !off_p ||
                   !off_q || !metric))
        return FJ_ERR_ARG;
    FjDfArgs g{lP, 0, lQ, n_pair, alpha, A, nullptr, nullptr, s_aux, A, alpha, s_aux,
               p_rec0, p_rec1, q_rec0, q_rec1, off_p, nullptr, off_q, 1, 0, n_aux, metric};
    return fj_df_launch(g, s, threads);
}

// (mn|P) of the cells (AO shell pair, aux shell) of class (la, lb, lP), the
// upstream three-center kernel's parameter list (kernel_eri3.py:805-838).
extern "C" int fj_eri3_loop(hipStream_t s, int la, int lb, int lP, int threads, int n_cell,
                            int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab,
                            const double* P, const double* XPA, const double* XPB, const double* s_ab,
                            const double* A, const double* alpha, const double* s_aux,
                            const int* ab_rec0, const int* ab_rec1, const int* p_rec0,
                            const int* p_rec1, const int* off_a, const int* off_b, const int* off_p,
                            double* eri3) {
    if (la < 0 || lb < 0 || lP < 0 || la > FJ_KL_LMAX || lb > FJ_KL_LMAX || lP > DF_LMAX_AUX)
        return FJ_ERR_CLASS;
    if (n_cell < 0 || n_bra_rec < 0 || n_aux_rec < 0 || n_ao < 0 || n_aux < 0) return FJ_ERR_SHAPE;
    if (n_cell && (!zeta_ab || !P || !XPA || !XPB || !s_ab || !A || !alpha || !s_aux || !ab_rec0 ||
                   !ab_rec1 || !p_rec0 || !p_rec1 || !off_a || !off_b || !off_p || !eri3))
        return FJ_ERR_ARG;
    FjDfArgs g{la, lb, lP, n_cell, zeta_ab, P, XPA, XPB, s_ab, A, alpha, s_aux,
               ab_rec0, ab_rec1, p_rec0, p_rec1, off_a, off_b, off_p, 0, n_ao, n_aux, eri3};
    return fj_df_launch(g, s, threads);
}
