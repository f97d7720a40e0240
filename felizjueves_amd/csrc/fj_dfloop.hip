// fj_dfloop.hip -- the launchers of the density-fitting kernel (fj_dfloop.h)
// with the upstream metric and three-center parameter lists; the per-class
// entries fj_metric_<lP><lQ> and fj_eri3_<la><lb><lP> (gen/df_entries.hip)
// call them.
#include "fj_device.h"
#include "fj_dfloop.h"

static constexpr int DF_GRID = 8192;   // workgroups per launch, at most

// max_rp bounds the record pairs of every cell (it sizes the batch B; a
// cell with more takes several batches): DF_BMAX when the caller does not
// know it.
static int fj_df_launch(const FjDfArgs& g, hipStream_t s, int threads, int max_rp) {
    if (threads <= 0) threads = 64;
    if (threads % 64 || threads > 256) return FJ_ERR_THREADS;
    if (g.n_cell < 0 || g.n_aux < 0 || g.n_ao < 0) return FJ_ERR_SHAPE;
    if (g.n_cell == 0) return 0;
    if (max_rp < 1) max_rp = DF_BMAX;
    const FjDfLds o = fj_df_plan(g.la, g.lb, g.lc, max_rp);
    const int grid = g.n_cell < DF_GRID ? g.n_cell : DF_GRID;
    hipLaunchKernelGGL(fj_dfloop_k, dim3((unsigned)grid), dim3(threads), (size_t)o.bytes, s, g, o);
    return (int)hipGetLastError();
}

// (P|Q) of the aux shell pairs of class (lP, lQ), the upstream metric
// kernel's parameter list (kernel_metric.py:460-480); fj_metric_loop_rp
// also takes the largest record-pair count of a pair (its batch size).
extern "C" int fj_metric_loop_rp(hipStream_t s, int lP, int lQ, int threads, int n_pair, int n_rec,
                                 int n_aux, const double* A, const double* alpha, const double* s_aux,
                                 const int* p_rec0, const int* p_rec1, const int* q_rec0,
                                 const int* q_rec1, const int* off_p, const int* off_q, double* metric,
                                 int max_rp) {
    if (lP < 0 || lQ < 0 || lP > DF_LMAX_AUX || lQ > DF_LMAX_AUX) return FJ_ERR_CLASS;
    if (n_pair < 0 || n_rec < 0 || n_aux < 0) return FJ_ERR_SHAPE;
    if (n_pair && (!A || !alpha || !s_aux || !p_rec0 || !p_rec1 || !q_rec0 || !q_rec1 || !off_p ||
                   !off_q || !metric))
        return FJ_ERR_ARG;
    FjDfArgs g{lP, 0, lQ, n_pair, alpha, A, nullptr, nullptr, s_aux, A, alpha, s_aux,
               p_rec0, p_rec1, q_rec0, q_rec1, off_p, nullptr, off_q, 1, 0, n_aux, metric};
    return fj_df_launch(g, s, threads, max_rp);
}

extern "C" int fj_metric_loop(hipStream_t s, int lP, int lQ, int threads, int n_pair, int n_rec,
                              int n_aux, const double* A, const double* alpha, const double* s_aux,
                              const int* p_rec0, const int* p_rec1, const int* q_rec0,
                              const int* q_rec1, const int* off_p, const int* off_q, double* metric) {
    return fj_metric_loop_rp(s, lP, lQ, threads, n_pair, n_rec, n_aux, A, alpha, s_aux, p_rec0, p_rec1,
                             q_rec0, q_rec1, off_p, off_q, metric, 0);
}

// (mn|P) of the cells (AO shell pair, aux shell) of class (la, lb, lP), the
// upstream three-center kernel's parameter list (kernel_eri3.py:805-838);
// fj_eri3_loop_rp also takes the largest record-pair count of a cell.
extern "C" int fj_eri3_loop_rp(hipStream_t s, int la, int lb, int lP, int threads, int n_cell,
                               int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab,
                               const double* P, const double* XPA, const double* XPB, const double* s_ab,
                               const double* A, const double* alpha, const double* s_aux,
                               const int* ab_rec0, const int* ab_rec1, const int* p_rec0,
                               const int* p_rec1, const int* off_a, const int* off_b, const int* off_p,
                               double* eri3, int max_rp) {
    if (la < 0 || lb < 0 || lP < 0 || la > FJ_KL_LMAX || lb > FJ_KL_LMAX || lP > DF_LMAX_AUX)
        return FJ_ERR_CLASS;
    if (n_cell < 0 || n_bra_rec < 0 || n_aux_rec < 0 || n_ao < 0 || n_aux < 0) return FJ_ERR_SHAPE;
    if (n_cell && (!zeta_ab || !P || !XPA || !XPB || !s_ab || !A || !alpha || !s_aux || !ab_rec0 ||
                   !ab_rec1 || !p_rec0 || !p_rec1 || !off_a || !off_b || !off_p || !eri3))
        return FJ_ERR_ARG;
    FjDfArgs g{la, lb, lP, n_cell, zeta_ab, P, XPA, XPB, s_ab, A, alpha, s_aux,
               ab_rec0, ab_rec1, p_rec0, p_rec1, off_a, off_b, off_p, 0, n_ao, n_aux, eri3};
    return fj_df_launch(g, s, threads, max_rp);
}

extern "C" int fj_eri3_loop(hipStream_t s, int la, int lb, int lP, int threads, int n_cell,
                            int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab,
                            const double* P, const double* XPA, const double* XPB, const double* s_ab,
                            const double* A, const double* alpha, const double* s_aux,
                            const int* ab_rec0, const int* ab_rec1, const int* p_rec0,
                            const int* p_rec1, const int* off_a, const int* off_b, const int* off_p,
                            double* eri3) {
    return fj_eri3_loop_rp(s, la, lb, lP, threads, n_cell, n_bra_rec, n_aux_rec, n_ao, n_aux, zeta_ab, P,
                           XPA, XPB, s_ab, A, alpha, s_aux, ab_rec0, ab_rec1, p_rec0, p_rec1, off_a,
                           off_b, off_p, eri3, 0);
}
