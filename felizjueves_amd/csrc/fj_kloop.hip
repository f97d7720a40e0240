// fj_kloop.hip -- launchers of the loop-form MMD kernels (fj_kloop.h): K
// (the entries fj_k_<la><lc><lb><ld> of the classes with an f shell and the
// generic fj_k_loop) and Schwarz (fj_schwarz_<la><lb> past la + lb = 4 and
// the generic fj_schwarz_loop) and the fused F kernel of the classes with
// an f shell (their fj_f_ / fj_fx_ / fj_fxm_ entries, generated).
#include "fj_device.h"
#include "fj_kloop.h"
#include "fj_floop.h"

// x[0 .. n) = 0, grid-stride (vector stores).
__global__ void fj_kl_zero_k(double* __restrict__ x, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] = 0.0;
}

int fj_kloop_launch(hipStream_t s, int la, int lc, int lb, int ld, int threads, int n_bra, int n_ket,
                    int n_dc, double tau, double maxD, const double* P, const double* zeta_ac,
                    const double* XPA, const double* XPC, const double* s_ac, const double* Ip,
                    const int* dc, const double* Q, const double* eta, const double* XQB,
                    const double* XQD, const double* s_bd, const double* Iq, const int* dd,
                    const double* Dcd, double* K) {
    if (la < 0 || la > FJ_KL_LMAX || lc < 0 || lc > FJ_KL_LMAX || lb < 0 || lb > FJ_KL_LMAX ||
        ld < 0 || ld > FJ_KL_LMAX)
        return FJ_ERR_ARG;
    // threads is checked as for the straight-line kernels; this kernel runs
    // FJ_KL_BLOCK lanes per workgroup whatever it asks
    if (threads <= 0) threads = 256;
    if (threads % 64 || threads > 256) return FJ_ERR_THREADS;
    if (n_bra < 0 || n_ket < 0 || n_dc < 0) return FJ_ERR_SHAPE;
    if (n_bra == 0) return 0;
    // the kets cut into `split` slices, one workgroup per (bra, slice); with
    // more than one the slices add into K, cleared first by a fill kernel
    // (no memset node, so the launch can be captured in a graph)
    const int split = std::max(1, std::min(n_ket, FJ_KL_WORKGROUPS / n_bra));
    if (split > 1) {
        const long long n = (long long)n_bra * ((la + 1) * (la + 2) / 2) * ((lb + 1) * (lb + 2) / 2);
        const int blocks = (int)std::min<long long>((n + 255) / 256, 1024);
        hipLaunchKernelGGL(fj_kl_zero_k, dim3((unsigned)blocks), dim3(256), 0, s, K, n);
    }
    const FjKlLds lay = fj_kloop_lds(la, lc, lb, ld);
    hipLaunchKernelGGL(fj_kloop_k, dim3((unsigned)n_bra, (unsigned)split), dim3(FJ_KL_BLOCK), (size_t)lay.bytes, s,
                       la, lc, lb, ld, n_bra, n_ket, n_dc, tau, maxD, P, zeta_ac, XPA, XPC, s_ac, Ip, dc, Q,
                       eta, XQB, XQD, s_bd, Iq, dd, Dcd, K);
    return (int)hipGetLastError();
}

extern "C" int fj_k_loop(hipStream_t s, int la, int lc, int lb, int ld, int threads, int n_bra,
                         int n_ket, int n_dc, double tau, double maxD, const double* P,
                         const double* zeta_ac, const double* XPA, const double* XPC,
                         const double* s_ac, const double* Ip, const int* dc, const double* Q,
                         const double* eta, const double* XQB, const double* XQD,
                         const double* s_bd, const double* Iq, const int* dd, const double* Dcd,
                         double* K) {
    return fj_kloop_launch(s, la, lc, lb, ld, threads, n_bra, n_ket, n_dc, tau, maxD, P, zeta_ac, XPA,
                           XPC, s_ac, Ip, dc, Q, eta, XQB, XQD, s_bd, Iq, dd, Dcd, K);
}

int fj_schwarz_loop_launch(hipStream_t s, int la, int lb, int threads, int n_pair, int n_rec, int n_ao,
                           const double* zeta_ab, const double* P, const double* XPA,
                           const double* XPB, const double* s_ab, const int* rec0, const int* rec1,
                           const int* off_a, const int* off_b, double* schwarz_bound) {
    if (la < 0 || la > FJ_KL_LMAX || lb < 0 || lb > FJ_KL_LMAX) return FJ_ERR_ARG;
    // threads is checked as for the straight-line kernels; this kernel runs
    // one wavefront per shell pair whatever it asks
    if (threads <= 0) threads = 64;
    if (threads % 64 || threads > 256) return FJ_ERR_THREADS;
    if (n_pair < 0 || n_rec < 0 || n_ao <= 0) return FJ_ERR_SHAPE;
    if (n_pair == 0) return 0;
    hipLaunchKernelGGL(fj_schwarz_loop_k, dim3((unsigned)n_pair), dim3(64), 0, s, la, lb, n_pair, n_ao,
                       zeta_ab, P, XPA, XPB, s_ab, rec0, rec1, off_a, off_b, schwarz_bound);
    return (int)hipGetLastError();
}

extern "C" int fj_schwarz_loop(hipStream_t s, int la, int lb, int threads, int n_pair, int n_rec,
                               int n_ao, const double* zeta_ab, const double* P, const double* XPA,
                               const double* XPB, const double* s_ab, const int* rec0,
                               const int* rec1, const int* off_a, const int* off_b,
                               double* schwarz_bound) {
    return fj_schwarz_loop_launch(s, la, lb, threads, n_pair, n_rec, n_ao, zeta_ab, P, XPA, XPB, s_ab,
                                  rec0, rec1, off_a, off_b, schwarz_bound);
}

int fj_floop_launch(hipStream_t s, int la, int lb, int lc, int ld, int threads, FjFArgs a, int n_blocks) {
    if (la < 0 || la > FJ_KL_LMAX || lb < 0 || lb > FJ_KL_LMAX || lc < 0 || lc > FJ_KL_LMAX ||
        ld < 0 || ld > FJ_KL_LMAX)
        return FJ_ERR_ARG;
    // threads is the item width (kets per work item) the plan was made for;
    // the kernel takes an item's kets one at a time on one workgroup of
    // FJ_FL_BLOCK lanes
    if (threads <= 0) threads = FJ_DEFAULT_THREADS;
    if (threads % 64 || threads > FJ_MAX_THREADS) return FJ_ERR_THREADS;
    if (n_blocks <= 0 || a.ket.n <= 0) return 0;
    // with a work list, each item's kets are cut into up to `threads` slices,
    // one workgroup each, so that a launch of few long items still fills the
    // chip (a workgroup takes its kets one after another)
    const int split = a.work ? std::max(1, std::min(threads, FJ_FL_WORKGROUPS / n_blocks)) : 0;
    const dim3 grid((unsigned)n_blocks,
                    a.work ? (unsigned)split : (unsigned)((a.ket.n + threads - 1) / threads));
    a.n_bx = (int)grid.x;
    a.n_items = (int)(grid.x * grid.y);
    // the batched kernel where B >= 2 primitive quartets fit its LDS budget,
    // else one primitive quartet at a time
    const FjFbLds fb = fj_fbloop_plan(la, lb, lc, ld);
    if (fb.B >= 2) {
        hipLaunchKernelGGL(fj_fbloop_k, grid, dim3(FJ_FL_BLOCK), (size_t)fb.bytes, s, la, lb, lc, ld, threads,
                           fb.B, a);
    } else {
        const FjFlLds lay = fj_floop_lds(la, lb, lc, ld);
        hipLaunchKernelGGL(fj_floop_k, grid, dim3(FJ_FL_BLOCK), (size_t)lay.bytes, s, la, lb, lc, ld, threads, a);
    }
    return (int)hipGetLastError();
}
