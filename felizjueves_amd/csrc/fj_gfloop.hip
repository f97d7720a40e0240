// fj_gfloop.hip -- the launcher of the fused two-electron gradient kernel
// (fj_gfloop.h) with the upstream gf_ parameter list; the per-class entries
// fj_gf_<la><lb><lc><ld> (gen/gf_entries.hip) call it.
#include "fj_device.h"
#include "fj_gfloop.h"

static constexpr long long GF_GRID = 8192;   // workgroups per launch, at most
static constexpr long long GF_ITEMS = 2048;  // work items a launch aims for

// Primitive quartets per work item: whole quartets when there are
// GF_ITEMS of them, else the quartets split into about GF_ITEMS / n_q
// pieces, each a multiple of the batch B.
static int fj_gf_chunk(long long n_q, int n_pq, int B) {
    if (n_q >= GF_ITEMS || n_pq <= B) return n_pq;
    long long pieces = (GF_ITEMS + n_q - 1) / n_q;
    const long long most = (n_pq + B - 1) / B;
    if (pieces > most) pieces = most;
    int chunk = (int)((n_pq + pieces - 1) / pieces);
    chunk = (chunk + B - 1) / B * B;
    return chunk < n_pq ? chunk : n_pq;
}

// One launch of class (la lb | lc ld) over packed sides (the upstream gf_
// kernel's parameter list, kernel_gf.py:501-512; bra_sh / ket_sh may be NULL
// when la != lb / lc != ld, as the upstream omits them there). G [n_atom][3]
// is accumulated, zeroed by the caller before the first launch.
extern "C" int fj_gf_loop(hipStream_t s, int la, int lb, int lc, int ld, int threads, int n_bra,
                          int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag,
                          const double* bra_prim, const double* AB, const int* bra_sh,
                          const int* bra_ao, const int* bra_at, const double* ket_prim,
                          const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at,
                          const double* D, double* G) {
    if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > GF_LMAX || lb > GF_LMAX || lc > GF_LMAX ||
        ld > GF_LMAX)
        return FJ_ERR_CLASS;
    if (threads <= 0) threads = 64;
    if (threads % 64 || threads > 64 * GF_MAXW) return FJ_ERR_THREADS;
    if (n_bra < 0 || n_ket < 0 || n_atom < 0 || n_ao < 0 || (diag != 0 && diag != 1)) return FJ_ERR_SHAPE;
    if (diag && (n_bra != n_ket || la != lc || lb != ld)) return FJ_ERR_SHAPE;
    if (n_bra == 0 || n_ket == 0) return 0;
    if (k_ab < 1 || k_cd < 1) return FJ_ERR_DEGREE;
    if (!bra_prim || !AB || !bra_ao || !bra_at || !ket_prim || !CD || !ket_ao || !ket_at || !D || !G ||
        (la == lb && !bra_sh) || (lc == ld && !ket_sh))
        return FJ_ERR_ARG;
    if ((long long)k_ab * k_cd > 0x7fffffffLL) return FJ_ERR_DEGREE;
    const long long n_q = diag ? (long long)n_bra * (n_bra + 1) / 2 : (long long)n_bra * n_ket;
    const int n_pq = k_ab * k_cd;
    const FjGfLds o = fj_gf_plan(la, lb, lc, ld, n_pq, threads);
    // a launch of few quartets splits each quartet's primitive quartets over
    // several workgroups, so that about GF_ITEMS work items fill the chip
    const int chunk = fj_gf_chunk(n_q, n_pq, o.B);
    FjGfArgs g{la, lb, lc, ld, n_bra, n_ket, k_ab, k_cd, n_atom, n_ao, diag, chunk, bra_prim, AB, bra_sh,
               bra_ao, bra_at, ket_prim, CD, ket_sh, ket_ao, ket_at, D, G};
    const long long n_items = n_q * ((n_pq + chunk - 1) / chunk);
    // grid-stride over the items: enough workgroups to fill the chip
    const long long grid = n_items < GF_GRID ? n_items : GF_GRID;
    hipLaunchKernelGGL(fj_gfloop_k, dim3((unsigned)grid), dim3(threads), (size_t)o.bytes, s, g, o);
    return (int)hipGetLastError();
}

// The LDS bytes one workgroup of class (la lb | lc ld) takes for quartets of
// n_pq primitive quartets at `threads` lanes (0 past d).
extern "C" int fj_gf_lds_bytes(int la, int lb, int lc, int ld, int n_pq, int threads) {
    if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > GF_LMAX || lb > GF_LMAX || lc > GF_LMAX ||
        ld > GF_LMAX || n_pq < 1 || threads < 64)
        return 0;
    return fj_gf_plan(la, lb, lc, ld, n_pq, threads).bytes;
}

// The batch B of primitive quartets a workgroup of the class takes at once
// for quartets of n_pq primitive quartets at `threads` lanes (0 past d).
extern "C" int fj_gf_batch(int la, int lb, int lc, int ld, int n_pq, int threads) {
    if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > GF_LMAX || lb > GF_LMAX || lc > GF_LMAX ||
        ld > GF_LMAX || n_pq < 1 || threads < 64)
        return 0;
    return fj_gf_plan(la, lb, lc, ld, n_pq, threads).B;
}

extern "C" int fj_gf_jobs(hipStream_t s, int la, int lb, int lc, int ld, int threads, int n_jobs,
                          const FjGfJob* jobs, long long n_items, int max_pq, int n_atom, int n_ao,
                          const double* prim, const double* sep, const int* sh, const int* ao,
                          const int* at, const double* D, double* G) {
    if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > GF_LMAX || lb > GF_LMAX || lc > GF_LMAX ||
        ld > GF_LMAX)
        return FJ_ERR_CLASS;
    if (threads <= 0) threads = 64;
    if (threads % 64 || threads > 64 * GF_MAXW) return FJ_ERR_THREADS;
    if (n_jobs < 0 || n_items < 0 || n_atom < 0 || n_ao < 0) return FJ_ERR_SHAPE;
    if (n_jobs == 0 || n_items == 0) return 0;
    if (max_pq < 1) return FJ_ERR_DEGREE;
    if (!jobs || !prim || !sep || !sh || !ao || !at || !D || !G) return FJ_ERR_ARG;
    const FjGfLds o = fj_gf_plan(la, lb, lc, ld, max_pq, threads);
    FjGfArgs g{la, lb, lc, ld, 0, 0, 0, 0, n_atom, n_ao, 0, 1, prim, sep, sh, ao, at,
               prim, sep, sh, ao, at, D, G};
    const long long grid = n_items < GF_GRID ? n_items : GF_GRID;
    hipLaunchKernelGGL(fj_gfjobs_k, dim3((unsigned)grid), dim3(threads), (size_t)o.bytes, s, g, o, jobs,
                       n_jobs, n_items);
    return (int)hipGetLastError();
}
