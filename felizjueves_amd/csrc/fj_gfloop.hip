// fj_gfloop.hip -- the launcher of the fused two-electron gradient kernel
// (fj_gfloop.h) with the upstream gf_ parameter list; the per-class entries
// fj_gf_<la><lb><lc><ld> (gen/gf_entries.hip) call it.
#include "fj_device.h"
#include "fj_gfloop.h"

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
    const long long n_q = diag ? (long long)n_bra * (n_bra + 1) / 2 : (long long)n_bra * n_ket;
    if (n_q > 0x7fffffffLL) return FJ_ERR_SHAPE;
    const FjGfLds o = fj_gf_lds(la, lb, lc, ld);
    FjGfArgs g{la, lb, lc, ld, n_bra, n_ket, k_ab, k_cd, n_atom, n_ao, diag, bra_prim, AB, bra_sh,
               bra_ao, bra_at, ket_prim, CD, ket_sh, ket_ao, ket_at, D, G};
    hipLaunchKernelGGL(fj_gfloop_k, dim3((unsigned)n_q), dim3(threads), (size_t)o.bytes, s, g, o);
    return (int)hipGetLastError();
}

// The LDS bytes one workgroup of class (la lb | lc ld) takes (0 past d).
extern "C" int fj_gf_lds_bytes(int la, int lb, int lc, int ld) {
    if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > GF_LMAX || lb > GF_LMAX || lc > GF_LMAX ||
        ld > GF_LMAX)
        return 0;
    return fj_gf_lds(la, lb, lc, ld).bytes;
}
