// fj_noalloc.hip -- fj_f_noalloc: the upstream-ABI fused-F launch
// (fj_f_<la><lb><lc><ld>) that allocates nothing, for callers that cannot
// call hipMalloc (code under stream capture, tight serving loops, harnesses
// that forbid it).
//
// fj_f_<cls> keeps the staged classes' contracted boundary, half-transformed
// block and F_cd accumulators in a library-owned workspace that it allocates
// (and grows) on first use. This variant uses a fixed file-scope device
// array instead, `scratch`, FJ_NOALLOC_SCRATCH_BYTES per GPU, held by every
// GPU the library runs on until the process ends. The persistent kernel is
// cut to the workgroups whose slots fit in it (the allocating entry cuts it
// at FJ_LIB_WS_BLOCKS), so any launch size runs; a class whose single
// workgroup of `threads` lanes needs more than the array is refused with
// FJ_ERR_WORKSPACE. No part of the array needs to start at zero: every slot
// is written before it is read, within each work item. Every call on one GPU
// shares that GPU's array, so calls that overlap in time on one GPU race
// (run them on one stream, or order them).
#include "fj_device.h"
#include "gen/f_any_table.h"

#include <mutex>

static __device__ __align__(256) double scratch[FJ_NOALLOC_SCRATCH_BYTES / 8];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per
// device (with that device current). The array's host-side name is only a
// stand-in: kernels get this pointer.
int scratch_ptr(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= MAX_DEVICES) return FJ_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return FJ_ERR_WORKSPACE;
        g_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_ptr[dev];
    return 0;
}

}  // namespace

extern "C" long long fj_f_noalloc_scratch_bytes(void) { return FJ_NOALLOC_SCRATCH_BYTES; }

extern "C" int fj_f_noalloc(hipStream_t s, int la, int lb, int lc, int ld, int threads, int n_bra,
                            int n_ket, int k_ab, int k_cd, int n_sh, int n_ao, int diag, double tau,
                            const double* bra_prim, const double* AB, const double* G_ab,
                            const int* bra_sh, const int* bra_ao, const double* ket_prim,
                            const double* CD, const double* G_cd, const int* ket_sh,
                            const int* ket_ao, const double* Dsh, const double* D, double* F) {
    const fj_f_any_t fn = fj_f_any_for(la, lb, lc, ld);
    if (!fn) return FJ_ERR_CLASS;
    if ((lb > 0 && !AB) || (ld > 0 && !CD)) return FJ_ERR_ARG;
    if (threads <= 0) threads = FJ_DEFAULT_THREADS;
    if (threads % 64 || threads > FJ_MAX_THREADS) return FJ_ERR_THREADS;
    const long long per_lane = fj_fx_workspace_bytes(la, lb, lc, ld, threads, 1) / threads;
    if (per_lane * threads > FJ_NOALLOC_SCRATCH_BYTES) return FJ_ERR_WORKSPACE;
    double* ws = nullptr;
    const int rc = scratch_ptr(&ws);
    if (rc) return rc;
    return fn(s, threads, n_bra, n_ket, k_ab, k_cd, n_sh, n_ao, diag, tau, bra_prim, AB, G_ab,
              bra_sh, bra_ao, ket_prim, CD, G_cd, ket_sh, ket_ao, Dsh, D, F, ws,
              FJ_NOALLOC_SCRATCH_BYTES);
}
