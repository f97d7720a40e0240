// fj_runtime.hip -- the library's small non-class entries: version, the
// closing transpose, the density block maxima, the Boys test entry.
#include "fj_device.h"

#include <mutex>

#define FJ_VERSION 1

extern "C" int fj_version(void) { return FJ_VERSION; }

// One grow-only workspace per (device, stream) for the upstream-ABI entries
// of the workspace classes. Launches on one stream run in order and may
// share it; launches on different streams get different ones (the native
// builder passes its own). At most FJ_LIB_WS_SLOTS (device, stream) pairs.
#define FJ_LIB_WS_SLOTS 64
double* fj_library_workspace(hipStream_t s, long long bytes) {
    static std::mutex mu;
    static struct { int dev; hipStream_t stream; double* buf; long long have; } slot[FJ_LIB_WS_SLOTS];
    static int n_slot = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || bytes <= 0) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    int i = 0;
    while (i < n_slot && !(slot[i].dev == dev && slot[i].stream == s)) ++i;
    if (i == n_slot) {
        if (n_slot == FJ_LIB_WS_SLOTS) return nullptr;
        slot[i] = {dev, s, nullptr, 0};
        ++n_slot;
    }
    if (slot[i].have < bytes) {
        if (slot[i].buf) {
            if (hipDeviceSynchronize() != hipSuccess) return nullptr;
            (void)hipFree(slot[i].buf);
            slot[i].buf = nullptr;
            slot[i].have = 0;
        }
        void* p = nullptr;
        if (hipMalloc(&p, (size_t)bytes) != hipSuccess) return nullptr;
        slot[i].buf = static_cast<double*>(p);
        slot[i].have = bytes;
    }
    return slot[i].buf;
}

// The 8-element ERI symmetry group as shell-position maps g: the class g.cls
// has cls[g[j]] in position j (upstream hgp/orbits.py:78-87, the same order,
// so a perm index means the same orientation on both sides).
static const int FJ_GROUP[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                   {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};

int fj_split_dig(int screen, int perm, int n_ao, FjDig* dg) {
    if (!dg || perm < 0 || perm > 7 || (screen != 0 && screen != 1) || n_ao <= 0) return FJ_ERR_ARG;
    int inv[4];
    for (int j = 0; j < 4; ++j) inv[FJ_GROUP[perm][j]] = j;
    // J: K_ab += D_cd (ab|cd) on the ordered quartet's positions (0,1)/(2,3);
    // K: K_ac += D_bd (ab|cd), positions (0,2)/(1,3); each carried to the
    // built quartet's positions by g^-1
    const int o0 = inv[0], o1 = screen ? inv[2] : inv[1];
    const int q0 = screen ? inv[1] : inv[2], q1 = inv[3];
    const int lo = o0 < o1 ? o0 : o1, hi = o0 < o1 ? o1 : o0;
    int blk = -1;
    if (lo == 0 && hi == 1) blk = 0;
    else if (lo == 2 && hi == 3) blk = 1;
    else if (lo == 0 && hi == 2) blk = 2;
    else if (lo == 0 && hi == 3) blk = 3;
    else if (lo == 1 && hi == 2) blk = 4;
    else if (lo == 1 && hi == 3) blk = 5;
    if (blk < 0) return FJ_ERR_ARG;
    for (int i = 0; i < 6; ++i) dg->w[i] = i == blk ? 1.0 : 0.0;
    // the kernels read and write every block in (lower, higher) position order
    dg->fr = o0 < o1 ? n_ao : 1;
    dg->fc = o0 < o1 ? 1 : n_ao;
    dg->dr = q0 < q1 ? n_ao : 1;
    dg->dc = q0 < q1 ? 1 : n_ao;
    dg->q0 = q0;
    dg->q1 = q1;
    return 0;
}

// F <- F + F^T: each thread owns one (i, j), i < j, pair and the diagonal.
__global__ void fj_close_k(int n, double* __restrict__ F) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)n * (n + 1) / 2;
    if (t >= total) return;
    // row-major upper triangle index -> (i, j), i <= j
    const double nn = 2.0 * n + 1.0;
    int i = (int)((nn - sqrt(nn * nn - 8.0 * (double)t)) * 0.5);
    long long start = (long long)i * n - (long long)i * (i - 1) / 2;
    while (start > t) { --i; start = (long long)i * n - (long long)i * (i - 1) / 2; }
    while (start + (n - i) <= t) { start += n - i; ++i; }
    const int j = i + (int)(t - start);
    if (i == j) {
        F[(size_t)i * n + i] *= 2.0;
    } else {
        const double s = F[(size_t)i * n + j] + F[(size_t)j * n + i];
        F[(size_t)i * n + j] = s;
        F[(size_t)j * n + i] = s;
    }
}

extern "C" int fj_close(hipStream_t s, int n, double* F) {
    if (n <= 0 || !F) return FJ_ERR_ARG;
    const long long total = (long long)n * (n + 1) / 2;
    const int threads = 256;
    hipLaunchKernelGGL(fj_close_k, dim3((unsigned)((total + threads - 1) / threads)), dim3(threads), 0, s,
                       n, F);
    return (int)hipGetLastError();
}

// Dsh[s][t] = max |D| over the AO block of shells (s, t): one thread per block.
__global__ void fj_density_blocks_k(int n_sh, int n_ao, const int* __restrict__ ao,
                                    const int* __restrict__ sz, const double* __restrict__ D,
                                    double* __restrict__ Dsh) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_sh * n_sh) return;
    const int s = t / n_sh, u = t % n_sh;
    double m = 0.0;
    for (int i = 0; i < sz[s]; ++i)
        for (int j = 0; j < sz[u]; ++j) m = fmax(m, fabs(D[(size_t)(ao[s] + i) * n_ao + ao[u] + j]));
    Dsh[t] = m;
}

extern "C" int fj_density_blocks(hipStream_t s, int n_sh, int n_ao, const int* shell_ao,
                                 const int* shell_size, const double* D, double* Dsh) {
    if (n_sh <= 0 || n_ao <= 0) return FJ_ERR_ARG;
    const int threads = 256, total = n_sh * n_sh;
    hipLaunchKernelGGL(fj_density_blocks_k, dim3((total + threads - 1) / threads), dim3(threads), 0, s,
                       n_sh, n_ao, shell_ao, shell_size, D, Dsh);
    return (int)hipGetLastError();
}

template <int M>
__global__ void fj_boys_k(int n, const double* __restrict__ T, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double F[M + 1];
    fj_boys<M>(T[i], F);
#pragma unroll
    for (int m = 0; m <= M; ++m) out[(size_t)i * (M + 1) + m] = F[m];
}

extern "C" int fj_boys_eval(hipStream_t s, int M, int n, const double* T, double* out) {
    if (n < 0) return FJ_ERR_ARG;
    if (n == 0) return 0;
    dim3 g((n + 255) / 256), b(256);
    switch (M) {
#define FJ_B(m) case m: hipLaunchKernelGGL(fj_boys_k<m>, g, b, 0, s, n, T, out); break;
        FJ_B(0) FJ_B(1) FJ_B(2) FJ_B(3) FJ_B(4) FJ_B(5) FJ_B(6) FJ_B(7) FJ_B(8) FJ_B(9) FJ_B(10)
        FJ_B(11) FJ_B(12)
#undef FJ_B
        default: return FJ_ERR_ARG;
    }
    return (int)hipGetLastError();
}

// K[oa[r] + i, ob + j] += Kr[r, i * nb + j]: the K kernel's per-bra-record
// blocks added into the exchange matrix (several records share a block).
__global__ void fj_k_scatter_k(int n_rec, int na, int nb, int n_ao, const int* __restrict__ oa, int ob,
                               const double* __restrict__ Kr, double* __restrict__ K) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nab = na * nb;
    if (t >= n_rec * nab) return;
    const int r = t / nab, e = t % nab;
    unsafeAtomicAdd(K + (size_t)(oa[r] + e / nb) * n_ao + ob + e % nb, Kr[t]);
}

extern "C" int fj_k_scatter(hipStream_t s, int n_rec, int na, int nb, int n_ao, const int* oa, int ob,
                            const double* Kr, double* K) {
    if (n_rec < 0 || na <= 0 || nb <= 0 || n_ao <= 0) return FJ_ERR_ARG;
    const long long total = (long long)n_rec * na * nb;
    if (total == 0) return 0;
    hipLaunchKernelGGL(fj_k_scatter_k, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       n_rec, na, nb, n_ao, oa, ob, Kr, K);
    return (int)hipGetLastError();
}
