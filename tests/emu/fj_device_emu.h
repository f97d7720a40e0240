// CPU emulation of felizjueves_amd/csrc/fj_device.h for the generated F
// kernels (TEST INFRASTRUCTURE): one std::thread per lane of one workgroup,
// a barrier for __syncthreads and for the wavefront shuffles. Lets a test
// run a generated kernel unchanged on the host under AddressSanitizer.
#pragma once
#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
using std::min; using std::max; using std::fma; using std::sqrt; using std::exp; using std::fmax;
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(x)
#define __shared__ static
#define __restrict__ __restrict
struct dim3 { unsigned x, y, z; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
struct uint3e { unsigned x, y, z; };
extern thread_local uint3e threadIdx;
extern uint3e blockIdx, blockDim, gridDim;
struct int4 { int x, y, z, w; };
typedef struct ihipStream_t* hipStream_t;
#include "fj_mi355x.h"
#include "gen/fj_boys_table.h"
#define FJ_SREC 9
#define FJ_MAX_PRIM_PAIRS 128
#define FJ_MAX_THREADS 256
#define FJ_DEFAULT_THREADS 64
#define FJ_TWO_OVER_SQRT_PI 1.1283791670955125739
#define FJ_PI 3.141592653589793238
#define FJ_BOYS_ROW 24
#define FJ_BOYS_TMAX 40.0
struct FjFArgs {
    FjSide bra, ket; const int4* work; const int* ket_list; const FjSys* sys; FjSys one;
    int diag; double tau; double* ws; long long ws_doubles; int n_bx; int n_items; int n_fp64;
    const int* tier;
};
inline bool fj_item_fp64(const FjFArgs& a, int i) { return a.tier ? a.tier[i] == 0 : i < a.n_fp64; }
extern std::barrier<>* g_bar;
extern double g_shfl[256];
#define FJ_LAUNDER(p, s) ((void)0)
// the dynamic LDS block: a heap buffer of the launch's exact size, set by the driver
extern double* g_dyn_lds;
#define FJ_DYN_SHARED(name) double* const name = g_dyn_lds
inline void __syncthreads() { g_bar->arrive_and_wait(); }
inline double __shfl_xor(double v, int o, int w) {
    g_bar->arrive_and_wait();
    g_shfl[threadIdx.x] = v;
    g_bar->arrive_and_wait();
    double r = g_shfl[threadIdx.x ^ o];
    g_bar->arrive_and_wait();
    return r;
}
// The FP64 matrix-core step D = A B + C (16 x 4 by 4 x 16) of one
// wavefront: lane l gives A[l % 16][l / 16] and B[l / 16][l % 16] and gets
// D[l / 16 + 4 v][l % 16] in element v (the layout measured on gfx950).
// The wavefront's lanes meet at its own barrier (g_wbar, set by the driver),
// so every lane of a wavefront must call it together, as on the GPU.
typedef double fj_v4d __attribute__((vector_size(32)));
extern std::barrier<>* g_wbar[4];
extern double g_mfa[2][256], g_mfb[2][256];
extern thread_local int g_mf_parity;
inline fj_v4d fj_mfma_f64_16x16x4(double a, double b, fj_v4d c) {
    // two exchange buffers used in turn: a lane passes step k's barrier only
    // after writing step k, so after reading step k - 1 -- the buffer step
    // k + 1 writes is free once step k's barrier is passed
    const int l = (int)threadIdx.x, w = l >> 6, lw = l & 63, base = w << 6, ph = g_mf_parity;
    g_mf_parity ^= 1;
    g_mfa[ph][l] = a;
    g_mfb[ph][l] = b;
    g_wbar[w]->arrive_and_wait();
    const int j = lw & 15;
    for (int v = 0; v < 4; ++v) {
        const int i = (lw >> 4) + 4 * v;
        double s = c[v];
        for (int k = 0; k < 4; ++k) s += g_mfa[ph][base + i + 16 * k] * g_mfb[ph][base + 16 * k + j];
        c[v] = s;
    }
    return c;
}
inline void unsafeAtomicAdd(double* p, double v) { std::atomic_ref<double>(*p).fetch_add(v); }
inline double atomicAdd(double* p, double v) { return std::atomic_ref<double>(*p).fetch_add(v); }
inline void fj_side_range(const FjSide& s, int i, int& k0, int& k) {
    if (s.off) { k0 = s.off[i]; k = s.off[i + 1] - k0; } else { k0 = i * s.k; k = s.k; }
}
inline double fj_wave_sum(double v) { for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64); return v; }
template <int M>
inline void fj_boys(double T, double (&F)[M + 1]) {
    if (T < FJ_BOYS_TMAX) {
        const int i = (int)(T * 8.0 + 0.5);
        if (i < 0 || i >= FJ_BOYS_NROW) { fprintf(stderr, "boys row %d T %g\n", i, T); abort(); }
        const double x = (double)i * 0.125 - T;
        const double* r = fj_boys_tab + FJ_BOYS_ROW * i + 1 + M;
        double f = r[8];
        for (int k = 8; k >= 1; --k) f = fma(f, x * (1.0 / k), r[k - 1]);
        if constexpr (M == 0) { F[0] = f; } else {
            double ex = 1.0 / 40320.0;
            const double cs[8] = {1.0/5040, 1.0/720, 1.0/120, 1.0/24, 1.0/6, 0.5, 1.0, 1.0};
            for (double c : cs) ex = fma(ex, x, c);
            const double e = fj_boys_tab[FJ_BOYS_ROW * i] * ex;
            F[M] = f;
            for (int m = M; m > 0; --m) { f = (2.0 * T * f + e) / (double)(2 * m - 1); F[m - 1] = f; }
        }
    } else {
        double f = 0.5 * sqrt(FJ_PI / T); F[0] = f;
        if constexpr (M > 0) { const double e = T < 80.0 ? exp(-T) : 0.0; const double h = 0.5 / T;
            for (int m = 1; m <= M; ++m) { f = ((double)(2 * m - 1) * f - e) * h; F[m] = f; } }
    }
}
#define FJ_BOYS_LROW 10
template <int M>
inline void fj_boys_stage(double* tb) {
    static_assert(M + 8 <= 20, "Boys table holds F_0..F_20");
    for (int j = (int)threadIdx.x; j < FJ_BOYS_NROW * FJ_BOYS_LROW; j += (int)blockDim.x) {
        const int i = j / FJ_BOYS_LROW, c = j - i * FJ_BOYS_LROW;
        const int g = FJ_BOYS_ROW * i + (c ? M + c : 0);
        if (g < 0 || g >= FJ_BOYS_NROW * FJ_BOYS_ROW || (c && M + c >= FJ_BOYS_ROW)) {
            fprintf(stderr, "boys stage index %d\n", g); abort();
        }
        tb[j] = fj_boys_tab[g];
    }
}
// the LDS copy must hold exactly what the global table holds for order M
template <int M>
inline void fj_boys_lds(double T, double (&F)[M + 1], const double* tb) {
    if (T < FJ_BOYS_TMAX) {
        const int i = (int)(T * 8.0 + 0.5);
        if (i < 0 || i >= FJ_BOYS_NROW) { fprintf(stderr, "boys lds row %d T %g\n", i, T); abort(); }
        for (int c = 0; c < FJ_BOYS_LROW; ++c) {
            const double want = fj_boys_tab[FJ_BOYS_ROW * i + (c ? M + c : 0)];
            if (tb[FJ_BOYS_LROW * i + c] != want) { fprintf(stderr, "boys lds row %d col %d differs\n", i, c); abort(); }
        }
    }
    fj_boys<M>(T, F);
}
