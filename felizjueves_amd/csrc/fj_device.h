// fj_device.h -- shared device code of the gfx950 Fock-build library:
// argument records, the Boys function, wavefront reductions, error codes.
// Written for CDNA4 (gfx950, wave64) only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fj_mi355x.h"
#include "gen/fj_boys_table.h"

#define FJ_SREC 9                 // staged bra record: P(3) X(3) zeta U 1/zeta
#define FJ_MAX_PRIM_PAIRS 128     // primitive pairs per contracted pair
#define FJ_MAX_THREADS 256
#define FJ_DEFAULT_THREADS 64
#define FJ_TWO_OVER_SQRT_PI 1.1283791670955125739
#define FJ_PI 3.141592653589793238
#define FJ_BOYS_ROW 24
#define FJ_BOYS_TMAX 40.0

// Launch arguments shared by the per-class kernels (both C-ABI entries).
struct FjFArgs {
    FjSide bra, ket;
    const int4* work;      // optional: workgroup -> (bra pair, first ket, system, ket end)
    const int* ket_list;   // optional: work item kets are ket_list[first .. end)
    const FjSys* sys;      // the system table the work items index (device)
    FjSys one;             // the single system of a launch without a work list
    int diag;
    double tau;
    double* ws;            // launch workspace (classes with ws_doubles > 0)
    long long ws_doubles;  // its size
    int n_bx;              // work items: n_items = n_bx * (grid y of a walk launch)
    int n_items;
    int n_fp64;            // a mixed launch: items [0, n_fp64) FP64, the rest FP32
    const int* tier;       // a mixed launch: per item 0 FP64, 1 FP32 (overrides n_fp64)
};

// The digestion of a split J or K launch (fj_hj_ / fj_hk_): the weight of
// each output block of the built quartet (ab, cd, ac, ad, bc, bd; one is 1,
// the others 0), the strides of the density element (r, c) read, D[r * dr +
// c * dc], and of the output element written, F[r * fr + c * fc] (a block
// transposed by the launch's orientation), and the built positions of the
// density pair whose block maximum screens a quartet.
struct FjDig {
    double w[6];
    int dr, dc, fr, fc;
    int q0, q1;
};

// The FjDig of a split launch: screen 0 (J: the output on the orientation's
// positions (g^-1[0], g^-1[1]), the density on (g^-1[2], g^-1[3])) or 1 (K:
// output (g^-1[0], g^-1[2]), density (g^-1[1], g^-1[3])), g the orientation
// GROUP[perm] (host code).
int fj_split_dig(int screen, int perm, int n_ao, FjDig* dg);

// The tier of work item i of a mixed launch.
__device__ __forceinline__ bool fj_item_fp64(const FjFArgs& a, int i) {
    return a.tier ? a.tier[i] == 0 : i < a.n_fp64;
}

// Workgroups of a workspace class: at most FJ_WS_MAX_BLOCKS in one launch;
// the upstream-ABI entries (no workspace argument) use a library-owned one
// per stream, sized for FJ_LIB_WS_BLOCKS.
#define FJ_WS_MAX_BLOCKS 1024
#define FJ_LIB_WS_BLOCKS 256
// The fixed device array of fj_f_noalloc, per GPU: 64 MiB (the (dd|dd)
// kernel's 2,113 doubles per lane fit 62 workgroups of 64 lanes).
#define FJ_NOALLOC_SCRATCH_BYTES (64LL << 20)

// The library-owned workspace of at least `bytes` for launches on stream s
// of the current device: one per (device, stream), so launches on different
// streams may overlap, those on one stream run in order (allocated on first
// use, grown when asked for more; NULL on failure).
double* fj_library_workspace(hipStream_t s, long long bytes);

// A kernel's dynamic LDS block (its size given at launch).
#define FJ_DYN_SHARED(name) extern __shared__ double name[]

// One FP64 matrix-core step of a wavefront, D = A B + C with A 16 x 4, B
// 4 x 16, C and D 16 x 16: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[l / 16 + 4 v][l % 16] in element v. Every
// lane of the wavefront calls it together.
typedef double fj_v4d __attribute__((ext_vector_type(4)));
__device__ __forceinline__ fj_v4d fj_mfma_f64_16x16x4(double a, double b, fj_v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// An empty asm statement that the compiler must assume redefines the
// per-lane pointer p (a VGPR pair) and the uniform stride s (an SGPR pair):
// values derived from them are recomputed after it, not hoisted above it.
#define FJ_LAUNDER(p, s) __asm__ volatile("" : "+v"(p), "+s"(s))

// First primitive record and degree of pair i of one side.
__device__ __forceinline__ void fj_side_range(const FjSide& s, int i, int& k0, int& k) {
    if (s.off) {
        k0 = s.off[i];
        k = s.off[i + 1] - k0;
    } else {
        k0 = i * s.k;
        k = s.k;
    }
}

// Sum of v over the 64 lanes of the wavefront (butterfly, every lane gets it).
__device__ __forceinline__ double fj_wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Boys function F_m(T), m = 0..M, FP64.
//  T < 40: F_M by a 9-term Taylor expansion about the nearest tabulated
//          T_i = i/8 (|T - T_i| <= 1/16, truncation < 1e-16 relative),
//          e^-T = e^-T_i e^{T_i - T} by the same 9-term series, then the
//          downward recursion F_{m-1} = (2T F_m + e^-T) / (2m - 1), stable.
//  T >= 40: F_0 = (1/2) sqrt(pi/T) (erfc(sqrt 40) < 2e-19) and the upward
//          recursion F_{m+1} = ((2m+1) F_m - e^-T) / (2T), stable for T > m.
// The table holds correctly rounded values of the Kummer series.
// fj_boys_from reads row i of a table with ROW doubles per row: e^-T_i at
// column 0 and F_M(T_i) .. F_{M+8}(T_i) from column FO on.
template <int M, int ROW, int FO>
__device__ __forceinline__ void fj_boys_from(double T, double (&F)[M + 1],
                                             const double* __restrict__ tab) {
    static_assert(M + 8 <= 20, "Boys table holds F_0..F_20");
    if (T < FJ_BOYS_TMAX) {
        const int i = (int)(T * 8.0 + 0.5);
        const double x = (double)i * 0.125 - T;
        const double* __restrict__ r = tab + ROW * i + FO;
        double f = r[8];
        f = fma(f, x * (1.0 / 8.0), r[7]);
        f = fma(f, x * (1.0 / 7.0), r[6]);
        f = fma(f, x * (1.0 / 6.0), r[5]);
        f = fma(f, x * (1.0 / 5.0), r[4]);
        f = fma(f, x * (1.0 / 4.0), r[3]);
        f = fma(f, x * (1.0 / 3.0), r[2]);
        f = fma(f, x * (1.0 / 2.0), r[1]);
        f = fma(f, x, r[0]);
        if constexpr (M == 0) {
            F[0] = f;
        } else {
            double ex = 1.0 / 40320.0;
            ex = fma(ex, x, 1.0 / 5040.0);
            ex = fma(ex, x, 1.0 / 720.0);
            ex = fma(ex, x, 1.0 / 120.0);
            ex = fma(ex, x, 1.0 / 24.0);
            ex = fma(ex, x, 1.0 / 6.0);
            ex = fma(ex, x, 0.5);
            ex = fma(ex, x, 1.0);
            ex = fma(ex, x, 1.0);
            const double e = tab[ROW * i] * ex;
            const double t2 = 2.0 * T;
            F[M] = f;
#pragma unroll
            for (int m = M; m > 0; --m) {
                f = (t2 * f + e) * (1.0 / (double)(2 * m - 1));
                F[m - 1] = f;
            }
        }
    } else {
        double f = 0.5 * sqrt(FJ_PI / T);
        F[0] = f;
        if constexpr (M > 0) {
            const double e = T < 80.0 ? exp(-T) : 0.0;
            const double h = 0.5 / T;
#pragma unroll
            for (int m = 1; m <= M; ++m) {
                f = ((double)(2 * m - 1) * f - e) * h;
                F[m] = f;
            }
        }
    }
}

template <int M>
__device__ __forceinline__ void fj_boys(double T, double (&F)[M + 1]) {
    fj_boys_from<M, FJ_BOYS_ROW, 1 + M>(T, F, fj_boys_tab);
}

// The columns of the Boys table one order M needs (e^-T_i, F_M .. F_{M+8}),
// FJ_BOYS_LROW doubles per row, staged in LDS by a workgroup that evaluates
// many Boys functions (a persistent kernel stages it once): the Taylor step
// then reads LDS instead of scattered global loads.
#define FJ_BOYS_LROW 10

template <int M>
__device__ __forceinline__ void fj_boys_stage(double* __restrict__ tb) {
    static_assert(M + 8 <= 20, "Boys table holds F_0..F_20");
    for (int j = (int)threadIdx.x; j < FJ_BOYS_NROW * FJ_BOYS_LROW; j += (int)blockDim.x) {
        const int i = j / FJ_BOYS_LROW, c = j - i * FJ_BOYS_LROW;
        tb[j] = fj_boys_tab[FJ_BOYS_ROW * i + (c ? M + c : 0)];
    }
}

template <int M>
__device__ __forceinline__ void fj_boys_lds(double T, double (&F)[M + 1], const double* __restrict__ tb) {
    fj_boys_from<M, FJ_BOYS_LROW, 1>(T, F, tb);
}
