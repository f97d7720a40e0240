// fj_mlp.hip -- the precision selector's network: x[F] -> 64 -> 64 -> 1,
// ReLU, BF16 weights, FP32 activations and accumulation.
//
//   forward:   h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = w3 . h2 + b3
//   backward:  the loss L = mean_i (y_i - t_i)^2, gradients of the FP32
//              master weights through the BF16 weights the forward used
//   adam:      one Adam step on the FP32 master weights, then their BF16
//              copies (round to nearest even)
//
// One lane per sample; the weights are staged in LDS per workgroup. The
// weight gradients are reduced in LDS per workgroup (LDS atomics) and then
// added to global memory once per workgroup.
#include "fj_device.h"
#include <hip/hip_bf16.h>

#define FJ_MLP_H 64
#define FJ_MLP_FMAX 32

// parameter block layout (floats): W1[H][F] b1[H] W2[H][H] b2[H] w3[H] b3
__host__ __device__ static inline int fj_mlp_nparam(int F) {
    return FJ_MLP_H * F + FJ_MLP_H + FJ_MLP_H * FJ_MLP_H + FJ_MLP_H + FJ_MLP_H + 1;
}

__device__ __forceinline__ float fj_bf(const uint16_t v) {
    return __uint_as_float(((uint32_t)v) << 16);
}

__device__ __forceinline__ uint16_t fj_to_bf(float f) {
    uint32_t u = __float_as_uint(f);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;  // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}

// weights: bf16 copies of W1, W2, w3 (same layout, biases kept FP32 in P)
struct FjMlpW {
    const uint16_t* w;  // bf16 [nparam] (bias slots unused)
    const float* p;     // fp32 master [nparam] (biases read from here)
};

template <bool KEEP>
__device__ __forceinline__ float fj_mlp_eval(int F, const float* __restrict__ x, const float* sW,
                                             float* h1, float* h2) {
    const float* W1 = sW;
    const float* b1 = W1 + FJ_MLP_H * F;
    const float* W2 = b1 + FJ_MLP_H;
    const float* b2 = W2 + FJ_MLP_H * FJ_MLP_H;
    const float* w3 = b2 + FJ_MLP_H;
    const float b3 = w3[FJ_MLP_H];
    for (int j = 0; j < FJ_MLP_H; ++j) {
        float a = b1[j];
        for (int k = 0; k < F; ++k) a = fmaf(W1[j * F + k], x[k], a);
        h1[j] = fmaxf(a, 0.0f);
    }
    float y = b3;
    for (int j = 0; j < FJ_MLP_H; ++j) {
        float a = b2[j];
#pragma unroll 8
        for (int k = 0; k < FJ_MLP_H; ++k) a = fmaf(W2[j * FJ_MLP_H + k], h1[k], a);
        a = fmaxf(a, 0.0f);
        if (KEEP) h2[j] = a;
        y = fmaf(w3[j], a, y);
    }
    return y;
}

// Stage the effective weights (bf16 values as floats, fp32 biases) in LDS.
__device__ void fj_mlp_stage(int F, FjMlpW W, float* sW) {
    const int n = fj_mlp_nparam(F);
    const int b1 = FJ_MLP_H * F, b2 = b1 + FJ_MLP_H + FJ_MLP_H * FJ_MLP_H, w3 = b2 + FJ_MLP_H;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const bool bias = (i >= b1 && i < b1 + FJ_MLP_H) || (i >= b2 && i < w3) || i == n - 1;
        sW[i] = bias ? W.p[i] : fj_bf(W.w[i]);
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) fj_mlp_forward_k(int n, int F, const float* __restrict__ X,
                                                        FjMlpW W, float* __restrict__ Y) {
    extern __shared__ float sW[];
    fj_mlp_stage(F, W, sW);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x[FJ_MLP_FMAX], h1[FJ_MLP_H];
    for (int k = 0; k < F; ++k) x[k] = X[(size_t)i * F + k];
    Y[i] = fj_mlp_eval<false>(F, x, sW, h1, nullptr);
}

// quantile < 0: mean squared error; 0 < quantile < 1: the pinball loss of
// that quantile, mean over samples of q (t - y) if t >= y else (q - 1)(t - y)
__global__ void __launch_bounds__(256) fj_mlp_backward_k(int n, int F, const float* __restrict__ X,
                                                         const float* __restrict__ T, FjMlpW W,
                                                         float* __restrict__ G, float* __restrict__ loss,
                                                         float quantile) {
    extern __shared__ float smem[];
    const int np = fj_mlp_nparam(F);
    float* sW = smem;
    float* sG = smem + np;
    for (int k = threadIdx.x; k < np; k += blockDim.x) sG[k] = 0.0f;
    fj_mlp_stage(F, W, sW);  // ends with a barrier, so sG is zeroed too
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < n) {
        float x[FJ_MLP_FMAX], h1[FJ_MLP_H], h2[FJ_MLP_H];
        for (int k = 0; k < F; ++k) x[k] = X[(size_t)i * F + k];
        const float y = fj_mlp_eval<true>(F, x, sW, h1, h2);
        const float r = y - T[i];
        float dy;
        if (quantile < 0.0f) {
            l = r * r / (float)n;
            dy = 2.0f * r / (float)n;
        } else if (r <= 0.0f) {   // t >= y
            l = -quantile * r / (float)n;
            dy = -quantile / (float)n;
        } else {
            l = (1.0f - quantile) * r / (float)n;
            dy = (1.0f - quantile) / (float)n;
        }
        const float* W2 = sW + FJ_MLP_H * F + FJ_MLP_H;
        const float* w3 = W2 + FJ_MLP_H * FJ_MLP_H + FJ_MLP_H;
        const int o_b1 = FJ_MLP_H * F, o_W2 = o_b1 + FJ_MLP_H, o_b2 = o_W2 + FJ_MLP_H * FJ_MLP_H;
        const int o_w3 = o_b2 + FJ_MLP_H, o_b3 = o_w3 + FJ_MLP_H;
        atomicAdd(&sG[o_b3], dy);
        float d1[FJ_MLP_H];
        for (int k = 0; k < FJ_MLP_H; ++k) d1[k] = 0.0f;
        for (int j = 0; j < FJ_MLP_H; ++j) {
            atomicAdd(&sG[o_w3 + j], dy * h2[j]);
            const float d2 = h2[j] > 0.0f ? dy * w3[j] : 0.0f;
            if (d2 != 0.0f) {
                atomicAdd(&sG[o_b2 + j], d2);
                for (int k = 0; k < FJ_MLP_H; ++k) {
                    atomicAdd(&sG[o_W2 + j * FJ_MLP_H + k], d2 * h1[k]);
                    d1[k] = fmaf(W2[j * FJ_MLP_H + k], d2, d1[k]);
                }
            }
        }
        for (int k = 0; k < FJ_MLP_H; ++k) {
            const float g = h1[k] > 0.0f ? d1[k] : 0.0f;
            if (g != 0.0f) {
                atomicAdd(&sG[o_b1 + k], g);
                for (int f = 0; f < F; ++f) atomicAdd(&sG[k * F + f], g * x[f]);
            }
        }
    }
    // the workgroup's loss and gradient, once per workgroup
    l = fj_wave_sum(l);
    if ((threadIdx.x & 63) == 0 && l != 0.0f) atomicAdd(loss, l);
    __syncthreads();
    for (int k = threadIdx.x; k < np; k += blockDim.x)
        if (sG[k] != 0.0f) atomicAdd(&G[k], sG[k]);
}

__global__ void fj_mlp_adam_k(int np, float* __restrict__ P, const float* __restrict__ G,
                              float* __restrict__ M, float* __restrict__ V, uint16_t* __restrict__ Wb,
                              float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float g = G[i];
    const float m = b1 * M[i] + (1.0f - b1) * g;
    const float v = b2 * V[i] + (1.0f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    const float p = P[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[i] = p;
    Wb[i] = fj_to_bf(p);
}

__global__ void fj_mlp_round_k(int np, const float* __restrict__ P, uint16_t* __restrict__ Wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) Wb[i] = fj_to_bf(P[i]);
}

extern "C" int fj_mlp_nparams(int F) { return fj_mlp_nparam(F); }

extern "C" int fj_mlp_forward(hipStream_t s, int n, int F, const float* X, const uint16_t* Wb,
                              const float* P, float* Y) {
    if (n < 0 || F < 1 || F > FJ_MLP_FMAX) return FJ_ERR_ARG;
    if (n == 0) return 0;
    FjMlpW W{Wb, P};
    const size_t lds = sizeof(float) * fj_mlp_nparam(F);
    hipLaunchKernelGGL(fj_mlp_forward_k, dim3((n + 255) / 256), dim3(256), lds, s, n, F, X, W, Y);
    return (int)hipGetLastError();
}

extern "C" int fj_mlp_backward(hipStream_t s, int n, int F, const float* X, const float* T,
                               const uint16_t* Wb, const float* P, float* G, float* loss) {
    if (n < 0 || F < 1 || F > FJ_MLP_FMAX) return FJ_ERR_ARG;
    if (n == 0) return 0;
    FjMlpW W{Wb, P};
    const size_t lds = 2 * sizeof(float) * fj_mlp_nparam(F);
    hipLaunchKernelGGL(fj_mlp_backward_k, dim3((n + 255) / 256), dim3(256), lds, s, n, F, X, T, W, G,
                       loss, -1.0f);
    return (int)hipGetLastError();
}

extern "C" int fj_mlp_backward_quantile(hipStream_t s, int n, int F, const float* X, const float* T,
                                        const uint16_t* Wb, const float* P, float* G, float* loss,
                                        float quantile) {
    if (n < 0 || F < 1 || F > FJ_MLP_FMAX || !(quantile > 0.0f && quantile < 1.0f)) return FJ_ERR_ARG;
    if (n == 0) return 0;
    FjMlpW W{Wb, P};
    const size_t lds = 2 * sizeof(float) * fj_mlp_nparam(F);
    hipLaunchKernelGGL(fj_mlp_backward_k, dim3((n + 255) / 256), dim3(256), lds, s, n, F, X, T, W, G,
                       loss, quantile);
    return (int)hipGetLastError();
}

extern "C" int fj_mlp_adam(hipStream_t s, int np, float* P, const float* G, float* M, float* V,
                           uint16_t* Wb, float lr, float b1, float b2, float eps, int t) {
    if (np <= 0 || t < 1) return FJ_ERR_ARG;
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(fj_mlp_adam_k, dim3((np + 255) / 256), dim3(256), 0, s, np, P, G, M, V, Wb, lr, b1,
                       b2, eps, c1, c2);
    return (int)hipGetLastError();
}

extern "C" int fj_mlp_round(hipStream_t s, int np, const float* P, uint16_t* Wb) {
    if (np <= 0) return FJ_ERR_ARG;
    hipLaunchKernelGGL(fj_mlp_round_k, dim3((np + 255) / 256), dim3(256), 0, s, np, P, Wb);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-build tier choice of a mixed launch, on the device, ahead of the ERI
// launch that reads it: one wavefront per work item. The lanes take the
// item's kets, evaluate the kernel's own six-block density screen with the
// CURRENT density block maxima and keep, over the kets that pass it, the
// largest bound (or the bound rule's amplified error estimate) and the
// largest six-block weight. The bound rule sends the item to FP32 when that
// estimate is below the error budget. Otherwise, with the learned selector,
// the item's density-free features (standardised once at plan time) get the
// density feature log10(max weight) and the network runs in the same
// wavefront, one hidden unit per lane, with the fj_mlp_eval arithmetic in
// the same order: FP32 iff the predicted log10 error plus the margin is
// below log10(budget) and the features lie in the training range.
__device__ __forceinline__ double fj_wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// The network staged in LDS for one wavefront per sample: W1 and W2
// transposed (W1T[k][j], W2T[k][j]), so that lane j's reads of row k are
// consecutive words (no bank conflicts), then b1, b2, w3, b3 -- the same
// values fj_mlp_stage stages, so the arithmetic and its order are
// fj_mlp_eval's.
__device__ void fj_mlp_stage_t(int F, FjMlpW W, float* sW) {
    const int n = fj_mlp_nparam(F);
    const int o_b1 = FJ_MLP_H * F, o_W2 = o_b1 + FJ_MLP_H, o_b2 = o_W2 + FJ_MLP_H * FJ_MLP_H;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int dst = i;
        float v;
        if (i < o_b1) {                      // W1[j][k] -> W1T[k][j]
            const int j = i / F, k = i - j * F;
            dst = k * FJ_MLP_H + j;
            v = fj_bf(W.w[i]);
        } else if (i < o_W2) {               // b1
            v = W.p[i];
        } else if (i < o_b2) {               // W2[j][k] -> W2T[k][j]
            const int r = i - o_W2, j = r / FJ_MLP_H, k = r - j * FJ_MLP_H;
            dst = o_W2 + k * FJ_MLP_H + j;
            v = fj_bf(W.w[i]);
        } else if (i < o_b2 + FJ_MLP_H || i == n - 1) {   // b2, b3
            v = W.p[i];
        } else {                             // w3
            v = fj_bf(W.w[i]);
        }
        sW[dst] = v;
    }
    __syncthreads();
}

// One work item's tier, by one wavefront (hx: the wavefront's LDS slots,
// 2 * 64 + 32 floats; sW: the network staged by fj_mlp_stage_t).
__device__ __forceinline__ void fj_tier_item(const FjSide& bra, const FjSide& ket, const FjSys* sys,
                                             const FjTierSel& t, int it, const float* sW, float* hx) {
    const int lane = threadIdx.x & 63;
    const int4 w = reinterpret_cast<const int4*>(t.work)[it];
    const FjSys S = sys[w.z & 0xffff];
    const int b = w.x, sa = bra.sh[2 * b], sb = bra.sh[2 * b + 1], ns = S.n_sh;
    const double Gb = bra.G[b];
    const double amp_b = t.amp_bra[b];
    const double dab = 2.0 * S.Dsh[sa * ns + sb];
    double est = -1.0, dm_max = 0.0;
    for (int e = w.y + lane; e < w.w; e += 64) {
        const int k = t.ket_list[e] & 0x1ffff;
        if (t.diag && k < b) continue;
        const int sc = ket.sh[2 * k], sd = ket.sh[2 * k + 1];
        double dm = fmax(dab, 2.0 * S.Dsh[sc * ns + sd]);
        dm = fmax(dm, 0.5 * S.Dsh[sa * ns + sc]);
        dm = fmax(dm, 0.5 * S.Dsh[sa * ns + sd]);
        dm = fmax(dm, 0.5 * S.Dsh[sb * ns + sc]);
        dm = fmax(dm, 0.5 * S.Dsh[sb * ns + sd]);
        const double bound = Gb * (ket.G[k] * dm);
        if (bound < t.tau) continue;   // the ERI kernel's screen skips this quartet
        const double e_k = t.fp32_bound >= 0.0 ? bound : bound * amp_b * t.amp_ket[k] * t.est_scale;
        est = fmax(est, e_k);
        dm_max = fmax(dm_max, dm);
    }
    est = fj_wave_max(est);
    dm_max = fj_wave_max(dm_max);
    int fp32 = 0;
    if (est >= 0.0) {   // some quartet of the item passes the screen
        fp32 = t.fp32_bound >= 0.0 ? (est < t.fp32_bound) : (est < t.fp32_error);
        if (!fp32 && t.feat && t.inside[it]) {
            const double xd = log10(dm_max);
            if (xd >= t.d_lo && xd <= t.d_hi) {
                const int F = t.n_feat;
                float* x = hx + 2 * FJ_MLP_H;
                if (lane < F - 1) x[lane] = t.feat[(size_t)it * F + lane];
                if (lane == 0) x[F - 1] = (float)((xd - t.d_mean) / t.d_std);
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                const float* W1T = sW;
                const float* b1 = W1T + FJ_MLP_H * F;
                const float* W2T = b1 + FJ_MLP_H;
                const float* b2 = W2T + FJ_MLP_H * FJ_MLP_H;
                const float* w3 = b2 + FJ_MLP_H;
                float a = b1[lane];
                for (int k = 0; k < F; ++k) a = fmaf(W1T[k * FJ_MLP_H + lane], x[k], a);
                hx[lane] = fmaxf(a, 0.0f);
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                a = b2[lane];
#pragma unroll 8
                for (int k = 0; k < FJ_MLP_H; ++k) a = fmaf(W2T[k * FJ_MLP_H + lane], hx[k], a);
                hx[FJ_MLP_H + lane] = fmaxf(a, 0.0f);
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                float y = w3[FJ_MLP_H];
                for (int j = 0; j < FJ_MLP_H; ++j) y = fmaf(w3[j], hx[FJ_MLP_H + j], y);
                fp32 = (double)y + t.margin < t.log_budget;
                __builtin_amdgcn_wave_barrier();   // hx is reused by the next item
            }
        }
    }
    if (lane == 0) t.tier[it] = fp32;
}

__global__ void __launch_bounds__(256) fj_tier_select_k(FjSide bra, FjSide ket, const FjSys* sys,
                                                        FjTierSel t, FjMlpW W) {
    extern __shared__ float smem[];
    const int np = t.feat ? fj_mlp_nparam(t.n_feat) : 0;
    float* hx = smem + np + (threadIdx.x >> 6) * (2 * FJ_MLP_H + FJ_MLP_FMAX);  // per wavefront
    if (t.feat) fj_mlp_stage_t(t.n_feat, W, smem);   // ends with a barrier
    const int waves = (int)(gridDim.x * (blockDim.x >> 6));
    for (int it = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); it < t.n_items;
         it += waves)
        fj_tier_item(bra, ket, sys, t, it, smem, hx);
}

// Every mixed launch of a build in one launch: job j's items are the global
// items [first[j], first[j + 1]); one network (W, n_feat) for every job.
__global__ void __launch_bounds__(256) fj_tier_select_jobs_k(const FjTierJob* __restrict__ jobs,
                                                             int n_jobs, int n_items, const FjSys* sys,
                                                             int n_feat, FjMlpW W) {
    extern __shared__ float smem[];
    const int np = n_feat ? fj_mlp_nparam(n_feat) : 0;
    float* hx = smem + np + (threadIdx.x >> 6) * (2 * FJ_MLP_H + FJ_MLP_FMAX);
    if (n_feat) fj_mlp_stage_t(n_feat, W, smem);
    const int waves = (int)(gridDim.x * (blockDim.x >> 6));
    for (int g = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); g < n_items; g += waves) {
        int lo = 0, hi = n_jobs - 1;   // the last job whose first item is <= g
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobs[mid].first <= g) lo = mid;
            else hi = mid - 1;
        }
        const FjTierJob& J = jobs[lo];
        fj_tier_item(J.bra, J.ket, sys, J.sel, g - J.first, smem, hx);
    }
}

extern "C" int fj_tier_select(hipStream_t s, const FjSide* bra, const FjSide* ket, const FjSys* sys,
                              const FjTierSel* sel) {
    if (!bra || !ket || !sys || !sel || !sel->work || !sel->ket_list || !sel->tier ||
        !sel->amp_bra || !sel->amp_ket || sel->n_items < 0)
        return FJ_ERR_ARG;
    if (sel->feat && (sel->n_feat < 2 || sel->n_feat > FJ_MLP_FMAX || !sel->inside || !sel->Wb ||
                      !sel->P || !(sel->d_std > 0.0)))
        return FJ_ERR_ARG;
    if (sel->n_items == 0) return 0;
    FjMlpW W{sel->Wb, sel->P};
    const int np = sel->feat ? fj_mlp_nparam(sel->n_feat) : 0;
    const size_t lds = sizeof(float) * (np + 4 * (2 * FJ_MLP_H + FJ_MLP_FMAX));
    const int blocks = (sel->n_items + 3) / 4 < 1024 ? (sel->n_items + 3) / 4 : 1024;
    hipLaunchKernelGGL(fj_tier_select_k, dim3(blocks), dim3(256), lds, s, *bra, *ket, sys, *sel, W);
    return (int)hipGetLastError();
}

extern "C" int fj_tier_select_jobs(hipStream_t s, int n_jobs, const FjTierJob* jobs, int n_items,
                                   const FjSys* sys, int n_feat, const unsigned short* Wb, const float* P) {
    if (n_jobs < 0 || n_items < 0 || !sys || (n_jobs && !jobs) || n_feat < 0 || n_feat > FJ_MLP_FMAX ||
        n_feat == 1 || (n_feat && (!Wb || !P)))
        return FJ_ERR_ARG;
    if (n_jobs == 0 || n_items == 0) return 0;
    FjMlpW W{reinterpret_cast<const uint16_t*>(Wb), P};
    const int np = n_feat ? fj_mlp_nparam(n_feat) : 0;
    const size_t lds = sizeof(float) * (np + 4 * (2 * FJ_MLP_H + FJ_MLP_FMAX));
    const int blocks = (n_items + 3) / 4 < 1024 ? (n_items + 3) / 4 : 1024;
    hipLaunchKernelGGL(fj_tier_select_jobs_k, dim3(blocks), dim3(256), lds, s, jobs, n_jobs, n_items, sys,
                       n_feat, W);
    return (int)hipGetLastError();
}
