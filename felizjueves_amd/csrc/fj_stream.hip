// fj_stream.hip -- the fused-F pair stream on the device (upstream
// eri/fock/launch_f.py:390-470, f_launches; the primitive-pair bound and the
// prune of ao/pairs.py:270-290,528-560; density_blocks of kernel_f.py:608):
// for every unique shell pair (a, b) (canonical: la <= lb, the lower index
// first on a tie) the bound G_[ab] of each primitive pair, their exact sum
// G, the largest G and eps = tau / (2 G_max), the prune of each pair's
// smallest-bound tail summing below eps, the density block maximum dmax,
// the key G dmax, the eq. 29 magnitude index, the bins (la, lb, degree,
// magnitude) ascending with each bin's pairs by key descending (stably, in
// pair order), and the launches: every batch pair I <= J whose magnitude
// indices sum below n. All of it on the device: the host hands over the
// basis, the density block maxima, tau, n and the magnitudes tau^(i/n),
// and reads back the plan.
#include "fj_device.h"

#include <hipcub/hipcub.hpp>

// Every multiply and add rounded on its own (no fused multiply-adds), the
// host planner's operation order, so the bounds agree with it to the
// transcendental functions' last bits.
#pragma clang fp contract(off)

namespace {

constexpr int LMAX = 3;
constexpr int NT = (2 * LMAX + 1) * (2 * LMAX + 2) * (2 * LMAX + 3) / 6;   // |t| <= 6: 84
constexpr unsigned DELETED = 0xffffffffu;

__device__ __forceinline__ int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// Component k of the Cartesians of one shell, (i, j, k) with i descending,
// then j descending (the AO order).
__device__ __forceinline__ void cart_of(int l, int idx, int& x, int& y, int& z) {
    int n = 0;
    for (int i = l; i >= 0; --i)
        for (int j = l - i; j >= 0; --j) {
            if (n == idx) {
                x = i;
                y = j;
                z = l - i - j;
                return;
            }
            ++n;
        }
    x = y = z = 0;
}

__device__ __forceinline__ double dfact_odd(int n) {   // (2n - 1)!!
    double out = 1.0;
    for (int k = 1; k < 2 * n; k += 2) out *= (double)k;
    return out;
}

// Exact-to-rounding summation (Neumaier's compensated sum).
struct CSum {
    double s = 0.0, c = 0.0;
    __device__ void add(double x) {
        const double t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ double value() const { return s + c; }
};

__device__ __forceinline__ void atomic_max_pos(unsigned long long* p, double v) {
    // v >= 0: the bit patterns order as the values
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

}  // namespace

// One thread per primitive pair: G_[ab] = max over the pair's component
// rows (a, b) of sum_t |c_a c_b (pi/zeta)^{3/2} E^{ab}_t| I_t, t in the box
// t_d <= a_d + b_d in Hermite order, I_t^2 = (2/sqrt(pi)) sqrt(zeta/2)
// zeta^|t| (2t_x-1)!!(2t_y-1)!!(2t_z-1)!! / (2|t| + 1).
__global__ void __launch_bounds__(256) fj_stream_bounds_k(
    int n_prim, int n_pair, const int* __restrict__ pair_ab, const int* __restrict__ prim_off,
    const int* __restrict__ shell_l, const double* __restrict__ centre, const int* __restrict__ nprim,
    const int* __restrict__ pstart, const double* __restrict__ exps, const double* __restrict__ coefs,
    double* __restrict__ g) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= n_prim) return;
    int lo = 0, hi = n_pair - 1;   // the last pair whose first primitive pair is <= gi
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prim_off[mid] <= gi) lo = mid;
        else hi = mid - 1;
    }
    const int p = lo, sa = pair_ab[2 * p], sb = pair_ab[2 * p + 1];
    const int la = shell_l[sa], lb = shell_l[sb];
    const int loc = gi - prim_off[p], nb_ = nprim[sb];
    const int i = loc / nb_, j = loc - i * nb_;
    const double al = exps[pstart[sa] + i], be = exps[pstart[sb] + j];
    const double ca = coefs[pstart[sa] + i], cb = coefs[pstart[sb] + j];
    const double zeta = al + be;
    const double pref = ca * cb * pow(FJ_PI / zeta, 1.5);
    const double h = 0.5 / zeta;
    // per-axis E[i][j][t] (i <= la, j <= lb, t <= la + lb)
    double E[3][LMAX + 1][LMAX + 1][2 * LMAX + 1];
    const int T = la + lb + 1;
    for (int d = 0; d < 3; ++d) {
        const double A = centre[3 * sa + d], B = centre[3 * sb + d];
        const double P = (al * A + be * B) / zeta;
        const double pa = P - A, pb = P - B;
        for (int x = 0; x <= la; ++x)
            for (int y = 0; y <= lb; ++y)
                for (int t = 0; t < T; ++t) E[d][x][y][t] = 0.0;
        E[d][0][0][0] = exp(zeta * pa * pb);
        for (int x = 0; x <= la; ++x) {
            if (x > 0) {
                for (int t = 0; t < T; ++t) {
                    double v = pa * E[d][x - 1][0][t];
                    if (t >= 1) v += h * E[d][x - 1][0][t - 1];
                    if (t + 1 < T) v += (double)(t + 1) * E[d][x - 1][0][t + 1];
                    E[d][x][0][t] = v;
                }
            }
            for (int y = 1; y <= lb; ++y) {
                for (int t = 0; t < T; ++t) {
                    double v = pb * E[d][x][y - 1][t];
                    if (t >= 1) v += h * E[d][x][y - 1][t - 1];
                    if (t + 1 < T) v += (double)(t + 1) * E[d][x][y - 1][t + 1];
                    E[d][x][y][t] = v;
                }
            }
        }
    }
    // I_t over the Hermite order |t| <= la + lb
    const int L = la + lb;
    double Ip[NT];
    int tc[NT];
    int nh = 0;
    for (int s = 0; s <= L; ++s)
        for (int tx = s; tx >= 0; --tx)
            for (int ty = s - tx; ty >= 0; --ty) {
                const int tz = s - tx - ty;
                const double df = dfact_odd(tx) * dfact_odd(ty) * dfact_odd(tz);
                Ip[nh] = sqrt((2.0 / sqrt(FJ_PI)) * sqrt(zeta / 2.0) * pow(zeta, (double)s) * df /
                              (double)(2 * s + 1));
                tc[nh] = tx | (ty << 4) | (tz << 8);
                ++nh;
            }
    double best = 0.0;
    for (int a = 0; a < ncart(la); ++a) {
        int ax, ay, az;
        cart_of(la, a, ax, ay, az);
        for (int b = 0; b < ncart(lb); ++b) {
            int bx, by, bz;
            cart_of(lb, b, bx, by, bz);
            double s = 0.0;
            for (int k = 0; k < nh; ++k) {
                const int tx = tc[k] & 15, ty = (tc[k] >> 4) & 15, tz = tc[k] >> 8;
                if (tx > ax + bx || ty > ay + by || tz > az + bz) continue;
                const double e = E[0][ax][bx][tx] * E[1][ay][by][ty] * E[2][az][bz][tz];
                s += fabs(pref * e) * Ip[k];
            }
            best = fmax(best, s);
        }
    }
    g[gi] = best;
}

// One thread per pair: G = the (compensated) sum of its primitive bounds,
// and the largest G.
__global__ void fj_stream_sums_k(int n_pair, const int* __restrict__ prim_off, const double* __restrict__ g,
                                 double* __restrict__ G, unsigned long long* __restrict__ gmax_bits) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pair) return;
    CSum s;
    for (int k = prim_off[p]; k < prim_off[p + 1]; ++k) s.add(g[k]);
    const double v = s.value();
    G[p] = v;
    atomic_max_pos(gmax_bits, v);
}

// One thread per pair: the prune (kept primitive pairs in bound-descending
// order, stable; the longest tail whose exact sum is below eps dropped),
// then the key, the magnitude index and the bin id (DELETED when nothing
// is kept).
__global__ void fj_stream_prune_k(int n_pair, const int* __restrict__ pair_ab, const int* __restrict__ shell_l,
                                  const int* __restrict__ prim_off, const double* __restrict__ g,
                                  const double* __restrict__ G, const unsigned long long* __restrict__ gmax_bits,
                                  double tau, int n, const double* __restrict__ mags,
                                  const double* __restrict__ Dsh, int n_sh, int ket_break,
                                  int* __restrict__ degree, int* __restrict__ keep,
                                  double* __restrict__ dmax, double* __restrict__ key,
                                  int* __restrict__ mag, unsigned* __restrict__ bin,
                                  double* __restrict__ eps_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pair) return;
    const double gmax = __longlong_as_double((long long)*gmax_bits);
    const double eps = (tau > 0.0 && gmax > 0.0) ? tau / (2.0 * gmax) : 0.0;
    if (p == 0) *eps_out = eps;
    const int k0 = prim_off[p], n_k = prim_off[p + 1] - k0;
    int* __restrict__ kp = keep + k0;
    // stable insertion sort of the local indices by bound, descending
    for (int k = 0; k < n_k; ++k) {
        const double v = g[k0 + k];
        int m = k;
        while (m > 0 && g[k0 + kp[m - 1]] < v) {
            kp[m] = kp[m - 1];
            --m;
        }
        kp[m] = k;
    }
    int cut = n_k;
    if (eps > 0.0) {
        CSum t;
        while (cut > 0) {
            CSum u = t;
            u.add(g[k0 + kp[cut - 1]]);
            if (u.value() < eps) {
                t = u;
                --cut;
            } else {
                break;
            }
        }
    }
    degree[p] = cut;
    const int sa = pair_ab[2 * p], sb = pair_ab[2 * p + 1];
    const double dm = Dsh[(size_t)sa * n_sh + sb];
    const double Gp = G[p];
    dmax[p] = dm;
    key[p] = ket_break ? Gp : Gp * dm;
    int mi = 0;
    if (tau > 0.0) {
        for (int i = n - 1; i > 0; --i)
            if (Gp < mags[i]) {
                mi = i;
                break;
            }
    }
    mag[p] = mi;
    bin[p] = cut ? (unsigned)(((shell_l[sa] * 4 + shell_l[sb]) * (FJ_MAX_PRIM_PAIRS + 1) + cut) * n + mi)
                 : DELETED;
}

__global__ void fj_iota_k(int n, int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i;
}

__global__ void fj_gather_u32_k(int n, const int* __restrict__ idx, const unsigned* __restrict__ src,
                                unsigned* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

__global__ void fj_zero_u64_k(long long n, unsigned long long* __restrict__ p) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// The number of batches and of kept pairs (one lane).
__global__ void fj_stream_count_k(int n_pair, const unsigned* __restrict__ bin_sorted,
                                  const int* __restrict__ flag, const int* __restrict__ pos,
                                  int* __restrict__ n_batch, int* __restrict__ n_kept) {
    if (threadIdx.x != 0) return;
    *n_batch = pos[n_pair - 1] + flag[n_pair - 1];
    int lo = 0, hi = n_pair;   // the first DELETED position (bins sorted ascending)
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (bin_sorted[mid] == DELETED) hi = mid;
        else lo = mid + 1;
    }
    *n_kept = lo;
}

// Batch starts over the sorted bins, and each batch's magnitude index and
// largest G.
__global__ void fj_stream_flags_k(int n_pair, const unsigned* __restrict__ bin_sorted, int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pair) return;
    const unsigned b = bin_sorted[i];
    flag[i] = b != DELETED && (i == 0 || bin_sorted[i - 1] != b);
}

__global__ void fj_stream_batches_k(int n_pair, const unsigned* __restrict__ bin_sorted,
                                    const int* __restrict__ flag, const int* __restrict__ pos,
                                    const int* __restrict__ order, const double* __restrict__ G,
                                    int* __restrict__ batch_start, int* __restrict__ batch_mag,
                                    unsigned long long* __restrict__ batch_gmax, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pair) return;
    const unsigned b = bin_sorted[i];
    if (b == DELETED) return;
    // pos[i] = flags before i: the batch of position i is pos[i] + flag[i] - 1
    const int bt = pos[i] + flag[i] - 1;
    if (flag[i]) {
        batch_start[bt] = i;
        batch_mag[bt] = (int)(b % (unsigned)n);
    }
    atomic_max_pos(batch_gmax + bt, G[order[i]]);
}

// The launches: every batch pair I <= J (I outer, J inner) whose magnitude
// indices sum below n (eq. 29; the others are screened before launch). One
// workgroup: lane 0 sizes each row, then the lanes fill them.
__global__ void __launch_bounds__(256) fj_stream_launches_k(int n_batch, int n,
                                                            const int* __restrict__ batch_mag,
                                                            int* __restrict__ row_off,
                                                            int* __restrict__ launches,
                                                            int* __restrict__ n_launch) {
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int bi = 0; bi < n_batch; ++bi) {
            row_off[bi] = tot;
            for (int ki = bi; ki < n_batch; ++ki) tot += batch_mag[bi] + batch_mag[ki] < n;
        }
        row_off[n_batch] = tot;
        *n_launch = tot;
    }
    __syncthreads();
    for (int bi = threadIdx.x; bi < n_batch; bi += blockDim.x) {
        int o = row_off[bi];
        for (int ki = bi; ki < n_batch; ++ki)
            if (batch_mag[bi] + batch_mag[ki] < n) {
                launches[2 * o] = bi;
                launches[2 * o + 1] = ki;
                ++o;
            }
    }
}

extern "C" int fj_stream_launches(hipStream_t s, int n_batch, int n, const int* batch_mag, int* row_off,
                                  int* launches, int* n_launch) {
    if (n_batch < 0 || n < 1 || (n_batch && (!batch_mag || !row_off || !launches || !n_launch)))
        return FJ_ERR_ARG;
    if (n_batch == 0) return 0;
    hipLaunchKernelGGL(fj_stream_launches_k, dim3(1), dim3(256), 0, s, n_batch, n, batch_mag, row_off,
                       launches, n_launch);
    return (int)hipGetLastError();
}

extern "C" long long fj_stream_workspace_bytes(int n_pair, int n_prim) {
    if (n_pair < 0 || n_prim < 0) return -1;
    size_t a = 0, b = 0, c = 0, d = 0;
    const int m = n_pair > 0 ? n_pair : 1;
    hipcub::DeviceRadixSort::SortPairsDescending(nullptr, a, (const double*)nullptr, (double*)nullptr,
                                                 (const int*)nullptr, (int*)nullptr, m);
    hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const unsigned*)nullptr, (unsigned*)nullptr,
                                       (const int*)nullptr, (int*)nullptr, m);
    hipcub::DeviceScan::ExclusiveSum(nullptr, c, (const int*)nullptr, (int*)nullptr, m);
    long long tmp = (long long)(a > b ? a : b);
    if ((long long)c > tmp) tmp = (long long)c;
    (void)d;
    // sort scratch + key/bin copies + flags/positions, 256-byte aligned pieces
    const long long n8 = ((long long)m * 8 + 255) / 256 * 256, n4 = ((long long)m * 4 + 255) / 256 * 256;
    return (tmp + 255) / 256 * 256 + 2 * n8 + 6 * n4;
}

extern "C" int fj_stream_plan(hipStream_t s, const FjStreamIn* in, FjStreamOut* out, void* ws,
                              long long ws_bytes) {
    if (!in || !out) return FJ_ERR_ARG;
    const int np = in->n_pair, nk = in->n_prim, n = in->n;
    if (np < 1 || nk < 1 || n < 1 || !(in->tau >= 0.0)) return FJ_ERR_ARG;
    if (ws_bytes < fj_stream_workspace_bytes(np, nk) || !ws) return FJ_ERR_WORKSPACE;
    char* w = static_cast<char*>(ws);
    size_t a = 0, b = 0, c = 0;
    hipcub::DeviceRadixSort::SortPairsDescending(nullptr, a, (const double*)nullptr, (double*)nullptr,
                                                 (const int*)nullptr, (int*)nullptr, np);
    hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const unsigned*)nullptr, (unsigned*)nullptr,
                                       (const int*)nullptr, (int*)nullptr, np);
    hipcub::DeviceScan::ExclusiveSum(nullptr, c, (const int*)nullptr, (int*)nullptr, np);
    size_t tmp = a > b ? a : b;
    if (c > tmp) tmp = c;
    void* cub = w;
    w += (tmp + 255) / 256 * 256;
    const size_t n8 = ((size_t)np * 8 + 255) / 256 * 256, n4 = ((size_t)np * 4 + 255) / 256 * 256;
    double* key_s = reinterpret_cast<double*>(w); w += n8;
    double* key_c = reinterpret_cast<double*>(w); w += n8;
    int* idx = reinterpret_cast<int*>(w); w += n4;
    int* idx_s = reinterpret_cast<int*>(w); w += n4;
    unsigned* bin_g = reinterpret_cast<unsigned*>(w); w += n4;
    unsigned* bin_s = reinterpret_cast<unsigned*>(w); w += n4;
    int* flag = reinterpret_cast<int*>(w); w += n4;
    int* pos = reinterpret_cast<int*>(w); w += n4;

    const int T = 256;
    hipLaunchKernelGGL(fj_zero_u64_k, dim3(1), dim3(64), 0, s, 1LL, out->gmax_bits);
    hipError_t e;
    hipLaunchKernelGGL(fj_stream_bounds_k, dim3((nk + T - 1) / T), dim3(T), 0, s, nk, np, in->pair_ab,
                       in->prim_off, in->shell_l, in->centre, in->nprim, in->pstart, in->exps, in->coefs,
                       out->g);
    hipLaunchKernelGGL(fj_stream_sums_k, dim3((np + T - 1) / T), dim3(T), 0, s, np, in->prim_off, out->g,
                       out->G, out->gmax_bits);
    hipLaunchKernelGGL(fj_stream_prune_k, dim3((np + T - 1) / T), dim3(T), 0, s, np, in->pair_ab,
                       in->shell_l, in->prim_off, out->g, out->G, out->gmax_bits, in->tau, n, in->mags,
                       in->Dsh, in->n_sh, in->ket_break, out->degree, out->keep, out->dmax, out->key,
                       out->mag, bin_g, out->eps);
    // pair order 0..np-1, then key descending (stable), then bin ascending (stable)
    hipLaunchKernelGGL(fj_iota_k, dim3((np + T - 1) / T), dim3(T), 0, s, np, idx);
    e = hipcub::DeviceRadixSort::SortPairsDescending(cub, a, out->key, key_s, idx, idx_s, np, 0, 64, s);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(fj_gather_u32_k, dim3((np + T - 1) / T), dim3(T), 0, s, np, idx_s, bin_g, bin_s);
    e = hipcub::DeviceRadixSort::SortPairs(cub, b, bin_s, reinterpret_cast<unsigned*>(key_c), idx_s,
                                           out->order, np, 0, 32, s);
    if (e != hipSuccess) return (int)e;
    unsigned* bin_sorted = reinterpret_cast<unsigned*>(key_c);
    hipLaunchKernelGGL(fj_stream_flags_k, dim3((np + T - 1) / T), dim3(T), 0, s, np, bin_sorted, flag);
    e = hipcub::DeviceScan::ExclusiveSum(cub, c, flag, pos, np, s);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(fj_stream_count_k, dim3(1), dim3(64), 0, s, np, bin_sorted, flag, pos, out->n_batch,
                       out->n_kept);
    hipLaunchKernelGGL(fj_zero_u64_k, dim3((np + T - 1) / T), dim3(T), 0, s, (long long)np,
                       out->batch_gmax_bits);
    hipLaunchKernelGGL(fj_stream_batches_k, dim3((np + T - 1) / T), dim3(T), 0, s, np, bin_sorted, flag, pos,
                       out->order, out->G, out->batch_start, out->batch_mag, out->batch_gmax_bits, n);
    return (int)hipGetLastError();
}
