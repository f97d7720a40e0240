// fj_plan.hip -- the native Fock builder's planner for one class pair (host
// code): the six-block screen per (bra, ket) pair of the same system, the
// FP32/FP64 tier split of the bound rule, the primitive-loop split into
// lanes and work items, and the costliest-first order of the items.
//
// It produces exactly what felizjueves_amd/fock.py's Python planner produces
// (same screen, same rule, same lane encoding and order), for builds too
// large to plan in Python (the many-molecule batch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/fj_mi355x.h"

namespace {

constexpr size_t WAVE = 64;              // lanes of a wavefront
constexpr long long FILL_MIN_PRIMS = 4;  // halve a lane to fill idle lanes from this many pairs

struct Tier {
    std::vector<int> work;      // int4 per item
    std::vector<long long> cost;
    std::vector<int> kets;
    double n_quartets = 0, n_prim = 0, flops = 0;
    long long n_cold = 0;
};

inline double dsh(const double* const* Dsh, const int* n_sh, int m, int i, int j) {
    return Dsh[m][(size_t)i * n_sh[m] + j];
}

}  // namespace

extern "C" int fj_plan_pairs(const FjPlanSide* bra, const FjPlanSide* ket, const int* mol_lo,
                             const int* mol_hi, const double* const* Dsh, const int* n_sh,
                             int diag, double tau, int mixed, double fp32_error, double fp32_bound,
                             double est_scale, int threads, double lane_flops, double per_prim,
                             double per_quartet, int max_prim, FjPlanResult* out) {
    if (!bra || !ket || !mol_lo || !mol_hi || !Dsh || !n_sh || !out) return FJ_ERR_ARG;
    if (threads <= 0 || threads % 64 || threads > 256) return FJ_ERR_THREADS;
    std::memset(out, 0, sizeof(*out));
    // bra pairs in contiguous chunks, processed by a few host threads, the
    // chunks' results joined in bra order (the same plan as one thread)
    const int n_chunk = std::max(1, std::min(bra->n, 256));
    std::vector<Tier> part(2 * (size_t)n_chunk);
    auto run_chunk = [&](int ci) {
        const int b0 = (int)((long long)bra->n * ci / n_chunk);
        const int b1 = (int)((long long)bra->n * (ci + 1) / n_chunk);
        Tier* tier = &part[2 * (size_t)ci];
        std::vector<int> kl[3];
        std::vector<long long> cnt;
        std::vector<int> order, lk, llo;
        // stable sort of the lanes by primitive count, descending
        auto sort_lanes = [&]() {
            order.resize(lk.size());
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cnt[x] > cnt[y]; });
            std::vector<int> k2(lk.size()), l2(lk.size());
            std::vector<long long> c2(lk.size());
            for (size_t i = 0; i < order.size(); ++i) {
                k2[i] = lk[order[i]];
                l2[i] = llo[order[i]];
                c2[i] = cnt[order[i]];
            }
            lk.swap(k2);
            llo.swap(l2);
            cnt.swap(c2);
        };
        for (int b = b0; b < b1; ++b) {
            const int m = bra->mol[b];
            const int lo = mol_lo[m], hi = mol_hi[m];
            const int sa = bra->sh[2 * b], sb = bra->sh[2 * b + 1];
            const double Gb = bra->G[b];
            const double dab = 2.0 * dsh(Dsh, n_sh, m, sa, sb);
            kl[0].clear();
            kl[1].clear();
            kl[2].clear();
            const int k_first = diag ? std::max(lo, b) : lo;
            for (int k = k_first; k < hi; ++k) {
                // the density-free screen decides which quartets the plan
                // covers; the planning density only sorts them (cold: the
                // ones its six-block screen drops)
                if (!(Gb * ket->G[k] >= tau)) continue;
                const int sc = ket->sh[2 * k], sd = ket->sh[2 * k + 1];
                double dm = dab;
                dm = std::max(dm, 2.0 * dsh(Dsh, n_sh, m, sc, sd));
                dm = std::max(dm, 0.5 * dsh(Dsh, n_sh, m, sa, sc));
                dm = std::max(dm, 0.5 * dsh(Dsh, n_sh, m, sa, sd));
                dm = std::max(dm, 0.5 * dsh(Dsh, n_sh, m, sb, sc));
                dm = std::max(dm, 0.5 * dsh(Dsh, n_sh, m, sb, sd));
                const double bound = Gb * (ket->G[k] * dm);
                if (!(bound >= tau)) {
                    kl[2].push_back(k);
                    continue;
                }
                int t = 0;
                if (mixed) {
                    if (fp32_bound >= 0.0) {
                        t = bound < fp32_bound;
                    } else {
                        const double est = bound * bra->amp[b] * ket->amp[k] * est_scale;
                        t = est < fp32_error;
                    }
                }
                kl[t].push_back(k);
            }
            const int kab = bra->deg[b];
            for (int t = 0; t < 2; ++t) {
                if (kl[t].empty()) continue;
                Tier& T = tier[t];
                long long sum_kcd = 0;
                for (int k : kl[t]) sum_kcd += ket->deg[k];
                T.n_quartets += (double)kl[t].size();
                T.n_prim += (double)kab * (double)sum_kcd;
                T.flops += per_prim * (double)kab * (double)sum_kcd + per_quartet * (double)kl[t].size();
                if (lane_flops <= 0.0) continue;   // counting pass
                // bra primitive pairs per work item (cb), ket primitive pairs per lane (ck)
                const int cb = (kab * per_prim <= lane_flops)
                                   ? kab : std::max(1, (int)std::floor(lane_flops / per_prim));
                const double ckd = std::floor(lane_flops / (cb * per_prim));
                const int ck = std::max(1, (int)std::min((double)max_prim, ckd));
                // lanes (ket, first primitive pair, count), equal chunks per ket
                lk.clear();
                llo.clear();
                cnt.clear();
                for (int k : kl[t]) {
                    const int kcd = ket->deg[k];
                    const int n_ch = (kcd + ck - 1) / ck;
                    const int size = (kcd + n_ch - 1) / n_ch;
                    for (int j = 0; j < n_ch; ++j) {
                        const int lo_p = j * size;
                        lk.push_back(k);
                        llo.push_back(lo_p);
                        cnt.push_back(std::min(size, kcd - lo_p));
                    }
                }
                sort_lanes();
                // fill the last wavefront's idle lanes: halve the (first)
                // lane with the most primitive pairs while it has at least
                // FILL_MIN_PRIMS, ceil half in place, the rest appended
                const size_t target = (lk.size() + WAVE - 1) / WAVE * WAVE;
                if (target != lk.size()) {
                    while (lk.size() < target) {
                        size_t i = 0;
                        for (size_t x = 1; x < cnt.size(); ++x)
                            if (cnt[x] > cnt[i]) i = x;
                        const long long c = cnt[i];
                        if (c < FILL_MIN_PRIMS) break;
                        const long long h = (c + 1) / 2;
                        cnt[i] = h;
                        lk.push_back(lk[i]);
                        llo.push_back(llo[i] + (int)h);
                        cnt.push_back(c - h);
                    }
                    sort_lanes();
                }
                const long long st = (long long)T.kets.size();
                for (size_t i = 0; i < lk.size(); ++i) {
                    const int k = lk[i], lo_p = llo[i];
                    const int kcd = ket->deg[k];
                    long long v = k;
                    if (!(lo_p == 0 && cnt[i] == kcd)) {
                        const long long enc = (lo_p + cnt[i] >= kcd) ? 0 : cnt[i];
                        v |= ((long long)lo_p << 17) | (enc << 24);
                    }
                    T.kets.push_back((int)v);
                }
                for (size_t t0 = 0; t0 < lk.size(); t0 += threads) {
                    const size_t t1 = std::min(lk.size(), t0 + (size_t)threads);
                    const long long pmax = cnt[t0];
                    for (int blo = 0; blo < kab; blo += cb) {
                        const int bn = (cb >= kab) ? 0 : std::min(cb, kab - blo);
                        T.work.push_back(b);
                        T.work.push_back((int)(st + t0));
                        T.work.push_back(m | (blo << 16) | (bn << 24));
                        T.work.push_back((int)(st + t1));
                        T.cost.push_back((long long)std::min(cb, kab - blo) * pmax);
                    }
                }
            }
            // the cold kets: FP64 items of whole kets and the whole bra, no
            // cost (listed last) and no counted work
            if (!kl[2].empty() && lane_flops > 0.0) {
                Tier& T = tier[0];
                const long long st = (long long)T.kets.size();
                const size_t n = kl[2].size();
                for (int k : kl[2]) T.kets.push_back(k);
                for (size_t t0 = 0; t0 < n; t0 += threads) {
                    const size_t t1 = std::min(n, t0 + (size_t)threads);
                    T.work.push_back(b);
                    T.work.push_back((int)(st + t0));
                    T.work.push_back(m);
                    T.work.push_back((int)(st + t1));
                    T.cost.push_back(-1);   // below every counted item: listed last
                    ++T.n_cold;
                }
            }
        }
    };
    std::atomic<int> next{0};
    auto worker = [&]() {
        for (int ci = next++; ci < n_chunk; ci = next++) run_chunk(ci);
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < hw; ++i) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    Tier tier[2];
    for (int ci = 0; ci < n_chunk; ++ci) {
        for (int t = 0; t < 2; ++t) {
            Tier& src = part[2 * (size_t)ci + t];
            Tier& dst = tier[t];
            const int base = (int)dst.kets.size();
            for (size_t i = 0; i < src.work.size(); i += 4) {
                dst.work.push_back(src.work[i]);
                dst.work.push_back(src.work[i + 1] + base);
                dst.work.push_back(src.work[i + 2]);
                dst.work.push_back(src.work[i + 3] + base);
            }
            dst.cost.insert(dst.cost.end(), src.cost.begin(), src.cost.end());
            dst.kets.insert(dst.kets.end(), src.kets.begin(), src.kets.end());
            dst.n_quartets += src.n_quartets;
            dst.n_prim += src.n_prim;
            dst.flops += src.flops;
            dst.n_cold += src.n_cold;
        }
    }
    for (int t = 0; t < 2; ++t) {
        Tier& T = tier[t];
        out->n_quartets[t] = T.n_quartets;
        out->n_prim[t] = T.n_prim;
        out->flops[t] = T.flops;
        const long long n = (long long)T.cost.size();
        out->n_work[t] = n;
        out->n_kets[t] = (long long)T.kets.size();
        if (t == 0) out->n_cold = T.n_cold;
        if (lane_flops <= 0.0 || n == 0) continue;
        // costliest work items first (stable)
        std::vector<long long> idx(n);
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](long long x, long long y) { return T.cost[x] > T.cost[y]; });
        int* w = static_cast<int*>(std::malloc(sizeof(int) * 4 * n));
        int* k = static_cast<int*>(std::malloc(sizeof(int) * std::max<size_t>(1, T.kets.size())));
        if (!w || !k) {
            std::free(w);
            std::free(k);
            fj_plan_free(out);
            return FJ_ERR_ARG;
        }
        for (long long i = 0; i < n; ++i) std::memcpy(w + 4 * i, T.work.data() + 4 * idx[i], sizeof(int) * 4);
        std::memcpy(k, T.kets.data(), sizeof(int) * T.kets.size());
        out->work[t] = w;
        out->kets[t] = k;
    }
    return 0;
}

extern "C" void fj_plan_free(FjPlanResult* out) {
    if (!out) return;
    for (int t = 0; t < 2; ++t) {
        std::free(out->work[t]);
        std::free(out->kets[t]);
        out->work[t] = nullptr;
        out->kets[t] = nullptr;
    }
}
