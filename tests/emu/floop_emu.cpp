// Emulate the loop-form fused F kernel (felizjueves_amd/csrc/fj_floop.h) on
// the CPU over walker launches (upstream ABI: no work list, workgroup
// (bra, ket block)) read from a binary file; writes F after every launch.
// argv[3] threads (default 64), one per lane, one workgroup at a time.
// TEST INFRASTRUCTURE.
#include "fj_device.h"   // the emulation header, copied under this name
#include KLOOP_HEADER
#include FLOOP_HEADER
#include <fstream>
#include <string>
#include <thread>
#include <vector>
thread_local uint3e threadIdx;
uint3e blockIdx, blockDim, gridDim;
std::barrier<>* g_bar;
std::barrier<>* g_wbar[4];
double g_mfa[2][256], g_mfb[2][256];
thread_local int g_mf_parity;
double g_shfl[256];
double* g_dyn_lds;

template <typename T> T* rd(std::ifstream& f) {
    long long n;
    f.read((char*)&n, 8);
    T* p = (T*)malloc(n * sizeof(T) + 1);   // exactly sized, so ASan sees an overrun
    f.read((char*)p, n * sizeof(T));
    return p;
}

int main(int argc, char** argv) {
    std::ifstream f(argv[1], std::ios::binary);
    int hdr[2];   // n_launch n_ao
    f.read((char*)hdr, sizeof hdr);
    const int n_ao = hdr[1], span = 64;
    // argv[4] "one": every class on the one-at-a-time kernel; "batched":
    // only the classes the launcher gives the batched kernel (fj_fbloop_plan:
    // B >= 2), the others skipped; else the launcher's choice for every
    // class. Which launches ran goes to argv[2] + ".ran" (int32 per launch).
    const std::string mode = argc > 4 ? argv[4] : "auto";
    const bool one_only = mode == "one", batched_only = mode == "batched";
    // argv[5]: the largest batch (default the launcher's), so that a small
    // basis still has full batches, remainders and several batches per ket
    const int bmax = argc > 5 ? atoi(argv[5]) : FJ_FB_BMAX;
    const int block = argc > 3 ? atoi(argv[3]) : 64;
    struct L { int c[4]; FjFArgs a; double* lds; int B; int run; };
    std::vector<L> Ls(hdr[0]);
    double* F = (double*)calloc((size_t)n_ao * n_ao, 8);
    for (auto& l : Ls) {
        int h[11];   // la lb lc ld n_bra n_ket k_ab k_cd n_sh n_ao diag
        f.read((char*)h, sizeof h);
        double tau;
        f.read((char*)&tau, 8);
        for (int i = 0; i < 4; ++i) l.c[i] = h[i];
        // the launch's dynamic LDS at its exact size, so ASan sees an overrun
        const FjFbLds fb = fj_fbloop_plan(h[0], h[1], h[2], h[3], bmax);
        l.B = one_only ? 0 : (fb.B >= 2 ? fb.B : 0);
        l.run = !batched_only || l.B > 0;
        l.lds = (double*)malloc(l.B ? fb.bytes : fj_floop_lds(h[0], h[1], h[2], h[3]).bytes);
        double* bp = rd<double>(f); double* AB = h[1] ? rd<double>(f) : nullptr;
        double* Gab = rd<double>(f); int* bsh = rd<int>(f); int* bao = rd<int>(f);
        double* kp = rd<double>(f); double* CD = h[3] ? rd<double>(f) : nullptr;
        double* Gcd = rd<double>(f); int* ksh = rd<int>(f); int* kao = rd<int>(f);
        double* Dsh = rd<double>(f); double* D = rd<double>(f);
        FjFArgs& a = l.a;
        a.bra = FjSide{bp, nullptr, h[6], AB, Gab, bsh, bao, h[4]};
        a.ket = FjSide{kp, nullptr, h[7], CD, Gcd, ksh, kao, h[5]};
        a.work = nullptr; a.ket_list = nullptr; a.sys = nullptr;
        a.one = FjSys{D, Dsh, F, n_ao, h[8]};
        a.diag = h[10]; a.tau = tau;
        a.ws = nullptr; a.ws_doubles = 0; a.n_fp64 = 0; a.tier = nullptr;
        a.n_bx = h[4]; a.n_items = h[4] * ((h[5] + span - 1) / span);
    }
    const int threads = block;
    std::barrier<> bar(threads);
    g_bar = &bar;
    // one barrier per wavefront, for the matrix-core steps
    std::barrier<> wb0(64), wb1(64), wb2(64), wb3(64);
    g_wbar[0] = &wb0; g_wbar[1] = &wb1; g_wbar[2] = &wb2; g_wbar[3] = &wb3;
    if (threads % 64 || threads > 256) { fprintf(stderr, "threads %d\n", threads); return 1; }
    blockDim = {(unsigned)threads, 1, 1};
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t)
        ts.emplace_back([&, t] {
            threadIdx = {(unsigned)t, 0, 0};
            for (size_t li = 0; li < Ls.size(); ++li) {
                auto& l = Ls[li];
                if (!l.run) continue;
                if (t == 0 && getenv("FJ_EMU_PROGRESS"))
                    fprintf(stderr, "launch %zu class %d%d%d%d B %d\n", li, l.c[0], l.c[1], l.c[2], l.c[3], l.B);
                const unsigned gx = l.a.bra.n, gy = (l.a.ket.n + span - 1) / span;
                for (unsigned y = 0; y < gy; ++y)
                    for (unsigned x = 0; x < gx; ++x) {
                        if (t == 0) {
                            blockIdx = {x, y, 0};
                            gridDim = {gx, gy, 1};
                            g_dyn_lds = l.lds;
                        }
                        bar.arrive_and_wait();
                        if (l.B)
                            fj_fbloop_k(l.c[0], l.c[1], l.c[2], l.c[3], span, l.B, l.a);
                        else
                            fj_floop_k(l.c[0], l.c[1], l.c[2], l.c[3], span, l.a);
                        bar.arrive_and_wait();
                    }
            }
        });
    for (auto& t : ts) t.join();
    std::ofstream o(argv[2], std::ios::binary);
    o.write((char*)F, 8LL * n_ao * n_ao);
    std::ofstream r(std::string(argv[2]) + ".ran", std::ios::binary);
    for (auto& l : Ls) r.write((char*)&l.run, sizeof(int));
    fprintf(stderr, "done\n");
    return 0;
}
