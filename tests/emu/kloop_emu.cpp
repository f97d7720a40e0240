// Emulate the loop-form K kernel (felizjueves_amd/csrc/fj_kloop.h) on the
// CPU over a batch of launches read from a binary file (64 threads, one per
// lane, one workgroup at a time); writes every launch's K rows in order.
// TEST INFRASTRUCTURE.
#include "fj_device.h"   // the emulation header, copied under this name
#include KLOOP_HEADER
#include <fstream>
#include <thread>
#include <vector>
thread_local uint3e threadIdx;
uint3e blockIdx, blockDim, gridDim;
std::barrier<>* g_bar;
double g_shfl[256];
double* g_dyn_lds;

template <typename T> T* rd(std::ifstream& f) {
    long long n;
    f.read((char*)&n, 8);
    // exactly sized heap copies, so AddressSanitizer sees any overrun
    T* p = (T*)malloc(n * sizeof(T) + 1);
    f.read((char*)p, n * sizeof(T));
    return p;
}

struct Launch {
    int h[7];   // la lc lb ld n_bra n_ket n_dc
    double tau, maxD;
    double *P, *z, *XA, *XC, *s, *Ip, *Q, *eta, *XB, *XD, *sb, *Iq, *Dcd, *K, *lds;
    int *dc, *dd;
};

int main(int argc, char** argv) {
    std::ifstream f(argv[1], std::ios::binary);
    int n_launch;
    f.read((char*)&n_launch, 4);
    std::vector<Launch> Ls(n_launch);
    for (auto& L : Ls) {
        f.read((char*)L.h, sizeof L.h);
        f.read((char*)&L.tau, 8);
        f.read((char*)&L.maxD, 8);
        L.P = rd<double>(f); L.z = rd<double>(f); L.XA = rd<double>(f); L.XC = rd<double>(f);
        L.s = rd<double>(f); L.Ip = rd<double>(f); L.dc = rd<int>(f);
        L.Q = rd<double>(f); L.eta = rd<double>(f); L.XB = rd<double>(f); L.XD = rd<double>(f);
        L.sb = rd<double>(f); L.Iq = rd<double>(f); L.dd = rd<int>(f); L.Dcd = rd<double>(f);
        const int na = (L.h[0] + 1) * (L.h[0] + 2) / 2, nb = (L.h[2] + 1) * (L.h[2] + 2) / 2;
        L.K = (double*)malloc(sizeof(double) * (size_t)L.h[4] * na * nb + 1);
        // the launch's dynamic LDS at its exact size, so ASan sees an overrun
        L.lds = (double*)malloc(fj_kloop_lds(L.h[0], L.h[1], L.h[2], L.h[3]).bytes);
    }
    const int threads = argc > 3 ? atoi(argv[3]) : 64;
    // argv[4]: ket slices per bra (gridDim.y; default 1); with more than
    // one the slices add into K, cleared first, as the launcher does
    const int split = argc > 4 ? atoi(argv[4]) : 1;
    for (auto& L : Ls) {
        const int na = (L.h[0] + 1) * (L.h[0] + 2) / 2, nb = (L.h[2] + 1) * (L.h[2] + 2) / 2;
        for (size_t i = 0; i < (size_t)L.h[4] * na * nb; ++i) L.K[i] = 0.0;
    }
    std::barrier<> bar(threads);
    g_bar = &bar;
    blockDim = {(unsigned)threads, 1, 1};
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t)
        ts.emplace_back([&, t] {
            threadIdx = {(unsigned)t, 0, 0};
            for (auto& L : Ls)
                for (int y = 0; y < split; ++y)
                    for (int b = 0; b < L.h[4]; ++b) {
                        if (t == 0) {
                            blockIdx = {(unsigned)b, (unsigned)y, 0};
                            gridDim = {(unsigned)L.h[4], (unsigned)split, 1};
                            g_dyn_lds = L.lds;
                        }
                        bar.arrive_and_wait();
                        fj_kloop_k(L.h[0], L.h[1], L.h[2], L.h[3], L.h[4], L.h[5], L.h[6], L.tau, L.maxD,
                                   L.P, L.z, L.XA, L.XC, L.s, L.Ip, L.dc, L.Q, L.eta, L.XB, L.XD, L.sb,
                                   L.Iq, L.dd, L.Dcd, L.K);
                        bar.arrive_and_wait();
                    }
        });
    for (auto& t : ts) t.join();
    std::ofstream o(argv[2], std::ios::binary);
    for (auto& L : Ls) {
        const int na = (L.h[0] + 1) * (L.h[0] + 2) / 2, nb = (L.h[2] + 1) * (L.h[2] + 2) / 2;
        o.write((char*)L.K, 8LL * L.h[4] * na * nb);
    }
    fprintf(stderr, "done\n");
    return 0;
}
