// Emulate the fused-gradient kernel (felizjueves_amd/csrc/fj_gfloop.h) on
// the CPU over a batch of launches read from a binary file (one std::thread
// per lane, THREADS lanes per workgroup, one workgroup at a time, a grid of
// at most five workgroups so that each takes several quartets); writes G.
// TEST INFRASTRUCTURE.
#include "fj_device.h"   // the emulation header, copied under this name
#include GFLOOP_HEADER
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
    int h[12];   // la lb lc ld n_bra n_ket k_ab k_cd n_atom n_ao diag chunk
    double *bp, *AB, *kp, *CD, *lds;
    int *bs, *bao, *bat, *ks, *kao, *kat;
    long long n_q, grid;
};

int main(int argc, char** argv) {
    std::ifstream f(argv[1], std::ios::binary);
    const int threads = atoi(argv[3]);
    int n_launch, n_atom, n_ao;
    f.read((char*)&n_launch, 4);
    f.read((char*)&n_atom, 4);
    f.read((char*)&n_ao, 4);
    double* D = rd<double>(f);
    double* G = (double*)calloc((size_t)n_atom * 3 + 1, sizeof(double));
    std::vector<Launch> Ls(n_launch);
    for (auto& L : Ls) {
        f.read((char*)L.h, sizeof L.h);
        L.bp = rd<double>(f); L.AB = rd<double>(f); L.bs = rd<int>(f); L.bao = rd<int>(f);
        L.bat = rd<int>(f);
        L.kp = rd<double>(f); L.CD = rd<double>(f); L.ks = rd<int>(f); L.kao = rd<int>(f);
        L.kat = rd<int>(f);
        L.n_q = L.h[10] ? (long long)L.h[4] * (L.h[4] + 1) / 2 : (long long)L.h[4] * L.h[5];
        L.n_q *= (L.h[6] * L.h[7] + L.h[11] - 1) / L.h[11];   // work items
        L.grid = L.n_q < 5 ? L.n_q : 5;
        // the launch's dynamic LDS at its exact size, so ASan sees an overrun
        L.lds = (double*)malloc(fj_gf_plan(L.h[0], L.h[1], L.h[2], L.h[3], L.h[6] * L.h[7], threads).bytes);
    }
    std::barrier<> bar(threads);
    g_bar = &bar;
    blockDim = {(unsigned)threads, 1, 1};
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t)
        ts.emplace_back([&, t] {
            threadIdx = {(unsigned)t, 0, 0};
            for (auto& L : Ls) {
                const FjGfLds o = fj_gf_plan(L.h[0], L.h[1], L.h[2], L.h[3], L.h[6] * L.h[7], threads);
                FjGfArgs g{L.h[0], L.h[1], L.h[2], L.h[3], L.h[4], L.h[5], L.h[6], L.h[7], L.h[8],
                           L.h[9], L.h[10], L.h[11], L.bp, L.AB, L.bs, L.bao, L.bat, L.kp, L.CD, L.ks, L.kao,
                           L.kat, D, G};
                for (long long b = 0; b < L.grid; ++b) {
                    if (t == 0) {
                        blockIdx = {(unsigned)b, 0, 0};
                        gridDim = {(unsigned)L.grid, 1, 1};
                        g_dyn_lds = L.lds;
                    }
                    bar.arrive_and_wait();
                    fj_gfloop_k(g, o);
                    bar.arrive_and_wait();
                }
            }
        });
    for (auto& t : ts) t.join();
    std::ofstream out(argv[2], std::ios::binary);
    out.write((char*)G, 8LL * n_atom * 3);
    fprintf(stderr, "done\n");
    return 0;
}
