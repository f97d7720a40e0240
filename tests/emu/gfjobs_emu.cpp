// Emulate the fused-gradient kernel's merged launches (fj_gfjobs_k,
// felizjueves_amd/csrc/fj_gfloop.h) on the CPU: one launch per class over a
// job table, the sides concatenated, read from a binary file (one
// std::thread per lane, THREADS lanes, a grid of at most seven workgroups);
// writes G. TEST INFRASTRUCTURE.
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

struct Cls {
    int cls[4], n_jobs, max_pq;
    long long n_q, grid;
    FjGfJob* jobs;
    double* lds;
};

int main(int argc, char** argv) {
    std::ifstream f(argv[1], std::ios::binary);
    const int threads = atoi(argv[3]);
    int n_cls, n_atom, n_ao;
    f.read((char*)&n_cls, 4);
    f.read((char*)&n_atom, 4);
    f.read((char*)&n_ao, 4);
    double* D = rd<double>(f);
    double* prim = rd<double>(f);
    double* sep = rd<double>(f);
    int* sh = rd<int>(f);
    int* ao = rd<int>(f);
    int* at = rd<int>(f);
    double* G = (double*)calloc((size_t)n_atom * 3 + 1, sizeof(double));
    std::vector<Cls> Cs(n_cls);
    for (auto& c : Cs) {
        f.read((char*)c.cls, sizeof c.cls);
        f.read((char*)&c.n_jobs, 4);
        f.read((char*)&c.max_pq, 4);
        f.read((char*)&c.n_q, 8);
        c.jobs = rd<FjGfJob>(f);
        c.grid = c.n_q < 7 ? c.n_q : 7;
        c.lds = (double*)malloc(fj_gf_plan(c.cls[0], c.cls[1], c.cls[2], c.cls[3], c.max_pq, threads).bytes);
    }
    std::barrier<> bar(threads);
    g_bar = &bar;
    blockDim = {(unsigned)threads, 1, 1};
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t)
        ts.emplace_back([&, t] {
            threadIdx = {(unsigned)t, 0, 0};
            for (auto& c : Cs) {
                const FjGfLds o = fj_gf_plan(c.cls[0], c.cls[1], c.cls[2], c.cls[3], c.max_pq, threads);
                FjGfArgs g{c.cls[0], c.cls[1], c.cls[2], c.cls[3], 0, 0, 0, 0, n_atom, n_ao, 0, 1,
                           prim, sep, sh, ao, at, prim, sep, sh, ao, at, D, G};
                for (long long b = 0; b < c.grid; ++b) {
                    if (t == 0) {
                        blockIdx = {(unsigned)b, 0, 0};
                        gridDim = {(unsigned)c.grid, 1, 1};
                        g_dyn_lds = c.lds;
                    }
                    bar.arrive_and_wait();
                    fj_gfjobs_k(g, o, c.jobs, c.n_jobs, c.n_q);
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
