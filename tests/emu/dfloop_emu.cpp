// Emulate the density-fitting kernel (felizjueves_amd/csrc/fj_dfloop.h) on
// the CPU over a batch of class launches read from a binary file (one
// std::thread per lane, THREADS lanes per workgroup, a grid of at most five
// workgroups so that each takes several cells); writes the output array.
// TEST INFRASTRUCTURE.
#include "fj_device.h"   // the emulation header, copied under this name
#include DFLOOP_HEADER
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
    int h[5];   // la lb lc n_cell max_rp
    int *b0, *b1, *k0, *k1, *oa, *ob, *oc;
    double* lds;
    FjDfLds o;
};

int main(int argc, char** argv) {
    std::ifstream f(argv[1], std::ios::binary);
    const int threads = atoi(argv[3]);
    int hdr[4];   // mode n_ao n_aux n_cls
    f.read((char*)hdr, sizeof hdr);
    double* zeta = rd<double>(f);
    double* P = rd<double>(f);
    double* XPA = rd<double>(f);
    double* XPB = rd<double>(f);
    double* sb = rd<double>(f);
    double* A = rd<double>(f);
    double* alpha = rd<double>(f);
    double* sa = rd<double>(f);
    const size_t n_out = hdr[0] == 0 ? (size_t)hdr[1] * hdr[1] * hdr[2] : (size_t)hdr[2] * hdr[2];
    double* out = (double*)calloc(n_out + 1, sizeof(double));
    std::vector<Cls> Cs(hdr[3]);
    for (auto& c : Cs) {
        f.read((char*)c.h, sizeof c.h);
        c.b0 = rd<int>(f); c.b1 = rd<int>(f); c.k0 = rd<int>(f); c.k1 = rd<int>(f);
        c.oa = rd<int>(f); c.ob = rd<int>(f); c.oc = rd<int>(f);
        c.o = fj_df_plan(c.h[0], c.h[1], c.h[2], c.h[4]);
        // the launch's dynamic LDS at its exact size, so ASan sees an overrun
        c.lds = (double*)malloc(c.o.bytes);
    }
    std::barrier<> bar(threads);
    g_bar = &bar;
    blockDim = {(unsigned)threads, 1, 1};
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t)
        ts.emplace_back([&, t] {
            threadIdx = {(unsigned)t, 0, 0};
            for (auto& c : Cs) {
                const int grid = c.h[3] < 5 ? c.h[3] : 5;
                FjDfArgs g{c.h[0], c.h[1], c.h[2], c.h[3],
                           zeta, P, hdr[0] == 0 ? XPA : nullptr, hdr[0] == 0 ? XPB : nullptr, sb,
                           A, alpha, sa, c.b0, c.b1, c.k0, c.k1, c.oa, hdr[0] == 0 ? c.ob : nullptr,
                           c.oc, hdr[0], hdr[1], hdr[2], out};
                for (int b = 0; b < grid; ++b) {
                    if (t == 0) {
                        blockIdx = {(unsigned)b, 0, 0};
                        gridDim = {(unsigned)grid, 1, 1};
                        g_dyn_lds = c.lds;
                    }
                    bar.arrive_and_wait();
                    fj_dfloop_k(g, c.o);
                    bar.arrive_and_wait();
                }
            }
        });
    for (auto& t : ts) t.join();
    std::ofstream o(argv[2], std::ios::binary);
    o.write((char*)out, 8LL * n_out);
    fprintf(stderr, "done\n");
    return 0;
}
