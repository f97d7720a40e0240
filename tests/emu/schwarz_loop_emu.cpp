// Emulate the loop-form Schwarz kernel (felizjueves_amd/csrc/fj_kloop.h) on
// the CPU over the pair classes read from a binary file (64 threads, one per
// lane, one workgroup at a time); writes Q (n_ao x n_ao). TEST INFRASTRUCTURE.
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
    T* p = (T*)malloc(n * sizeof(T) + 1);   // exactly sized, so ASan sees an overrun
    f.read((char*)p, n * sizeof(T));
    return p;
}

struct Cls {
    int h[4];   // la lb n_pair n_rec
    double *z, *P, *XA, *XB, *s;
    int *rec0, *rec1, *oa, *ob;
};

int main(int argc, char** argv) {
    std::ifstream f(argv[1], std::ios::binary);
    int hdr[2];   // n_class n_ao
    f.read((char*)hdr, sizeof hdr);
    const int n_ao = hdr[1];
    std::vector<Cls> Cs(hdr[0]);
    for (auto& c : Cs) {
        f.read((char*)c.h, sizeof c.h);
        c.z = rd<double>(f); c.P = rd<double>(f); c.XA = rd<double>(f); c.XB = rd<double>(f);
        c.s = rd<double>(f); c.rec0 = rd<int>(f); c.rec1 = rd<int>(f); c.oa = rd<int>(f);
        c.ob = rd<int>(f);
    }
    double* Q = (double*)calloc((size_t)n_ao * n_ao, 8);
    const int threads = 64;
    std::barrier<> bar(threads);
    g_bar = &bar;
    blockDim = {(unsigned)threads, 1, 1};
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t)
        ts.emplace_back([&, t] {
            threadIdx = {(unsigned)t, 0, 0};
            for (auto& c : Cs)
                for (int b = 0; b < c.h[2]; ++b) {
                    if (t == 0) {
                        blockIdx = {(unsigned)b, 0, 0};
                        gridDim = {(unsigned)c.h[2], 1, 1};
                    }
                    bar.arrive_and_wait();
                    fj_schwarz_loop_k(c.h[0], c.h[1], c.h[2], n_ao, c.z, c.P, c.XA, c.XB, c.s, c.rec0,
                                      c.rec1, c.oa, c.ob, Q);
                    bar.arrive_and_wait();
                }
        });
    for (auto& t : ts) t.join();
    std::ofstream o(argv[2], std::ios::binary);
    o.write((char*)Q, 8LL * n_ao * n_ao);
    fprintf(stderr, "done\n");
    return 0;
}
