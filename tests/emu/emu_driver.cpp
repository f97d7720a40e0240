// Emulate fj_f_<CLS> (upstream ABI, walker path: one workgroup per (bra,
// ket block), persistent over the workspace slots) on the CPU for one launch
// read from a binary file; writes F. TEST INFRASTRUCTURE.
#include "fj_device.h"   // the emulation header, copied under this name
#include <thread>
#include <vector>
#include <fstream>
#include <string>
thread_local uint3e threadIdx;
uint3e blockIdx, blockDim, gridDim;
std::barrier<>* g_bar;
double g_shfl[256];
double* g_dyn_lds;
#include CLS_HEADER

template <typename T> std::vector<T> rd(std::ifstream& f) {
    long long n; f.read((char*)&n, 8); std::vector<T> v(n); f.read((char*)v.data(), n * sizeof(T)); return v;
}
int main(int argc, char** argv) {
    std::ifstream f(argv[1], std::ios::binary);
    int hdr[7]; f.read((char*)hdr, sizeof hdr); double tau; f.read((char*)&tau, 8);
    int n_bra = hdr[0], n_ket = hdr[1], k_ab = hdr[2], k_cd = hdr[3], n_sh = hdr[4], n_ao = hdr[5], diag = hdr[6];
    auto bra_prim = rd<double>(f); std::vector<double> AB; if (HAS_AB) AB = rd<double>(f);
    auto G_ab = rd<double>(f); auto bra_sh = rd<int>(f); auto bra_ao = rd<int>(f);
    auto ket_prim = rd<double>(f); std::vector<double> CD; if (HAS_CD) CD = rd<double>(f);
    auto G_cd = rd<double>(f); auto ket_sh = rd<int>(f); auto ket_ao = rd<int>(f);
    auto Dsh = rd<double>(f); auto D = rd<double>(f);
    // heap copies sized exactly, so ASan sees any overrun
    auto H = [](auto& v) { auto* p = (typename std::remove_reference_t<decltype(v)>::value_type*)malloc(v.size() * sizeof(v[0]) + 1); std::copy(v.begin(), v.end(), p); return p; };
    double* F = (double*)calloc((size_t)n_ao * n_ao, 8);
    const int threads = 64;
    FjFArgs a;
    a.bra = FjSide{H(bra_prim), nullptr, k_ab, HAS_AB ? H(AB) : nullptr, H(G_ab), H(bra_sh), H(bra_ao), n_bra};
    a.ket = FjSide{H(ket_prim), nullptr, k_cd, HAS_CD ? H(CD) : nullptr, H(G_cd), H(ket_sh), H(ket_ao), n_ket};
    a.work = nullptr; a.ket_list = nullptr; a.sys = nullptr; a.tier = nullptr;
    a.one = FjSys{H(D), H(Dsh), F, n_ao, n_sh};
    a.diag = diag; a.tau = tau;
    const long long W = WS_DOUBLES;
    unsigned gx = n_bra, gy = (n_ket + threads - 1) / threads;
    a.n_bx = gx; a.n_items = gx * gy;
    std::vector<std::pair<unsigned, unsigned>> blocks_xy;
    if (W > 0) {
        // the launcher's rule: the library workspace holds FJ_LIB_WS_BLOCKS
        // workgroups' slots; the persistent grid is cut to min(that, items)
        const long long need = 8 * W * threads * 256;
        a.ws = (double*)malloc(need); a.ws_doubles = need / 8;
        long long blocks = a.ws_doubles / W / threads;
        unsigned grid = (unsigned)std::min<long long>(blocks, a.n_items);
        gridDim = {grid, 1, 1};
        for (unsigned b = 0; b < grid; ++b) blocks_xy.push_back({b, 0});
    } else {
        a.ws = nullptr; a.ws_doubles = 0;
        gridDim = {gx, gy, 1};
        for (unsigned y = 0; y < gy; ++y)
            for (unsigned x = 0; x < gx; ++x) blocks_xy.push_back({x, y});
    }
    blockDim = {(unsigned)threads, 1, 1};
    for (auto [bx, by] : blocks_xy) {
        blockIdx = {bx, by, 0};
        std::barrier<> bar(threads); g_bar = &bar;
        std::vector<std::thread> ts;
        for (int t = 0; t < threads; ++t) ts.emplace_back([&, t] { threadIdx = {(unsigned)t, 0, 0}; KERNEL<double>(a); });
        for (auto& t : ts) t.join();
    }
    std::ofstream o(argv[2], std::ios::binary); o.write((char*)F, 8LL * n_ao * n_ao);
    fprintf(stderr, "done\n");
    return 0;
}
