// fj_gfloop.h -- the fused two-electron nuclear gradient on the GPU:
// dE_2e/dR for E_2e = 1/2 tr(D G), G = J - K/2 (upstream
// gradient/kernel_gf.py:1-108, the fused F kernel differentiated), over the
// same stream of unique shell quartets as the fused F build, every class
// through (dd|dd) (the upstream emits gf_ through l = 2,
// emit/__main__.py:842).
//
// Per unique quartet (a b | c d) with the derived weights (kernel_gf.py:25-40)
//
//   w_abcd = deg (4 D_ab D_cd - D_ac D_bd - D_ad D_bc)
//
// and per centre X and axis k, from the derivative of a Cartesian Gaussian
// (kernel_gf.py:44-50), d/dA_k (ab|cd) = (a+1_k b|cd)^{2 alpha} - a_k
// (a-1_k b|cd): all four centres differentiated, no translational
// invariance used (so sum_A G[A] = 0 stays a check).
//
// The integrals are McMurchie-Davidson's, per primitive quartet
//   (ab|cd) = U_ab U_cd sum_t sum_u E^{ab}_t (-1)^|u| E^{cd}_u [t + u]
// so a centre's derivative replaces that side's E^{ab}_t by the E of the
// differentiated product, per axis 2 alpha E^{a+1,b}_t - a E^{a-1,b}_t (the
// two other axes unchanged). The weights are contracted into Hermite space
// first, once per primitive quartet:
//
//   Wk[ab][u] = (-1)^|u| sum_cd w_abcd E^{cd}_u,   |u| <= lc + ld
//   Wb[cd][t] =          sum_ab w_abcd E^{ab}_t,   |t| <= la + lb
//   gA_k, gB_k += U_ab U_cd sum_ab sum_{|t| <= la+lb+1} dE^{ab}_t sum_u [t+u] Wk[ab][u]
//   gC_k, gD_k += U_ab U_cd sum_cd sum_{|u| <= lc+ld+1} (-1)^|u| dE^{cd}_u sum_t [t+u] Wb[cd][t]
//
// A workgroup (threads = 64 .. 256 cooperating through LDS) takes unique
// quartets in turn (grid-stride) and a quartet's primitive quartets in
// batches of B: per batch the per-axis E tables to l + 1 on every centre and
// the Hermite tower to L + 1 of each of its B primitive quartets, the E
// products, the two weight GEMMs (Wk, Wb), the tower gathered into H[t][u],
// and the two contraction GEMMs (Y, X) with the differentiated E in their
// epilogue, lanes over (b, 4 x 4 register tile) -- L + 6 barriers per batch,
// not per primitive quartet, so a highly contracted low class fills the
// lanes and a (dd|dd) quartet (B = 1) has work for 256. B is the largest
// that fits the LDS budget and the lanes (fj_gf_plan). The twelve
// sums are reduced over the workgroup and added into G[atom][k] with twelve
// atomics per quartet.
// No screen (kernel_gf.py:82-85: the F screen bounds an integral, not its
// derivative): every unique quartet of a launch is visited.
//
// The includer provides fj_device.h (the device one, or the CPU emulation);
// the launcher and the per-class entries are in fj_gfloop.hip.
#pragma once
#define FJ_KLOOP_HELPERS_ONLY
#include "fj_kloop.h"

namespace {

constexpr int GF_LMAX = 2;    // through (dd|dd), the tower to L + 1 <= 9
constexpr int GF_REC = 10;    // P(3), X = P - A (3), zeta, U, e1, e2
constexpr int GF_MAXW = 4;    // waves per workgroup (threads <= 256)

struct FjGfArgs {
    int la, lb, lc, ld;
    int n_bra, n_ket, k_ab, k_cd, n_atom, n_ao, diag;
    int chunk;   // primitive quartets per work item (>= k_ab k_cd: whole quartets)
    const double* bra_prim;
    const double* AB;
    const int* bra_sh;
    const int* bra_ao;
    const int* bra_at;
    const double* ket_prim;
    const double* CD;
    const int* ket_sh;
    const int* ket_ao;
    const int* ket_at;
    const double* D;
    double* G;
};

// The class's launch plan: the batch B of primitive quartets, the byte
// offsets of the LDS arrays in one dynamic block (per-quartet doubles, B
// copies of the per-batch doubles, then ints and the u16 index table) and
// the total.
struct FjGfLds {
    int B;
    int w, red, R0, R1, eb, ek, wf, wk, wb, sc, un, hc, tw, pab, pcd, add, bytes;
    int axb, axk, nR, nu, nt, nu1, nt1, uw;
};

constexpr int GF_LDS_BUDGET = 64 * 1024;   // bytes per workgroup
constexpr int GF_TM = 4, GF_TN = 4;        // register tile of the Hermite-space GEMMs
constexpr int GF_BMAX = 64;                // primitive quartets per batch
constexpr int GF_SC = 8;                   // per primitive quartet: U U, 2 x4 exponents, PQ

__host__ __device__ inline int fj_gf_herm(int n) { return n * (n + 1) * (n + 2) / 6; }

__host__ __device__ inline FjGfLds fj_gf_layout(int la, int lb, int lc, int ld, int B) {
    const int L = la + lb + lc + ld, lab = la + lb, lcd = lc + ld;
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
    const int nc = (lc + 1) * (lc + 2) / 2, nd = (ld + 1) * (ld + 2) / 2;
    FjGfLds o;
    o.B = B;
    o.nR = fj_gf_herm(L + 2);      // |t| <= L + 1
    o.nu = fj_gf_herm(lcd + 1);    // |u| <= lc + ld
    o.nt = fj_gf_herm(lab + 1);
    o.nu1 = fj_gf_herm(lcd + 2);   // |u| <= lc + ld + 1
    o.nt1 = fj_gf_herm(lab + 2);
    o.axb = (la + 2) * (lb + 2) * (lab + 3);
    o.axk = (lc + 2) * (ld + 2) * (lcd + 3);
    int at = 0;
    auto take = [&](int n, int size) { const int r = at; at += ((n * size + 7) / 8) * 8; return r; };
    o.w = take(na * nb * nc * nd, 8);
    o.red = take(12 * GF_MAXW, 8);
    o.R0 = take(B * o.nR, 8);
    o.R1 = take(B * o.nR, 8);
    o.eb = take(B * 3 * o.axb, 8);
    o.ek = take(B * 3 * o.axk, 8);
    o.wf = take(B * (L + 2), 8);
    o.wk = take(B * na * nb * o.nu, 8);
    o.wb = take(B * nc * nd * o.nt, 8);
    o.sc = take(B * GF_SC, 8);
    // per primitive quartet one region holding first the E products Ecd
    // [ncd][nu] and Eab [nab][nt], then the gathered tower H [nt1][nu1]
    o.uw = nc * nd * o.nu + na * nb * o.nt;
    if (o.nt1 * o.nu1 > o.uw) o.uw = o.nt1 * o.nu1;
    o.un = take(B * o.uw, 8);
    o.hc = take(o.nR, 4);
    o.tw = take(o.nR, 4);
    o.pab = take(na * nb, 4);
    o.pcd = take(nc * nd, 4);
    o.add = take(o.nt1 * o.nu1, 2);
    o.bytes = at;
    return o;
}

// B for a launch: the largest that fits GF_LDS_BUDGET, at most GF_BMAX and
// the quartet's n_pq primitive quartets, and no more than the lanes need
// (one register tile per lane in the last stage).
__host__ __device__ inline FjGfLds fj_gf_plan(int la, int lb, int lc, int ld, int n_pq, int threads) {
    const FjGfLds one = fj_gf_layout(la, lb, lc, ld, 1), two = fj_gf_layout(la, lb, lc, ld, 2);
    const int per_b = two.bytes - one.bytes, fixed = one.bytes - per_b;
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
    const int nc = (lc + 1) * (lc + 2) / 2, nd = (ld + 1) * (ld + 2) / 2;
    auto tiles = [](int m, int n) { return ((m + GF_TM - 1) / GF_TM) * ((n + GF_TN - 1) / GF_TN); };
    // register tiles of the last stage, one primitive quartet
    const int el = tiles(na * nb, one.nt1) + tiles(nc * nd, one.nu1);
    int B = (GF_LDS_BUDGET - fixed) / per_b;
    const int want = (threads + el - 1) / el;
    if (B > want) B = want;
    if (B > GF_BMAX) B = GF_BMAX;
    if (B > n_pq) B = n_pq;
    if (B < 1) B = 1;
    return fj_gf_layout(la, lb, lc, ld, B);
}

// Components of the i-th Cartesian of an l shell, packed x | y << 4 | z << 8.
__device__ __forceinline__ int fj_gf_cart(int l, int i) { return fj_herm_decode(fj_herm_base(l) + i); }

// The differentiated product's E along one axis at Hermite order t: e0 (the
// product itself), ea (its first centre's derivative, 2 x1 E^{i+1,j} - i
// E^{i-1,j}) and eb (the second's, 2 x2 E^{i,j+1} - j E^{i,j-1}); table e of
// stride (nj, nt) built to (i + 1, j + 1).
__device__ __forceinline__ void fj_gf_axis(const double* __restrict__ e, int nj, int nt, int i,
                                           int j, int t, double x1, double x2, double& e0,
                                           double& ea, double& eb) {
    e0 = e[(i * nj + j) * nt + t];
    ea = x1 * e[((i + 1) * nj + j) * nt + t];
    if (i) ea -= (double)i * e[((i - 1) * nj + j) * nt + t];
    eb = x2 * e[(i * nj + j + 1) * nt + t];
    if (j) eb -= (double)j * e[(i * nj + j - 1) * nt + t];
}

// One register tile of C = A B: c[i][j] = sum_k A(m0 + i, k) B(k, n0 + j),
// A(m, k) = A[m sam + k sak], B(k, n) = Bm[k sbk + n sbn], rows past M and
// columns past N read row M - 1 / column N - 1 (their results are unused).
__device__ __forceinline__ void fj_gf_tile(const double* __restrict__ A, int sam, int sak,
                                           const double* __restrict__ Bm, int sbk, int sbn, int M,
                                           int N, int K, int m0, int n0, double (&c)[GF_TM][GF_TN]) {
    int ra[GF_TM], rb[GF_TN];
#pragma unroll
    for (int i = 0; i < GF_TM; ++i) ra[i] = (m0 + i < M ? m0 + i : M - 1) * sam;
#pragma unroll
    for (int j = 0; j < GF_TN; ++j) rb[j] = (n0 + j < N ? n0 + j : N - 1) * sbn;
#pragma unroll
    for (int i = 0; i < GF_TM; ++i)
#pragma unroll
        for (int j = 0; j < GF_TN; ++j) c[i][j] = 0.0;
    for (int k = 0; k < K; ++k) {
        double a[GF_TM], bv[GF_TN];
#pragma unroll
        for (int i = 0; i < GF_TM; ++i) a[i] = A[ra[i] + k * sak];
#pragma unroll
        for (int j = 0; j < GF_TN; ++j) bv[j] = Bm[k * sbk + rb[j]];
#pragma unroll
        for (int i = 0; i < GF_TM; ++i)
#pragma unroll
            for (int j = 0; j < GF_TN; ++j) c[i][j] += a[i] * bv[j];
    }
}

}  // namespace

// One workgroup's share of a launch: work items blockIdx.x, + gridDim.x,
// ... of n_items. A work item is one quartet's primitive quartets c chunk ..
// (c + 1) chunk - 1 (a small launch's quartets split over several
// workgroups, each adding its share into G). With a job table (several
// launches of one class merged, jobs sorted by q0, each with its own chunk),
// item i belongs to the last job whose q0 <= i and reads that job's sides
// from g's arrays at the job's offsets; without one, g is the launch.
__device__ __forceinline__ void fj_gf_body(const FjGfArgs& g0, const FjGfLds& o, const FjGfJob* jobs,
                                           int n_jobs, long long n_items) {
    FJ_DYN_SHARED(fj_gfloop_lds_block);
    char* const lds = reinterpret_cast<char*>(fj_gfloop_lds_block);
    double* const W = reinterpret_cast<double*>(lds + o.w);
    double* const red = reinterpret_cast<double*>(lds + o.red);
    double* const R0 = reinterpret_cast<double*>(lds + o.R0);
    double* const R1 = reinterpret_cast<double*>(lds + o.R1);
    double* const eb = reinterpret_cast<double*>(lds + o.eb);
    double* const ek = reinterpret_cast<double*>(lds + o.ek);
    double* const wf = reinterpret_cast<double*>(lds + o.wf);
    double* const Wk = reinterpret_cast<double*>(lds + o.wk);
    double* const Wb = reinterpret_cast<double*>(lds + o.wb);
    double* const sc = reinterpret_cast<double*>(lds + o.sc);
    const int nc_ = (g0.lc + 1) * (g0.lc + 2) / 2, nd_ = (g0.ld + 1) * (g0.ld + 2) / 2;
    double* const un = reinterpret_cast<double*>(lds + o.un);
    double* const Ecd = un;                                   // [b][cd][u], then
    double* const Eab = un + (size_t)(nc_ * nd_) * o.nu;      // [b][ab][t], then
    double* const Hm = un;                                    // [b][t][u]
    int* const hc = reinterpret_cast<int*>(lds + o.hc);
    int* const tw = reinterpret_cast<int*>(lds + o.tw);
    int* const pab = reinterpret_cast<int*>(lds + o.pab);
    int* const pcd = reinterpret_cast<int*>(lds + o.pcd);
    unsigned short* const add = reinterpret_cast<unsigned short*>(lds + o.add);

    const int tid = (int)threadIdx.x, nthr = (int)blockDim.x;
    const int la = g0.la, lb = g0.lb, lc = g0.lc, ld = g0.ld;
    const int L = la + lb + lc + ld, Lt = L + 1;
    const int na = (la + 1) * (la + 2) / 2, nb = (lb + 1) * (lb + 2) / 2;
    const int nc = (lc + 1) * (lc + 2) / 2, nd = (ld + 1) * (ld + 2) / 2;
    const int nab = na * nb, ncd = nc * nd;
    const int nu = o.nu, nt = o.nt, nu1 = o.nu1, nt1 = o.nt1, nR = o.nR, B = o.B;
    const int axb3 = 3 * o.axb, axk3 = 3 * o.axk;
    const int njb = lb + 2, ntb = la + lb + 3, njk = ld + 2, ntk = lc + ld + 3;

    // the class's tables, once per workgroup: the tower's step table and the
    // components of every Hermite index, the Cartesian pairs, [t + u]
    for (int j = tid; j < nR; j += nthr) {
        const int c = fj_herm_decode(j);
        const int x = c & 15, y = (c >> 4) & 15, z = c >> 8;
        hc[j] = c;
        int w = 0;
        if (j > 0) {
            const int d = x ? 0 : (y ? 1 : 2);
            const int td = d == 0 ? x : (d == 1 ? y : z);
            const int i1 = fj_herm_index(x - (d == 0), y - (d == 1), z - (d == 2));
            const int i2 = td >= 2 ? fj_herm_index(x - 2 * (d == 0), y - 2 * (d == 1), z - 2 * (d == 2)) : 0;
            w = i1 | (i2 << 10) | (d << 20) | ((td - 1) << 22);
        }
        tw[j] = w;
    }
    for (int r = tid; r < nab; r += nthr) {
        const int a = r / nb, b = r - a * nb;
        pab[r] = fj_gf_cart(la, a) | (fj_gf_cart(lb, b) << 12);
    }
    for (int r = tid; r < ncd; r += nthr) {
        const int c = r / nd, d = r - c * nd;
        pcd[r] = fj_gf_cart(lc, c) | (fj_gf_cart(ld, d) << 12);
    }
    for (int r = tid; r < nt1 * nu1; r += nthr) {
        const int t = r / nu1, u = r - t * nu1;
        const int ct = fj_herm_decode(t), cu = fj_herm_decode(u);
        add[r] = (unsigned short)fj_herm_index((ct & 15) + (cu & 15), ((ct >> 4) & 15) + ((cu >> 4) & 15),
                                               (ct >> 8) + (cu >> 8));
    }

    for (long long q0 = (long long)blockIdx.x; q0 < n_items; q0 += (long long)gridDim.x) {
        FjGfArgs g = g0;
        long long q = q0;
        if (jobs) {
            int lo = 0, hi = n_jobs - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (jobs[mid].q0 <= q0) lo = mid;
                else hi = mid - 1;
            }
            const FjGfJob J = jobs[lo];
            q = q0 - J.q0;
            g.n_bra = J.n_bra;
            g.n_ket = J.n_ket;
            g.k_ab = J.k_ab;
            g.k_cd = J.k_cd;
            g.diag = J.diag;
            g.chunk = J.chunk;
            g.bra_prim = g0.bra_prim + J.bra_prim;
            g.ket_prim = g0.bra_prim + J.ket_prim;
            g.AB = g0.AB + 3 * (size_t)J.bra_row;
            g.CD = g0.AB + 3 * (size_t)J.ket_row;
            g.bra_sh = g0.bra_sh + 2 * (size_t)J.bra_row;
            g.ket_sh = g0.bra_sh + 2 * (size_t)J.ket_row;
            g.bra_ao = g0.bra_ao + 2 * (size_t)J.bra_row;
            g.ket_ao = g0.bra_ao + 2 * (size_t)J.ket_row;
            g.bra_at = g0.bra_at + 2 * (size_t)J.bra_row;
            g.ket_at = g0.bra_at + 2 * (size_t)J.ket_row;
        }
        // the item: its quartet q and primitive quartets pq0 .. pq1 - 1
        const int n_pq = g.k_ab * g.k_cd;
        const int n_ch = (n_pq + g.chunk - 1) / g.chunk;
        const long long qi = q / n_ch;
        const int pq0 = (int)(q - qi * n_ch) * g.chunk;
        const int pq1 = pq0 + g.chunk < n_pq ? pq0 + g.chunk : n_pq;
        q = qi;
        // the quartet: (bra, ket), ket >= bra on the diagonal launch (its
        // upper triangle, row-major), every pair otherwise
        int bra, ket;
        if (g.diag) {
            const long long n = g.n_bra;
            const double b2 = (double)(2 * n + 1);
            long long b = (long long)((b2 - sqrt(b2 * b2 - 8.0 * (double)q)) * 0.5);
            if (b < 0) b = 0;
            if (b > n - 1) b = n - 1;
            while (b > 0 && b * n - b * (b - 1) / 2 > q) --b;
            while (b + 1 < n && (b + 1) * n - (b + 1) * b / 2 <= q) ++b;
            bra = (int)b;
            ket = (int)(b + (q - (b * n - b * (b - 1) / 2)));
        } else {
            bra = (int)(q / g.n_ket);
            ket = (int)(q - (long long)bra * g.n_ket);
        }
        const int oa = g.bra_ao[2 * bra], ob = g.bra_ao[2 * bra + 1];
        const int oc = g.ket_ao[2 * ket], od = g.ket_ao[2 * ket + 1];
        // the degeneracy 1/|Stab| (kernel_gf.py:591-606): 1/2 for a = b, for
        // c = d, and for the pair met with itself on the diagonal launch
        double deg = 1.0;
        if (la == lb && g.bra_sh[2 * bra] == g.bra_sh[2 * bra + 1]) deg *= 0.5;
        if (lc == ld && g.ket_sh[2 * ket] == g.ket_sh[2 * ket + 1]) deg *= 0.5;
        if (g.diag && ket == bra) deg *= 0.5;

        __syncthreads();   // the previous quartet's readers of W and red are done
        {
            const size_t n = (size_t)g.n_ao;
            const double* __restrict__ D = g.D;
            for (int r = tid; r < nab * ncd; r += nthr) {
                const int ab = r / ncd, cd = r - ab * ncd;
                const int a = ab / nb, b = ab - a * nb, c = cd / nd, d = cd - c * nd;
                const double w = 4.0 * D[(oa + a) * n + ob + b] * D[(oc + c) * n + od + d] -
                                 D[(oa + a) * n + oc + c] * D[(ob + b) * n + od + d] -
                                 D[(oa + a) * n + od + d] * D[(ob + b) * n + oc + c];
                W[r] = deg * w;
            }
        }
        double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
        double gC[3] = {0.0, 0.0, 0.0}, gD[3] = {0.0, 0.0, 0.0};
        const double* __restrict__ AB = g.AB + 3 * bra;
        const double* __restrict__ CD = g.CD + 3 * ket;
        const double* __restrict__ bp = g.bra_prim + (size_t)bra * g.k_ab * GF_REC;
        const double* __restrict__ kp = g.ket_prim + (size_t)ket * g.k_cd * GF_REC;
        __syncthreads();

        for (int base = pq0; base < pq1; base += B) {
            const int nbq = pq1 - base < B ? pq1 - base : B;
            // per primitive quartet b: its six E tables and its Boys seed
            for (int r = tid; r < 7 * nbq; r += nthr) {
                const int b = r / 7, s = r - 7 * b;
                const int pq = base + b, i = pq / g.k_cd, k = pq - i * g.k_cd;
                const double* __restrict__ rb = bp + i * GF_REC;
                const double* __restrict__ rk = kp + k * GF_REC;
                if (s < 3) {
                    fj_kl_eaxis_from(eb + b * axb3 + s * o.axb, la + 1, lb + 1, rb[6], rb[3 + s],
                                     rb[3 + s] + AB[s], 1.0);
                } else if (s < 6) {
                    fj_kl_eaxis_from(ek + b * axk3 + (s - 3) * o.axk, lc + 1, ld + 1, rk[6], rk[s],
                                     rk[s] + CD[s - 3], 1.0);
                } else {
                    const double p = rb[6], qz = rk[6];
                    const double PQx = rb[0] - rk[0], PQy = rb[1] - rk[1], PQz = rb[2] - rk[2];
                    const double rho = p * qz / (p + qz);
                    double* const wb_ = wf + b * (L + 2);
                    fj_kl_seed(wb_, rho, rho * (PQx * PQx + PQy * PQy + PQz * PQz), Lt);
                    ((Lt & 1) ? R1 : R0)[b * nR] = wb_[Lt];
                    double* const s_ = sc + b * GF_SC;
                    s_[0] = rb[7] * rk[7];
                    s_[1] = 2.0 * rb[8];
                    s_[2] = 2.0 * rb[9];
                    s_[3] = 2.0 * rk[8];
                    s_[4] = 2.0 * rk[9];
                    s_[5] = PQx;
                    s_[6] = PQy;
                    s_[7] = PQz;
                }
            }
            __syncthreads();
            // the towers, levels m = Lt - 1 .. 0: lanes over (b, t)
            for (int m = Lt - 1; m >= 0; --m) {
                double* __restrict__ cur = (m & 1) ? R1 : R0;
                const double* __restrict__ prv = (m & 1) ? R0 : R1;
                const int n_m = fj_herm_base(Lt - m + 1);
                for (int r = tid; r < nbq * n_m; r += nthr) {
                    const int b = r / n_m, j = r - b * n_m;
                    double v;
                    if (j == 0) {
                        v = wf[b * (L + 2) + m];
                    } else {
                        const int w = tw[j];
                        const int d = (w >> 20) & 3, c1 = w >> 22;
                        const double pq = sc[b * GF_SC + 5 + d];
                        v = pq * prv[b * nR + (w & 1023)];
                        if (c1 > 0) v += (double)c1 * prv[b * nR + ((w >> 10) & 1023)];
                    }
                    cur[b * nR + j] = v;
                }
                __syncthreads();
            }
            // the E products of every primitive quartet: Ecd[b][cd][u] with
            // the ket's sign (-1)^|u| folded in, Eab[b][ab][t]
            for (int r = tid; r < nbq * ncd * nu; r += nthr) {
                const int b = r / (ncd * nu), rr = r - b * (ncd * nu);
                const int cd = rr / nu, u = rr - cd * nu;
                const int pc = pcd[cd], cu = hc[u];
                const int ux = cu & 15, uy = (cu >> 4) & 15, uz = cu >> 8;
                const double* __restrict__ e = ek + b * axk3;
                const double v = e[(((pc & 15) * njk + ((pc >> 12) & 15)) * ntk + ux)] *
                                 e[o.axk + ((((pc >> 4) & 15) * njk + ((pc >> 16) & 15)) * ntk + uy)] *
                                 e[2 * o.axk + ((((pc >> 8) & 15) * njk + (pc >> 20)) * ntk + uz)];
                Ecd[b * o.uw + rr] = ((ux + uy + uz) & 1) ? -v : v;
            }
            for (int r = tid; r < nbq * nab * nt; r += nthr) {
                const int b = r / (nab * nt), rr = r - b * (nab * nt);
                const int ab = rr / nt, t = rr - ab * nt;
                const int pa = pab[ab], ct = hc[t];
                const double* __restrict__ e = eb + b * axb3;
                Eab[b * o.uw + rr] = e[(((pa & 15) * njb + ((pa >> 12) & 15)) * ntb + (ct & 15))] *
                                     e[o.axb + ((((pa >> 4) & 15) * njb + ((pa >> 16) & 15)) * ntb + ((ct >> 4) & 15))] *
                                     e[2 * o.axb + ((((pa >> 8) & 15) * njb + (pa >> 20)) * ntb + (ct >> 8))];
            }
            __syncthreads();
            // the weights in Hermite space, two small GEMMs per primitive
            // quartet in register tiles: Wk = W Ecd (nab x nu, over cd) and
            // Wb = W^T Eab (ncd x nt, over ab)
            {
                const int tmk = (nab + GF_TM - 1) / GF_TM, tnk = (nu + GF_TN - 1) / GF_TN;
                const int tmb = (ncd + GF_TM - 1) / GF_TM, tnb = (nt + GF_TN - 1) / GF_TN;
                const int nk = tmk * tnk, nbt = tmb * tnb;
                for (int r = tid; r < nbq * (nk + nbt); r += nthr) {
                    const int b = r / (nk + nbt), rr = r - b * (nk + nbt);
                    double c[GF_TM][GF_TN];
                    if (rr < nk) {
                        const int m0 = (rr / tnk) * GF_TM, n0 = (rr - (rr / tnk) * tnk) * GF_TN;
                        fj_gf_tile(W, ncd, 1, Ecd + b * o.uw, nu, 1, nab, nu, ncd, m0, n0, c);
                        double* __restrict__ out = Wk + b * nab * nu;
#pragma unroll
                        for (int i = 0; i < GF_TM; ++i)
#pragma unroll
                            for (int j = 0; j < GF_TN; ++j)
                                if (m0 + i < nab && n0 + j < nu) out[(m0 + i) * nu + n0 + j] = c[i][j];
                    } else {
                        const int q_ = rr - nk;
                        const int m0 = (q_ / tnb) * GF_TM, n0 = (q_ - (q_ / tnb) * tnb) * GF_TN;
                        fj_gf_tile(W, 1, ncd, Eab + b * o.uw, nt, 1, ncd, nt, nab, m0, n0, c);
                        double* __restrict__ out = Wb + b * ncd * nt;
#pragma unroll
                        for (int i = 0; i < GF_TM; ++i)
#pragma unroll
                            for (int j = 0; j < GF_TN; ++j)
                                if (m0 + i < ncd && n0 + j < nt) out[(m0 + i) * nt + n0 + j] = c[i][j];
                    }
                }
            }
            __syncthreads();
            // H[b][t][u] = [t + u] of each tower, t < nt1, u < nu1 (over the
            // E products, which are done)
            for (int r = tid; r < nbq * nt1 * nu1; r += nthr) {
                const int b = r / (nt1 * nu1), rr = r - b * (nt1 * nu1);
                // |t| + |u| = L + 2 (the corner no contraction reads) is past the tower
                const int ix = add[rr];
                Hm[b * o.uw + rr] = ix < nR ? R0[b * nR + ix] : 0.0;
            }
            __syncthreads();

            // the centres: Y = Wk H^T (nab x nt1, over u) and X = Wb H
            // (ncd x nu1, over t) in register tiles, each tile's elements
            // contracted at once with the differentiated E into the sums
            {
                const int tmy = (nab + GF_TM - 1) / GF_TM, tny = (nt1 + GF_TN - 1) / GF_TN;
                const int tmx = (ncd + GF_TM - 1) / GF_TM, tnx = (nu1 + GF_TN - 1) / GF_TN;
                const int ny = tmy * tny, nx = tmx * tnx;
                for (int r = tid; r < nbq * (ny + nx); r += nthr) {
                    const int b = r / (ny + nx), rr = r - b * (ny + nx);
                    const double* __restrict__ s_ = sc + b * GF_SC;
                    double c[GF_TM][GF_TN];
                    if (rr < ny) {
                        const int m0 = (rr / tny) * GF_TM, n0 = (rr - (rr / tny) * tny) * GF_TN;
                        fj_gf_tile(Wk + b * nab * nu, nu, 1, Hm + b * o.uw, 1, nu1, nab, nt1, nu, m0, n0, c);
                        const double* __restrict__ e = eb + b * axb3;
#pragma unroll
                        for (int i = 0; i < GF_TM; ++i) {
                            const int ab = m0 + i;
                            if (ab >= nab) continue;
                            const int pa = pab[ab];
                            const int ax = pa & 15, ay = (pa >> 4) & 15, az = (pa >> 8) & 15;
                            const int bx = (pa >> 12) & 15, by = (pa >> 16) & 15, bz = pa >> 20;
#pragma unroll
                            for (int j = 0; j < GF_TN; ++j) {
                                const int t = n0 + j;
                                if (t >= nt1) continue;
                                const int ct = hc[t];
                                const int tx = ct & 15, ty = (ct >> 4) & 15, tz = ct >> 8;
                                // outside the box t_k <= a_k + b_k (+1 on at most one axis) every term is zero
                                const int ox = tx - ax - bx, oy = ty - ay - by, oz = tz - az - bz;
                                if (ox > 1 || oy > 1 || oz > 1 || (ox == 1) + (oy == 1) + (oz == 1) > 1) continue;
                                const double y = c[i][j] * s_[0];
                                double ex, exa, exb, ey, eya, eyb, ez, eza, ezb;
                                fj_gf_axis(e, njb, ntb, ax, bx, tx, s_[1], s_[2], ex, exa, exb);
                                fj_gf_axis(e + o.axb, njb, ntb, ay, by, ty, s_[1], s_[2], ey, eya, eyb);
                                fj_gf_axis(e + 2 * o.axb, njb, ntb, az, bz, tz, s_[1], s_[2], ez, eza, ezb);
                                gA[0] += y * exa * ey * ez;
                                gA[1] += y * ex * eya * ez;
                                gA[2] += y * ex * ey * eza;
                                gB[0] += y * exb * ey * ez;
                                gB[1] += y * ex * eyb * ez;
                                gB[2] += y * ex * ey * ezb;
                            }
                        }
                    } else {
                        const int q_ = rr - ny;
                        const int m0 = (q_ / tnx) * GF_TM, n0 = (q_ - (q_ / tnx) * tnx) * GF_TN;
                        fj_gf_tile(Wb + b * ncd * nt, nt, 1, Hm + b * o.uw, nu1, 1, ncd, nu1, nt, m0, n0, c);
                        const double* __restrict__ e = ek + b * axk3;
#pragma unroll
                        for (int i = 0; i < GF_TM; ++i) {
                            const int cd = m0 + i;
                            if (cd >= ncd) continue;
                            const int pc = pcd[cd];
                            const int cx = pc & 15, cy = (pc >> 4) & 15, cz = (pc >> 8) & 15;
                            const int dx = (pc >> 12) & 15, dy = (pc >> 16) & 15, dz = pc >> 20;
#pragma unroll
                            for (int j = 0; j < GF_TN; ++j) {
                                const int u = n0 + j;
                                if (u >= nu1) continue;
                                const int cu = hc[u];
                                const int ux = cu & 15, uy = (cu >> 4) & 15, uz = cu >> 8;
                                const int ox = ux - cx - dx, oy = uy - cy - dy, oz = uz - cz - dz;
                                if (ox > 1 || oy > 1 || oz > 1 || (ox == 1) + (oy == 1) + (oz == 1) > 1) continue;
                                const double x = c[i][j] * (((ux + uy + uz) & 1) ? -s_[0] : s_[0]);
                                double ex, exc, exd, ey, eyc, eyd, ez, ezc, ezd;
                                fj_gf_axis(e, njk, ntk, cx, dx, ux, s_[3], s_[4], ex, exc, exd);
                                fj_gf_axis(e + o.axk, njk, ntk, cy, dy, uy, s_[3], s_[4], ey, eyc, eyd);
                                fj_gf_axis(e + 2 * o.axk, njk, ntk, cz, dz, uz, s_[3], s_[4], ez, ezc, ezd);
                                gC[0] += x * exc * ey * ez;
                                gC[1] += x * ex * eyc * ez;
                                gC[2] += x * ex * ey * ezc;
                                gD[0] += x * exd * ey * ez;
                                gD[1] += x * ex * eyd * ez;
                                gD[2] += x * ex * ey * ezd;
                            }
                        }
                    }
                }
            }
            __syncthreads();   // before the next batch rewrites the tables
        }

        // the twelve sums over the workgroup, then into G
        double v[12] = {gA[0], gA[1], gA[2], gB[0], gB[1], gB[2], gC[0], gC[1], gC[2], gD[0], gD[1], gD[2]};
        const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
        for (int c = 0; c < 12; ++c) {
            const double s = fj_wave_sum(v[c]);
            if (lane == 0) red[wave * 12 + c] = s;
        }
        __syncthreads();
        if (tid < 12) {
            double s = 0.0;
            for (int w = 0; w < (nthr >> 6); ++w) s += red[w * 12 + tid];
            const int pos = tid / 3, k = tid - pos * 3;
            const int atom = pos == 0 ? g.bra_at[2 * bra] : pos == 1 ? g.bra_at[2 * bra + 1]
                           : pos == 2 ? g.ket_at[2 * ket] : g.ket_at[2 * ket + 1];
            atomicAdd(g.G + 3 * atom + k, s);
        }
    }
}

__global__ void __launch_bounds__(256) fj_gfloop_k(FjGfArgs g, FjGfLds o) {
    const long long n_q = g.diag ? (long long)g.n_bra * (g.n_bra + 1) / 2 : (long long)g.n_bra * g.n_ket;
    const int n_pq = g.k_ab * g.k_cd;
    fj_gf_body(g, o, nullptr, 0, n_q * ((n_pq + g.chunk - 1) / g.chunk));
}

// Several launches of one class in one: g holds the class, n_atom, n_ao, D, G
// and the base pointers of the concatenated sides (bra_prim, AB, bra_sh,
// bra_ao, bra_at), which each job offsets for its bra and its ket.
__global__ void __launch_bounds__(256) fj_gfjobs_k(FjGfArgs g, FjGfLds o, const FjGfJob* jobs, int n_jobs,
                                                   long long n_items) {
    fj_gf_body(g, o, jobs, n_jobs, n_items);
}
