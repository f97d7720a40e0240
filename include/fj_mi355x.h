/*
 * fj_mi355x.h -- C-ABI of libfj_mi355x.so, the gfx950 (MI355X) two-electron
 * Fock-build library.  Plain pointers and sizes; every pointer argument is
 * DEVICE memory unless stated; every entry is asynchronous on `stream` and
 * returns 0 or an FJ_ERR_* / hipError_t code (the library validates shapes
 * before launching; the device has no error channel).
 *
 * Upstream interfaces replaced (paths relative to the FelizJueves checkout):
 *   fj_f_<la><lb><lc><ld>   the emitted fused-F unit f_<la><lb><lc><ld>
 *                           (eri/fock/kernel_f.py:655-692; parameter list
 *                           eri/fock/levers/seams.py:228-235,739-770), one
 *                           launch of walk_f's run(kernel, grid, block, args)
 *                           (eri/fock/walk_f.py:385-401)
 *   fj_fx_<la><lb><lc><ld>  the same kernel over ragged-degree pair tables
 *                           and a precision tier (no upstream counterpart:
 *                           the native builder's entry; the FP32 tier plays
 *                           the role of the fmixed lever,
 *                           eri/fock/levers/mixed.py:117-255)
 *
 * Ownership (upstream eri/j_coulomb/kernel_jab.py:38-43, walk_f.py:412-413):
 * the caller owns every buffer; F is ACCUMULATED (+=) and not closed -- the
 * caller zeroes it before the first launch of a build and closes
 * F <- F + F^T after the last (fj_close).
 */
#ifndef FJ_MI355X_H
#define FJ_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* fj_stream_t; /* a hipStream_t; NULL = default stream */

/* error codes (hipError_t values are passed through unchanged) */
#define FJ_OK 0
#define FJ_ERR_THREADS 9001  /* threads not a positive multiple of 64, <= 256 */
#define FJ_ERR_DEGREE 9002   /* primitive-pair degree outside 1..128 */
#define FJ_ERR_SHAPE 9003    /* inconsistent sizes (e.g. diag with n_bra != n_ket) */
#define FJ_ERR_CLASS 9004    /* no kernel for the requested class */
#define FJ_ERR_ARG 9005      /* another invalid argument */
#define FJ_ERR_WORKSPACE 9006 /* the class needs a launch workspace and none (or too
                                small a one for a single workgroup) was given */

/* One side of an F launch, as device pointers.  Pair i's primitive records
 * are prim[off[i] .. off[i+1]) (off == NULL: uniform degree k, records
 * prim[i*k .. (i+1)*k)); a record is 8 doubles (P_x,P_y,P_z, X_x,X_y,X_z,
 * zeta, U) with X = P - A (bra) or Q - C (ket) and
 * U = c_a c_b (pi/zeta)^{3/2} exp(-alpha beta/zeta |AB|^2)
 * (upstream eri/fock/kernel_f.py:529-606, eri/fock/walk_f.py:188-210).
 * sep[i] = A - B (3 doubles; may be NULL when the side's second shell is s),
 * G[i] the contracted Schwarz-type bound, sh[i] = (a, b) shell indices,
 * ao[i] = first AOs of the two shells. */
typedef struct FjSide {
    const double* prim;
    const int* off;
    int k;
    const double* sep;
    const double* G;
    const int* sh;
    const int* ao;
    int n;
} FjSide;

/* One system (molecule) of an fj_fx_* launch: its density, density block
 * maxima and Fock accumulator (device pointers) and sizes.  A launch's work
 * items each name one system; the pair tables carry that system's own shell
 * and AO indices. */
typedef struct FjSys {
    const double* D;     /* [n_ao][n_ao] */
    const double* Dsh;   /* [n_sh][n_sh] */
    double* F;           /* [n_ao][n_ao], accumulated */
    int n_ao;
    int n_sh;
} FjSys;

/* Library version (also checks that the code object loaded). */
int fj_version(void);

/* F <- F + F^T for an n x n row-major device matrix (upstream
 * eri/fock/kernel_f.py:621-628, close). */
int fj_close(fj_stream_t stream, int n, double* F);

/* Dsh[s][t] = max |D| over shell block (s, t) (upstream
 * eri/fock/kernel_f.py:608-618, density_blocks).  shell_ao / shell_size
 * are device int arrays of n_sh entries. */
int fj_density_blocks(fj_stream_t stream, int n_sh, int n_ao, const int* shell_ao,
                      const int* shell_size, const double* D, double* Dsh);

/* Boys function F_m(T), m = 0..M (M <= 12), for n values of T: out[i][m]
 * (upstream emit/_c.py:144-208, fj_boys<M>; host truth
 * reference/boys.py:184-218).  Test entry for the device function. */
int fj_boys_eval(fj_stream_t stream, int M, int n, const double* T, double* out);

/* fj_fx_<cls>: `bra`, `ket` are HOST pointers to FjSide records of device
 * pointers; `work` is a device int array [n_work][4] of (bra pair, first
 * ket, system index into the device array `sys`, ket end exclusive), one
 * workgroup each, `threads` consecutive kets from the first (the system
 * field packs the system index in bits 0-15 and, optionally, a chunk of the
 * bra pair's primitive pairs: first in bits 16-23, count in bits 24-31,
 * 0 = all; chunks of one quartet add up exactly in F); with a device
 * `ket_list` (may be NULL) the (first, end) range indexes that list instead,
 * one lane per entry, so a work item can name a compacted set of kets; an
 * entry packs the ket pair in bits 0-16 and, optionally, a chunk of its
 * primitive pairs: first in bits 17-23, count in bits 24-30 (0 = to the
 * end), so one quartet can be spread over several lanes; `precision` 0
 * is the FP64 kernel, 1 the FP32 tier (VRR and boundary in FP32; Boys, the
 * seed, the HRR and every F update in FP64).  `ws` (ws_bytes) is a device
 * workspace: the classes whose contracted boundary is too large for private
 * memory (fj_fx_workspace_bytes > 0: the f classes) keep it there, one slot
 * per resident lane, and run persistent over the work items with as many
 * workgroups as the workspace holds slots for; other classes ignore it (may
 * be NULL).  Launches that share a workspace must not overlap in time. */
/* Workspace bytes fj_fx_<la lb lc ld> uses for n_work items of `threads`
 * lanes (0: the class needs none; never more than FJ_WS_MAX_BLOCKS
 * workgroups' worth). */
long long fj_fx_workspace_bytes(int la, int lb, int lc, int ld, int threads, int n_work);
/* The native planner of one class pair (host code; no upstream counterpart:
 * the upstream plans in Python, eri/fock/launch_f.py:390-470, per
 * (degree, magnitude) bin).  Per bra pair of `bra`, the kets of the same
 * system (ket pairs mol_lo[m] .. mol_hi[m]; with `diag`, kets k >= b only)
 * that pass the density-free screen G_ab G_cd >= tau (the upstream
 * pre-launch screen, launch_f.py:448-466, per quartet).  Those that fail
 * the planning density's six-block screen G_ab G_cd dmax >= tau (dmax from
 * Dsh[m], n_sh[m] x n_sh[m]) are the cold kets: FP64 work items of whole
 * kets, listed last (the kernel's own screen drops them while the density
 * stays as planned; they are not counted as work).  The others are split
 * by the FP32 rule when `mixed` (fp32_bound >= 0:
 * bound < fp32_bound; else bound amp_b amp_k est_scale < fp32_error), are
 * expanded into lanes (ket pair, chunk of <= lane_flops / (cb per_prim)
 * primitive pairs, equal chunks per ket) sorted by primitive count, cut
 * into work items of `threads` lanes and bra chunks of cb primitive pairs,
 * costliest first.  lane_flops <= 0 only counts (n_quartets, n_prim,
 * flops).  Output arrays are allocated here; release them with
 * fj_plan_free. */
typedef struct FjPlanSide {
    const double* G;      /* pair bounds [n] */
    const int* deg;       /* primitive pairs per pair [n] */
    const int* sh;        /* shell indices within the pair's system [n][2] */
    const int* mol;       /* system of each pair [n] (pairs grouped by system) */
    const double* amp;    /* FP32 error amplification per pair [n] */
    int n;
} FjPlanSide;
typedef struct FjPlanResult {
    int* work[2];         /* per tier (0 FP64, 1 FP32): int4 work items */
    long long n_work[2];
    int* kets[2];         /* per tier: ket-list entries */
    long long n_kets[2];
    double n_quartets[2], n_prim[2], flops[2];
    long long n_cold;     /* the last n_cold FP64 items: the cold kets */
} FjPlanResult;
int fj_plan_pairs(const FjPlanSide* bra, const FjPlanSide* ket, const int* mol_lo,
    const int* mol_hi, const double* const* Dsh, const int* n_sh, int diag, double tau, int mixed,
    double fp32_error, double fp32_bound, double est_scale, int threads, double lane_flops,
    double per_prim, double per_quartet, int max_prim, FjPlanResult* out);
void fj_plan_free(FjPlanResult* out);

/* BEGIN F CLASSES (written by felizjueves_amd/gen_hip.py) */
int fj_f_0000(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0000(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0000(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0001(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0001(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0001(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0002(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0002(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0002(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0003(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0003(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0003(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0010(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0010(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0011(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0011(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0011(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0012(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0012(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0012(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0013(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0013(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0013(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0020(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0020(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0021(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0021(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0022(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0022(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0022(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0023(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0023(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0023(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0030(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0030(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0031(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0031(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0032(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0032(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0033(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0033(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0033(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0101(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0101(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0101(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0102(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0102(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0102(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0103(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0103(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0103(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0110(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0110(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0111(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0111(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0111(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0112(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0112(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0112(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0113(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0113(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0113(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0120(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0120(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0121(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0121(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0122(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0122(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0122(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0123(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0123(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0123(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0130(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0130(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0131(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0131(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0132(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0132(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0133(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0133(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0133(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0202(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0202(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0202(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0203(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0203(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0203(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0210(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0210(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0211(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0211(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0211(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0212(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0212(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0212(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0213(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0213(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0213(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0220(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0220(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0221(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0221(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0222(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0222(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0222(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0223(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0223(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0223(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0230(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0230(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0231(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0231(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0232(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0232(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0233(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0233(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0233(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0303(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0303(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0303(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0310(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0310(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0311(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0311(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0311(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0312(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0312(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0312(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0313(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0313(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0313(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0320(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0320(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0321(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0321(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0322(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0322(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0322(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0323(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0323(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0323(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0330(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0330(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0331(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0331(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_0332(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0332(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_0333(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_0333(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_0333(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1010(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1010(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1011(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1011(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1012(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1012(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1013(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1013(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1020(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1020(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1021(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1021(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1022(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1022(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1023(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1023(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1030(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1030(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1031(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1031(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1032(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1032(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1033(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1033(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1111(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1111(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1111(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1112(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1112(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1112(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1113(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1113(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1113(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1120(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1120(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1121(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1121(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1122(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1122(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1122(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1123(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1123(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1123(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1130(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1130(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1131(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1131(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1132(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1132(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1133(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1133(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1133(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1212(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1212(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1212(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1213(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1213(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1213(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1220(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1220(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1221(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1221(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1222(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1222(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1222(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1223(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1223(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1223(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1230(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1230(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1231(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1231(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1232(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1232(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1233(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1233(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1233(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1313(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1313(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1313(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1320(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1320(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1321(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1321(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1322(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1322(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1322(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1323(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1323(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1323(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1330(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1330(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1331(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1331(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_1332(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1332(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_1333(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_1333(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_1333(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2020(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2020(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2021(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2021(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2022(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2022(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2023(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2023(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2030(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2030(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2031(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2031(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2032(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2032(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2033(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2033(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2121(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2121(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2122(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2122(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2123(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2123(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2130(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2130(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2131(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2131(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2132(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2132(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2133(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2133(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_2222(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_2222(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2222(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_2223(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_2223(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2223(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2230(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2230(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2231(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2231(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2232(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2232(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_2233(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_2233(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2233(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_2323(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_2323(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2323(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2330(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2330(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2331(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2331(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_2332(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2332(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_2333(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_2333(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_2333(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_3030(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_3030(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_3031(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_3031(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_3032(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_3032(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_3033(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_3033(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_3131(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_3131(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_3132(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_3132(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_3133(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_3133(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_3232(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_3232(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_fx_3233(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_3233(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
int fj_f_3333(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int diag, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* F);
int fj_fx_3333(fj_stream_t stream, int threads, int precision, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,
    double tau, double* ws, long long ws_bytes);
int fj_fxm_3333(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,
    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,
    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);
/* END F CLASSES */

/* The MMD families, one entry per class, each with the upstream kernel's
 * parameter list after (stream, threads):
 *   fj_schwarz_<la><lb>  Q_mn = sqrt(max((mn|mn), 0)), one thread per shell
 *                        pair (upstream eri/kernel_schwarz.py:76-101); (df)
 *                        and (ff) on the loop-form kernel of fj_schwarz_loop
 *   fj_dq_<lc>_<ld>      D_q and Dmax per ket record
 *                        (upstream eri/j_coulomb/kernel_dq.py:10-60)
 *   fj_pq_L<L>_p<lp>_q<lq>  J_p = sum_k sum_q [p|q] D_q, kets sorted by
 *                        Iq Dmax, one wavefront per bra record
 *                        (upstream mmd/kernel.py:16-60,476-486)
 *   fj_jab_<la><lb>      J_ab = sum_p (s E_p) J_p scattered into J, atomics
 *                        (upstream eri/j_coulomb/kernel_jab.py:10-60)
 *   fj_k_<la><lc><lb><ld>  K[bra, a, b] = sum_ket D_cd (ac|bd), kets sorted
 *                        by Iq (upstream eri/k_exchange/kernel_k.py:1-110);
 *                        straight-line kernels through d, the classes with
 *                        an f shell on the loop-form kernel of fj_k_loop */
/* The Schwarz kernel in loop form for any pair class (la lb) through (ff),
 * the class as arguments, otherwise the fj_schwarz_<la><lb> parameter list
 * (upstream eri/kernel_schwarz.py:249-412). One wavefront per shell pair
 * whatever `threads` asks; the entries of (df) and (ff) call it. */
int fj_schwarz_loop(fj_stream_t stream, int la, int lb, int threads, int n_pair, int n_rec, int n_ao,
                    const double* zeta_ab, const double* P, const double* XPA, const double* XPB,
                    const double* s_ab, const int* rec0, const int* rec1, const int* off_a,
                    const int* off_b, double* schwarz_bound);
/* The K kernel in loop form for any class (la lc | lb ld) through f, the
 * class as arguments, otherwise the fj_k_<la><lc><lb><ld> parameter list
 * (upstream eri/k_exchange/kernel_k.py:314, host k_quartet k.py:395). One
 * workgroup of two wavefronts per bra record and slice of its kets
 * whatever `threads` asks (checked as for the straight-line kernels); with
 * several slices the launch first clears K's n_bra rows with a fill kernel
 * and the slices add into them, so K is written either way. The f classes'
 * entries call it, and it also runs the classes through d for checking one
 * form against the other. */
int fj_k_loop(fj_stream_t stream, int la, int lc, int lb, int ld, int threads, int n_bra, int n_ket,
              int n_dc, double tau, double maxD, const double* P, const double* zeta_ac,
              const double* XPA, const double* XPC, const double* s_ac, const double* Ip,
              const int* dc, const double* Q, const double* eta, const double* XQB,
              const double* XQD, const double* s_bd, const double* Iq, const int* dd,
              const double* Dcd, double* K);
/* K[oa[r] + i][ob + j] += Kr[r][i * nb + j] for n_rec records: the K
 * kernel's per-record blocks summed into K (the host-side sum and scatter
 * of upstream eri/k_exchange/launch_k.py:401-479, on the device). */
int fj_k_scatter(fj_stream_t stream, int n_rec, int na, int nb, int n_ao, const int* oa, int ob,
                 const double* Kr, double* K);

/* The precision selector's network (no upstream counterpart: the learned
 * FP32/FP64 tier choice per work item).  x[F] (F <= 32, float32, already
 * normalised) -> 64 -> 64 -> 1 with ReLU; parameters in one float32 block
 * P[fj_mlp_nparams(F)] = W1[64][F] b1[64] W2[64][64] b2[64] w3[64] b3 and its
 * BF16 copy Wb (uint16 bit patterns; the forward pass uses the BF16 weights
 * and the FP32 biases, accumulating in FP32). */
int fj_mlp_nparams(int F);
int fj_mlp_forward(fj_stream_t stream, int n, int F, const float* X, const unsigned short* Wb,
                   const float* P, float* Y);
/* G += dL/dP and *loss += L for L = mean_i (y_i - t_i)^2 (G, loss zeroed by
 * the caller). */
int fj_mlp_backward(fj_stream_t stream, int n, int F, const float* X, const float* T,
                    const unsigned short* Wb, const float* P, float* G, float* loss);
/* The same gradients for the pinball loss of `quantile` (0 < q < 1): mean
 * over samples of q (t - y) when t >= y, else (q - 1)(t - y). */
int fj_mlp_backward_quantile(fj_stream_t stream, int n, int F, const float* X, const float* T,
    const unsigned short* Wb, const float* P, float* G, float* loss, float quantile);
/* One Adam step (step t >= 1) on P from G, moments M, V; refreshes Wb. */
int fj_mlp_adam(fj_stream_t stream, int np, float* P, const float* G, float* M, float* V,
                unsigned short* Wb, float lr, float beta1, float beta2, float eps, int t);
/* Wb <- BF16(P), round to nearest even. */
int fj_mlp_round(fj_stream_t stream, int np, const float* P, unsigned short* Wb);

/* fj_f_noalloc: one launch of fj_f_<la><lb><lc><ld> (same arguments and
 * result; AB / CD may be NULL when lb / ld is 0) that allocates nothing.
 * No upstream counterpart: it is for callers that cannot allocate (code
 * under stream capture, serving loops, harnesses that forbid hipMalloc);
 * fj_f_<cls> allocates its staged classes' workspace on first use.  The
 * workspace here is a fixed device array of fj_f_noalloc_scratch_bytes()
 * (64 MiB) per GPU, held until the process ends; the persistent kernel runs
 * on as many workgroups as its slots fit (fewer than fj_f_<cls> uses for
 * the largest classes: (dd|dd) gets 62 workgroups of 64 lanes).  A class
 * whose one workgroup needs more than the array returns FJ_ERR_WORKSPACE.
 * Calls that overlap in time on one GPU share the array and race. */
int fj_f_noalloc(fj_stream_t stream, int la, int lb, int lc, int ld, int threads, int n_bra,
                 int n_ket, int k_ab, int k_cd, int n_sh, int n_ao, int diag, double tau,
                 const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh,
                 const int* bra_ao, const double* ket_prim, const double* CD, const double* G_cd,
                 const int* ket_sh, const int* ket_ao, const double* Dsh, const double* D,
                 double* F);
long long fj_f_noalloc_scratch_bytes(void);

/* Per-build precision tiers of one mixed launch (fj_fxm_<cls> with a `tier`
 * array), chosen on the device from the CURRENT density block maxima, so a
 * plan made with one density stays correct for another (no upstream
 * counterpart: the learned selector's inference, fused ahead of the ERI
 * launch).  One wavefront per work item: over the item's kets that pass the
 * ERI kernel's six-block screen (G_ab G_cd dmax >= tau), the largest bound
 * (fp32_bound >= 0: FP32 iff it is < fp32_bound) or bound amp_ab amp_cd
 * est_scale (FP32 iff < fp32_error).  With `feat` (the learned selector):
 * an item the rule keeps in FP64 goes to FP32 iff inside[item], the density
 * feature log10(max dmax) lies in [d_lo, d_hi], and the network's
 * prediction (fj_mlp_forward's arithmetic) on feat[item] (n_feat floats,
 * standardised; the last is replaced by (log10(max dmax) - d_mean) / d_std)
 * plus margin is below log_budget.  tier[item] = 1 for FP32, 0 for FP64 (an
 * item none of whose quartets passes the screen: 0).  `bra`, `ket`, `sel`
 * are HOST pointers to records of device pointers; `sys` as fj_fx_. */
typedef struct FjTierSel {
    const int* work;              /* [n_items][4] the launch's work items */
    const int* ket_list;          /* the launch's ket list */
    int n_items;
    int diag;
    double tau;
    const double* amp_bra;        /* [bra.n] FP32 error amplification per pair */
    const double* amp_ket;        /* [ket.n] */
    double est_scale, fp32_error, fp32_bound;
    const float* feat;            /* NULL: the bound rule alone */
    const unsigned char* inside;  /* [n_items] */
    int n_feat;
    double d_mean, d_std, d_lo, d_hi;
    const unsigned short* Wb;     /* the network, as fj_mlp_forward */
    const float* P;
    double margin, log_budget;
    int* tier;                    /* out: [n_items] */
} FjTierSel;
int fj_tier_select(fj_stream_t stream, const FjSide* bra, const FjSide* ket, const FjSys* sys,
                   const FjTierSel* sel);
/* Every mixed launch of a build in ONE launch: `jobs` is a DEVICE array of
 * n_jobs records (the two sides and the FjTierSel of one mixed launch, its
 * items the global items [first, next job's first)), n_items the total;
 * every job with `feat` set uses the one network Wb / P of n_feat inputs
 * (n_feat 0: no job has a network). */
typedef struct FjTierJob {
    FjSide bra, ket;
    FjTierSel sel;
    int first;
    int pad;
} FjTierJob;
int fj_tier_select_jobs(fj_stream_t stream, int n_jobs, const FjTierJob* jobs, int n_items,
                        const FjSys* sys, int n_feat, const unsigned short* Wb, const float* P);

/* The split HGP J and K kernels (upstream eri/hjk/kernel_hj.py:1-140 and
 * kernel_hk.py:1-110, canonical class mode; parameter list params_of,
 * kernel_hj.py:305-315): one kernel per orbit of the 8-fold ERI symmetry,
 * named after its representative (orbits.representative), through (dd|dd).
 * A launch is one ORDERED batch pair of the split stream (hjk.hjk_launches,
 * upstream eri/hjk/launch_hjk.py:280-342) presented in the representative's
 * positions, `perm` the orientation's index in the symmetry group (upstream
 * hgp/orbits.py GROUP order).  fj_hj_: J[a_i][b_j] += D[c_k][d_l] (ab|cd)
 * for the ordered quartet (a b|c d); fj_hk_: K[a_i][c_k] += D[b_j][d_l]
 * (ab|cd) (K's Hermite pairing: the class is (la, lc | lb, ld) of K_ab +=
 * D_cd (ac|bd)).  One block per quartet, no weights, no degeneracy, no
 * canonical half; a quartet is skipped when G_ab G_cd Dsh[density pair] <
 * tau; a perm outside the orbit's orientations deposits nothing.  J / K are
 * accumulated, never closed or mirrored. */
/* BEGIN SPLIT CLASSES (written by felizjueves_amd/gen_hip.py) */
int fj_hj_0000(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_0000(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_0010(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_0010(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_0011(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_0011(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_0020(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_0020(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_0021(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_0021(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_0022(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_0022(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_1010(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_1010(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_1011(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_1011(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_1111(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_1111(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_1121(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_1121(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_1122(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_1122(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_2010(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_2010(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_2011(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_2011(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_2020(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_2020(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_2021(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_2021(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_2110(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_2110(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_2121(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_2121(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_2210(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_2210(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_2220(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_2220(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_2221(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_2221(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
int fj_hj_2222(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* J);
int fj_hk_2222(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
    int n_sh, int n_ao, int perm, double tau,
    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh, const int* bra_ao,
    const double* ket_prim, const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao,
    const double* Dsh, const double* D, double* K);
/* END SPLIT CLASSES */

/* The fused-F pair stream on the device (upstream eri/fock/launch_f.py:
 * 390-470, f_launches; ao/pairs.py:270-290,528-560 for the primitive-pair
 * bound and the prune): for the unique shell pairs pair_ab (canonical, in
 * the upstream's unique-pair order) with their primitive pairs at
 * prim_off, the primitive bounds g, the (compensated) sums G, eps = tau /
 * (2 max G), each pair's kept primitive pairs (local indices, bound-
 * descending, stably; the longest tail with sum < eps dropped), dmax from
 * Dsh, the key G dmax (G with ket_break), the eq. 29 magnitude index
 * against mags[i] = tau^(i/n), and the batches: the kept pairs in bin order
 * (la, lb, degree, magnitude) ascending, each bin by key descending (stable
 * in pair order), with each batch's start, magnitude index and largest G.
 * Everything is written to the DEVICE arrays of `out`; `ws` is a device
 * workspace of fj_stream_workspace_bytes(n_pair, n_prim).
 * fj_stream_launches then lists the batch pairs I <= J (I outer) whose
 * magnitude indices sum below n: launches[k] = (I, J), n_launch of them
 * (row_off [n_batch + 1], launches [n_batch (n_batch + 1) / 2][2]). */
typedef struct FjStreamIn {
    int n_pair, n_prim, n_sh, n, ket_break;
    double tau;
    const int* pair_ab;
    const int* prim_off;
    const int* shell_l;
    const double* centre;
    const int* nprim;
    const int* pstart;
    const double* exps;
    const double* coefs;
    const double* Dsh;
    const double* mags;
} FjStreamIn;
typedef struct FjStreamOut {
    double* g;
    double* G;
    double* dmax;
    double* key;
    int* degree;
    int* mag;
    int* keep;
    int* order;
    int* batch_start;
    int* batch_mag;
    unsigned long long* batch_gmax_bits;
    unsigned long long* gmax_bits;
    double* eps;
    int* n_batch;
    int* n_kept;
} FjStreamOut;
long long fj_stream_workspace_bytes(int n_pair, int n_prim);
int fj_stream_plan(fj_stream_t stream, const FjStreamIn* in, FjStreamOut* out, void* ws,
                   long long ws_bytes);
int fj_stream_launches(fj_stream_t stream, int n_batch, int n, const int* batch_mag, int* row_off,
                       int* launches, int* n_launch);

/* Density fitting (upstream eri/density_fit/, the metric and three-center
 * integrals; DF-J and DF-K are out of the upstream's scope too).
 * fj_metric_loop / fj_metric_<lP><lQ>: (P|Q) of the aux shell pairs of class
 * (lP <= lQ, each <= 6), the upstream kernel_metric parameter list
 * (kernel_metric.py:460-480): aux records A[n_rec][3], alpha, s_aux = c
 * (pi/alpha)^{3/2}; per pair the record ranges [p_rec0, p_rec1), [q_rec0,
 * q_rec1) and the first aux functions off_p, off_q; metric [n_aux][n_aux]
 * written (both triangles; a diagonal pair writes a >= b, mirrored).
 * fj_eri3_loop / fj_eri3_<la><lb><lP>: (mn|P) of the cells (AO shell pair
 * la <= lb <= 3, aux shell lP <= 6), the upstream kernel_eri3 parameter
 * list (kernel_eri3.py:805-838): bra records (zeta_ab, P, XPA, XPB, s_ab =
 * c_m c_n (pi/zeta)^{3/2}), aux records, per cell the two record ranges and
 * off_a, off_b, off_p; eri3 [n_ao][n_ao][n_aux] written, mirrored in (m, n)
 * (a diagonal shell pair writes a >= b).  Plain stores, no accumulation:
 * every element belongs to one cell.  The loop entries take the class as
 * arguments (FJ_ERR_CLASS past it). */
int fj_metric_loop(fj_stream_t stream, int lP, int lQ, int threads, int n_pair, int n_rec,
                   int n_aux, const double* A, const double* alpha, const double* s_aux,
                   const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1,
                   const int* off_p, const int* off_q, double* metric);
int fj_eri3_loop(fj_stream_t stream, int la, int lb, int lP, int threads, int n_cell,
                 int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab,
                 const double* P, const double* XPA, const double* XPB, const double* s_ab,
                 const double* A, const double* alpha, const double* s_aux, const int* ab_rec0,
                 const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a,
                 const int* off_b, const int* off_p, double* eri3);
/* The same with max_rp, the largest record-pair count of a pair / cell (it
 * sizes the batch of record pairs a workgroup takes at once; the entries
 * above assume the most). No upstream counterpart. */
int fj_metric_loop_rp(fj_stream_t stream, int lP, int lQ, int threads, int n_pair, int n_rec,
                      int n_aux, const double* A, const double* alpha, const double* s_aux,
                      const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1,
                      const int* off_p, const int* off_q, double* metric, int max_rp);
int fj_eri3_loop_rp(fj_stream_t stream, int la, int lb, int lP, int threads, int n_cell,
                    int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab,
                    const double* P, const double* XPA, const double* XPB, const double* s_ab,
                    const double* A, const double* alpha, const double* s_aux, const int* ab_rec0,
                    const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a,
                    const int* off_b, const int* off_p, double* eri3, int max_rp);
/* BEGIN DF CLASSES (written by felizjueves_amd/gen_hip.py) */
int fj_metric_00(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_01(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_02(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_03(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_04(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_05(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_06(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_11(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_12(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_13(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_14(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_15(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_16(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_22(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_23(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_24(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_25(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_26(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_33(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_34(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_35(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_36(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_44(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_45(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_46(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_55(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_56(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_metric_66(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha, const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0, const int* q_rec1, const int* off_p, const int* off_q, double* metric);
int fj_eri3_000(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_001(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_002(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_003(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_004(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_005(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_006(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_010(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_011(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_012(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_013(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_014(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_015(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_016(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_020(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_021(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_022(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_023(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_024(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_025(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_026(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_030(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_031(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_032(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_033(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_034(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_035(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_036(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_110(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_111(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_112(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_113(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_114(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_115(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_116(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_120(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_121(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_122(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_123(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_124(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_125(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_126(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_130(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_131(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_132(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_133(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_134(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_135(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_136(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_220(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_221(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_222(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_223(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_224(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_225(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_226(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_230(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_231(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_232(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_233(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_234(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_235(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_236(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_330(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_331(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_332(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_333(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_334(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_335(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
int fj_eri3_336(fj_stream_t stream, int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux, const double* zeta_ab, const double* P, const double* XPA, const double* XPB, const double* s_ab, const double* A, const double* alpha, const double* s_aux, const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1, const int* off_a, const int* off_b, const int* off_p, double* eri3);
/* END DF CLASSES */

/* ---------------------------------------------------------------------------
 * The fused two-electron nuclear gradient (replaces the upstream gf_ kernels,
 * gradient/kernel_gf.py:523 build_gf_kernel, its parameter list
 * kernel_gf.py:501-512 and host stream gradient/interpreted.py:293-351):
 *
 *   G[atom][k] += dE_2e / dR_{atom,k},  E_2e = 1/2 tr(D (J - K/2))
 *
 * over one launch's unique quartets, bra pair I x ket pair J (J >= I on the
 * diagonal launch), with the weights w = deg (4 D_ab D_cd - D_ac D_bd -
 * D_ad D_bc), all four centres differentiated. A side is packed as the
 * upstream's gf_ side (kernel_gf.py:366-395): prim [n][k][10] = P(3),
 * X = P - first centre (3), zeta, U = c1 c2 (pi/zeta)^{3/2}
 * exp(-e1 e2/zeta |sep|^2), e1, e2 (the two primitive exponents); sep [n][3]
 * = first centre - second (AB, CD); sh, ao, at [n][2] the shell indices,
 * AO offsets and atoms. bra_sh is read only when la == lb and ket_sh only
 * when lc == ld (the upstream omits them otherwise; NULL is accepted then).
 * G [n_atom][3] is accumulated with atomics and zeroed by the caller before
 * the first launch; there is no screen and no tau (the upstream's gf_ has
 * none). Classes through (dd|dd); FJ_ERR_CLASS past d.
 * ------------------------------------------------------------------------- */
int fj_gf_loop(fj_stream_t stream, int la, int lb, int lc, int ld, int threads, int n_bra,
               int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag,
               const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao,
               const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh,
               const int* ket_ao, const int* ket_at, const double* D, double* G);
/* The LDS bytes of one workgroup of the class for quartets of n_pq primitive
 * quartets at `threads` lanes (0 past d). */
int fj_gf_lds_bytes(int la, int lb, int lc, int ld, int n_pq, int threads);
/* The batch B of primitive quartets one workgroup of the class takes at once
 * for quartets of n_pq primitive quartets at `threads` lanes (0 past d). */
int fj_gf_batch(int la, int lb, int lc, int ld, int n_pq, int threads);
/* Every launch of one class in one kernel launch (the stream's launches of a
 * class are many and small: one per pair of contraction degrees). The sides
 * of all launches are concatenated: prim (GREC doubles per primitive pair),
 * sep [rows][3], sh / ao / at [rows][2]; job j reads its bra at row bra_row,
 * record offset bra_prim (doubles), and its ket likewise. A job's work items
 * are its quartets (in the order of one fj_gf_<cls> launch of those sides),
 * each split into ceil(k_ab k_cd / chunk) items of `chunk` primitive
 * quartets; job j owns items q0 .. q0 + (its count) - 1 of n_items. jobs
 * (device memory) are sorted by q0, q0[0] = 0, chunk >= 1, and max_pq bounds
 * every job's k_ab k_cd. The table is trusted as the host stream builds it
 * (gradient.py); the per-class entries are the validated single launches.
 * No upstream counterpart: the upstream launches each batch pair separately
 * (gradient/interpreted.py:348-350). */
typedef struct {
    int bra_row, ket_row;
    int n_bra, n_ket, k_ab, k_cd, diag, chunk;
    long long bra_prim, ket_prim, q0;
} FjGfJob;
int fj_gf_jobs(fj_stream_t stream, int la, int lb, int lc, int ld, int threads, int n_jobs,
               const FjGfJob* jobs, long long n_items, int max_pq, int n_atom, int n_ao,
               const double* prim, const double* sep, const int* sh, const int* ao, const int* at,
               const double* D, double* G);
/* BEGIN GRADIENT CLASSES (written by felizjueves_amd/gen_hip.py) */
int fj_gf_0000(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0001(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0002(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0011(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0012(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0022(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0101(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0102(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0111(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0112(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0122(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0202(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0211(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0212(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_0222(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_1111(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_1112(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_1122(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_1212(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_1222(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
int fj_gf_2222(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag, const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao, const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh, const int* ket_ao, const int* ket_at, const double* D, double* G);
/* END GRADIENT CLASSES */

/* BEGIN MMD CLASSES (written by felizjueves_amd/gen_mmd.py) */
int fj_schwarz_00(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_ao,
    const double* zeta_ab, const double* P, const double* XPA, const double* XPB,
    const double* s_ab, const int* rec0, const int* rec1, const int* off_a,
    const int* off_b, double* schwarz_bound);
int fj_dq_0_0(fj_stream_t stream, int threads, int n_rec, int n_pair,
    const double* eta, const double* XPA, const double* XPB, const double* s,
    const double* w, const int* pair, const double* Dcd, double* D, double* Dmax);
int fj_jab_00(fj_stream_t stream, int threads, int n_rec, int n_ao, int n_rows,
    const double* zeta_ab, const double* XPA, const double* XPB, const double* s,
    const int* off_a, const int* off_b, const int* row, const double* Jp, double* J);
int fj_schwarz_01(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_ao,
    const double* zeta_ab, const double* P, const double* XPA, const double* XPB,
    const double* s_ab, const int* rec0, const int* rec1, const int* off_a,
    const int* off_b, double* schwarz_bound);
int fj_dq_0_1(fj_stream_t stream, int threads, int n_rec, int n_pair,
    const double* eta, const double* XPA, const double* XPB, const double* s,
    const double* w, const int* pair, const double* Dcd, double* D, double* Dmax);
int fj_jab_01(fj_stream_t stream, int threads, int n_rec, int n_ao, int n_rows,
    const double* zeta_ab, const double* XPA, const double* XPB, const double* s,
    const int* off_a, const int* off_b, const int* row, const double* Jp, double* J);
int fj_schwarz_02(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_ao,
    const double* zeta_ab, const double* P, const double* XPA, const double* XPB,
    const double* s_ab, const int* rec0, const int* rec1, const int* off_a,
    const int* off_b, double* schwarz_bound);
int fj_dq_0_2(fj_stream_t stream, int threads, int n_rec, int n_pair,
    const double* eta, const double* XPA, const double* XPB, const double* s,
    const double* w, const int* pair, const double* Dcd, double* D, double* Dmax);
int fj_jab_02(fj_stream_t stream, int threads, int n_rec, int n_ao, int n_rows,
    const double* zeta_ab, const double* XPA, const double* XPB, const double* s,
    const int* off_a, const int* off_b, const int* row, const double* Jp, double* J);
int fj_schwarz_03(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_ao,
    const double* zeta_ab, const double* P, const double* XPA, const double* XPB,
    const double* s_ab, const int* rec0, const int* rec1, const int* off_a,
    const int* off_b, double* schwarz_bound);
int fj_dq_0_3(fj_stream_t stream, int threads, int n_rec, int n_pair,
    const double* eta, const double* XPA, const double* XPB, const double* s,
    const double* w, const int* pair, const double* Dcd, double* D, double* Dmax);
int fj_jab_03(fj_stream_t stream, int threads, int n_rec, int n_ao, int n_rows,
    const double* zeta_ab, const double* XPA, const double* XPB, const double* s,
    const int* off_a, const int* off_b, const int* row, const double* Jp, double* J);
int fj_schwarz_11(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_ao,
    const double* zeta_ab, const double* P, const double* XPA, const double* XPB,
    const double* s_ab, const int* rec0, const int* rec1, const int* off_a,
    const int* off_b, double* schwarz_bound);
int fj_dq_1_1(fj_stream_t stream, int threads, int n_rec, int n_pair,
    const double* eta, const double* XPA, const double* XPB, const double* s,
    const double* w, const int* pair, const double* Dcd, double* D, double* Dmax);
int fj_jab_11(fj_stream_t stream, int threads, int n_rec, int n_ao, int n_rows,
    const double* zeta_ab, const double* XPA, const double* XPB, const double* s,
    const int* off_a, const int* off_b, const int* row, const double* Jp, double* J);
int fj_schwarz_12(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_ao,
    const double* zeta_ab, const double* P, const double* XPA, const double* XPB,
    const double* s_ab, const int* rec0, const int* rec1, const int* off_a,
    const int* off_b, double* schwarz_bound);
int fj_dq_1_2(fj_stream_t stream, int threads, int n_rec, int n_pair,
    const double* eta, const double* XPA, const double* XPB, const double* s,
    const double* w, const int* pair, const double* Dcd, double* D, double* Dmax);
int fj_jab_12(fj_stream_t stream, int threads, int n_rec, int n_ao, int n_rows,
    const double* zeta_ab, const double* XPA, const double* XPB, const double* s,
    const int* off_a, const int* off_b, const int* row, const double* Jp, double* J);
int fj_schwarz_13(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_ao,
    const double* zeta_ab, const double* P, const double* XPA, const double* XPB,
    const double* s_ab, const int* rec0, const int* rec1, const int* off_a,
    const int* off_b, double* schwarz_bound);
int fj_dq_1_3(fj_stream_t stream, int threads, int n_rec, int n_pair,
    const double* eta, const double* XPA, const double* XPB, const double* s,
    const double* w, const int* pair, const double* Dcd, double* D, double* Dmax);
int fj_jab_13(fj_stream_t stream, int threads, int n_rec, int n_ao, int n_rows,
    const double* zeta_ab, const double* XPA, const double* XPB, const double* s,
    const int* off_a, const int* off_b, const int* row, const double* Jp, double* J);
int fj_schwarz_22(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_ao,
    const double* zeta_ab, const double* P, const double* XPA, const double* XPB,
    const double* s_ab, const int* rec0, const int* rec1, const int* off_a,
    const int* off_b, double* schwarz_bound);
int fj_dq_2_2(fj_stream_t stream, int threads, int n_rec, int n_pair,
    const double* eta, const double* XPA, const double* XPB, const double* s,
    const double* w, const int* pair, const double* Dcd, double* D, double* Dmax);
int fj_jab_22(fj_stream_t stream, int threads, int n_rec, int n_ao, int n_rows,
    const double* zeta_ab, const double* XPA, const double* XPB, const double* s,
    const int* off_a, const int* off_b, const int* row, const double* Jp, double* J);
int fj_schwarz_23(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_ao,
    const double* zeta_ab, const double* P, const double* XPA, const double* XPB,
    const double* s_ab, const int* rec0, const int* rec1, const int* off_a,
    const int* off_b, double* schwarz_bound);
int fj_dq_2_3(fj_stream_t stream, int threads, int n_rec, int n_pair,
    const double* eta, const double* XPA, const double* XPB, const double* s,
    const double* w, const int* pair, const double* Dcd, double* D, double* Dmax);
int fj_jab_23(fj_stream_t stream, int threads, int n_rec, int n_ao, int n_rows,
    const double* zeta_ab, const double* XPA, const double* XPB, const double* s,
    const int* off_a, const int* off_b, const int* row, const double* Jp, double* J);
int fj_schwarz_33(fj_stream_t stream, int threads, int n_pair, int n_rec, int n_ao,
    const double* zeta_ab, const double* P, const double* XPA, const double* XPB,
    const double* s_ab, const int* rec0, const int* rec1, const int* off_a,
    const int* off_b, double* schwarz_bound);
int fj_dq_3_3(fj_stream_t stream, int threads, int n_rec, int n_pair,
    const double* eta, const double* XPA, const double* XPB, const double* s,
    const double* w, const int* pair, const double* Dcd, double* D, double* Dmax);
int fj_jab_33(fj_stream_t stream, int threads, int n_rec, int n_ao, int n_rows,
    const double* zeta_ab, const double* XPA, const double* XPB, const double* s,
    const int* off_a, const int* off_b, const int* row, const double* Jp, double* J);
int fj_pq_L0_p0_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L0_p0_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L1_p0_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L1_p0_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L2_p0_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L2_p0_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L3_p0_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L3_p0_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L4_p0_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L4_p0_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L5_p0_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L5_p0_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L6_p0_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L6_p0_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L1_p1_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L1_p1_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L2_p1_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L2_p1_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L3_p1_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L3_p1_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L4_p1_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L4_p1_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L5_p1_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L5_p1_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L6_p1_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L6_p1_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L7_p1_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L7_p1_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L2_p2_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L2_p2_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L3_p2_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L3_p2_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L4_p2_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L4_p2_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L5_p2_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L5_p2_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L6_p2_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L6_p2_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L7_p2_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L7_p2_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L8_p2_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L8_p2_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L3_p3_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L3_p3_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L4_p3_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L4_p3_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L5_p3_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L5_p3_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L6_p3_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L6_p3_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L7_p3_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L7_p3_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L8_p3_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L8_p3_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L9_p3_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L9_p3_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L4_p4_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L4_p4_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L5_p4_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L5_p4_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L6_p4_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L6_p4_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L7_p4_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L7_p4_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L8_p4_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L8_p4_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L9_p4_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L9_p4_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L10_p4_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L10_p4_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L5_p5_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L5_p5_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L6_p5_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L6_p5_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L7_p5_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L7_p5_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L8_p5_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L8_p5_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L9_p5_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L9_p5_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L10_p5_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L10_p5_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L11_p5_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L11_p5_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L6_p6_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L6_p6_q0(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L7_p6_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L7_p6_q1(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L8_p6_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L8_p6_q2(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L9_p6_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L9_p6_q3(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L10_p6_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L10_p6_q4(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L11_p6_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L11_p6_q5(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq_L12_p6_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_pq32_L12_p6_q6(fj_stream_t stream, int threads, int n_bra, int n_ket, double tau,
    const double* P, const double* zeta, const double* Ip, const double* Q,
    const double* eta, const double* Iq, const double* Dmax, const double* D, double* J);
int fj_k_0000(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0001(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0002(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0003(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0010(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0011(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0012(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0013(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0020(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0021(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0022(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0023(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0030(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0031(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0032(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0033(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0100(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0101(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0102(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0103(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0110(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0111(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0112(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0113(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0120(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0121(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0122(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0123(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0130(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0131(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0132(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0133(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0200(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0201(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0202(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0203(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0210(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0211(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0212(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0213(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0220(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0221(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0222(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0223(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0230(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0231(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0232(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0233(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0300(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0301(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0302(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0303(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0310(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0311(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0312(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0313(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0320(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0321(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0322(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0323(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0330(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0331(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0332(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_0333(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1000(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1001(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1002(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1003(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1010(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1011(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1012(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1013(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1020(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1021(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1022(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1023(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1030(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1031(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1032(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1033(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1100(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1101(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1102(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1103(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1110(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1111(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1112(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1113(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1120(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1121(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1122(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1123(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1130(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1131(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1132(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1133(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1200(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1201(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1202(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1203(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1210(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1211(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1212(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1213(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1220(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1221(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1222(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1223(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1230(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1231(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1232(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1233(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1300(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1301(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1302(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1303(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1310(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1311(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1312(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1313(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1320(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1321(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1322(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1323(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1330(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1331(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1332(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_1333(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2000(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2001(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2002(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2003(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2010(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2011(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2012(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2013(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2020(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2021(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2022(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2023(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2030(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2031(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2032(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2033(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2100(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2101(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2102(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2103(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2110(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2111(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2112(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2113(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2120(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2121(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2122(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2123(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2130(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2131(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2132(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2133(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2200(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2201(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2202(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2203(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2210(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2211(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2212(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2213(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2220(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2221(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2222(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2223(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2230(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2231(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2232(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2233(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2300(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2301(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2302(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2303(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2310(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2311(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2312(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2313(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2320(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2321(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2322(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2323(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2330(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2331(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2332(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_2333(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3000(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3001(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3002(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3003(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3010(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3011(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3012(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3013(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3020(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3021(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3022(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3023(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3030(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3031(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3032(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3033(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3100(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3101(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3102(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3103(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3110(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3111(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3112(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3113(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3120(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3121(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3122(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3123(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3130(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3131(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3132(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3133(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3200(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3201(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3202(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3203(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3210(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3211(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3212(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3213(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3220(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3221(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3222(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3223(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3230(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3231(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3232(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3233(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3300(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3301(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3302(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3303(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3310(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3311(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3312(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3313(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3320(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3321(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3322(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3323(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3330(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3331(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3332(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k_3333(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0000(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0001(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0002(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0010(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0011(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0012(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0020(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0021(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0022(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0100(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0101(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0102(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0110(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0111(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0112(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0120(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0121(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0122(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0200(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0201(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0202(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0210(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0211(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0212(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0220(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0221(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_0222(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1000(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1001(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1002(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1010(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1011(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1012(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1020(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1021(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1022(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1100(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1101(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1102(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1110(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1111(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1112(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1120(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1121(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1122(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1200(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1201(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1202(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1210(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1211(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1212(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1220(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1221(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_1222(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2000(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2001(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2002(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2010(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2011(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2012(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2020(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2021(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2022(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2100(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2101(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2102(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2110(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2111(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2112(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2120(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2121(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2122(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2200(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2201(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2202(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2210(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2211(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2212(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2220(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2221(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
int fj_k32_2222(fj_stream_t stream, int threads, int n_bra, int n_ket, int n_dc,
    double tau, double maxD, const double* P, const double* zeta_ac, const double* XPA,
    const double* XPC, const double* s_ac, const double* Ip, const int* dc,
    const double* Q, const double* eta, const double* XQB, const double* XQD,
    const double* s_bd, const double* Iq, const int* dd, const double* Dcd, double* K);
/* END MMD CLASSES */

#ifdef __cplusplus
}
#endif
#endif
