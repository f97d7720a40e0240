"""HIP generator for the fused Fock kernels, gfx950 only (wave64).

One translation unit per class ``(la lb|lc ld)`` (AO order, la <= lb,
lc <= ld, bra pair <= ket pair), written to ``csrc/gen/f_<la><lb><lc><ld>.hip``.
Each unit holds

* ``fj_q_<cls><R>``: the body for one contracted quartet -- the K_CD x K_AB
  primitive nest (quartet scalars, Boys, HGP VRR, boundary accumulation),
  the HRR on the contracted boundary and the six-block digestion -- emitted
  straight-line from the DAGs of ``hgp.py``, templated on the VRR/HRR real
  type R (double, or float for the FP32 tier; Boys, the seed, the weighted
  densities and every F accumulation stay double);
* the kernel ``fj_f_<cls>_k<R>``: one workgroup per bra pair (its primitive
  records staged once in LDS with 1/p precomputed), one lane per ket pair,
  the canonical half and the six-block screen as per-lane predicates, the
  bra block ``F_ab`` reduced across the wavefront before its atomics;
* the C-ABI: ``fj_f_<cls>`` with the upstream per-class parameter list
  (upstream `eri/fock/levers/seams.py:228-235,739-770`) and ``fj_fx_<cls>``,
  the ragged-degree, precision-selectable entry the native Fock builder
  uses.

The semantics (what is computed, weights, degeneracy, screen) are upstream
`eri/fock/kernel_f.py:1-165`; the code layout is this generator's own.
"""

from __future__ import annotations

from pathlib import Path

from .basis import cartesians, n_cart
from .hgp import ZERO, canonical, hrr_dag, vrr_dag

HERE = Path(__file__).resolve().parent
GEN = HERE / "csrc" / "gen"

MAX_PRIM_PAIRS = 128   # staged bra records per workgroup (upstream cap)
BOYS_H = 0.125         # Boys table spacing (exact in binary)
BOYS_TMAX = 40.0       # table range; above it, the asymptotic branch
BOYS_MTAB = 20         # orders tabulated: L <= 12 plus 8 more for the 9-term Taylor
BOYS_ROW = 24          # doubles per row: e^-T, F_0..F_MTAB, padding


def cls_name(cls) -> str:
    return "".join(str(x) for x in cls)


def _fidx(c) -> int:
    """index of a Cartesian vector in the table of every |v| <= n"""
    n = sum(c)
    return n * (n + 1) * (n + 2) // 6 + cartesians(n).index(tuple(c))


class _Emit:
    def __init__(self):
        self.lines: list[str] = []

    def __call__(self, s: str = "", ind: int = 1):
        self.lines.append("    " * ind + s)

    def text(self) -> str:
        return "\n".join(self.lines) + "\n"


def _vname(node) -> str:
    e, f, m = node
    return f"v{_fidx(e)}_{_fidx(f)}_{m}"


def _hname(node) -> str:
    a, b, c, d = node
    return f"h{_fidx(a)}_{_fidx(b)}_{_fidx(c)}_{_fidx(d)}"


AX = "xyz"


# The split digestion (fj_hj_ / fj_hk_): the output block each density block
# serves in the six-block form, and the output blocks' weights in the split
# form (s_<block>, one of them 1 per launch, the others 0).
SERVES = {"cd": "ab", "ab": "cd", "ac": "bd", "ad": "bc", "bc": "ad", "bd": "ac"}
SPLIT_WEIGHTS = ("    const double s_ab = dg.w[0], s_cd = dg.w[1], s_ac = dg.w[2], s_ad = dg.w[3],\n"
                 "                 s_bc = dg.w[4], s_bd = dg.w[5];")


def _dw(split: bool, dens: str) -> str:
    """The weight a density block is multiplied by: s_j (Coulomb) or s_k
    (exchange) in the fused form, its output block's s_<block> in the split
    form."""
    if split:
        return f"s_{SERVES[dens]}"
    return "s_j" if dens in ("cd", "ab") else "s_k"


def _didx(split: bool, r: str, c: str) -> str:
    """D's element (r, c): row-major, or through the split launch's strides
    (a transposed density block)."""
    return f"({r}) * dg.dr + ({c}) * dg.dc" if split else f"({r}) * n_ao + {c}"


def _fidx_out(split: bool, r: str, c: str) -> str:
    """F's element (r, c), likewise for the split launch's output block."""
    return f"({r}) * dg.fr + ({c}) * dg.fc" if split else f"({r}) * n_ao + {c}"


def emit_quartet(cls, split: bool = False) -> str:
    """The device function for one contracted quartet of class `cls`
    (split: the one-block digestion of the split J / K kernels, fj_qs_)."""
    la, lb, lc, ld = cls
    L = sum(cls)
    v = vrr_dag(tuple(cls))
    order, terms, targets, boundary = hrr_dag(tuple(cls))
    name = cls_name(cls)
    na, nb, nc, nd = (n_cart(x) for x in cls)
    # which quartet factors the VRR reads
    used = set()
    for n in v.order:
        r = v.terms[n]
        if r:
            d, ts = r
            for t in ts:
                used.add((t.coef, d if t.coef in ("PA", "WP", "QC", "WQ") else None))
    need = lambda c: any(u[0] == c for u in used)  # noqa: E731
    hrr_bra = any(r and r[0] == "bra" for r in terms.values())
    hrr_ket = any(r and r[0] == "ket" for r in terms.values())
    acc = {ef: f"c{i}" for i, ef in enumerate(v.targets)}

    E = _Emit()
    E(f"// ({la}{lb}|{lc}{ld}): {v.n_nodes} VRR nodes, {len(v.targets)} boundary "
      f"values, {len(order)} HRR statements, {len(targets)} ERIs", 0)
    E("template <typename R>", 0)
    E(f"__device__ __forceinline__ void fj_q{'s' if split else ''}_{name}(", 0)
    E("const double* __restrict__ bp, int k_ab,       // bra records, LDS, FJ_SREC doubles each")
    E("const double* __restrict__ kp, int k_cd,       // ket records, global, 8 doubles each")
    E("double ABx, double ABy, double ABz, double CDx, double CDy, double CDz,")
    E("const double* __restrict__ D, int n_ao, int oc, int od, int oa, int ob,")
    E("const FjDig& dg, const double* __restrict__ wab," if split else
      "double s_j, double s_k, const double* __restrict__ wab,")
    E("double* __restrict__ fab, double* __restrict__ F, double* __restrict__ ws, size_t st) {", 1)
    if split:
        E(SPLIT_WEIGHTS, 0)
    for i in range(len(v.targets)):
        E(f"R c{i} = R(0);")
    E("for (int kc = 0; kc < k_cd; ++kc) {")
    E("const double* __restrict__ q_ = kp + 8 * kc;", 2)
    E("const double Qx = q_[0], Qy = q_[1], Qz = q_[2], q = q_[6], Uq = q_[7];", 2)
    for d in range(3):
        if ("QC", d) in used:
            E(f"const R QC{AX[d]} = R(q_[{3 + d}]);", 2)
    if need("i2q") or need("i2q_r"):
        E("const double iq = 1.0 / q;", 2)
    E("for (int kb = 0; kb < k_ab; ++kb) {", 2)
    E("const double* __restrict__ p_ = bp + FJ_SREC * kb;", 3)
    E("const double PQx = p_[0] - Qx, PQy = p_[1] - Qy, PQz = p_[2] - Qz;", 3)
    E("const double p = p_[6];", 3)
    E("const double ippq = 1.0 / (p + q);", 3)
    E("const double rho = p * q * ippq;", 3)
    E("const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);", 3)
    E(f"double Fm[{L + 1}];", 3)
    E(f"fj_boys<{L}>(T, Fm);", 3)
    E("const double S = p_[7] * Uq * (FJ_TWO_OVER_SQRT_PI * sqrt(rho));", 3)
    for d in range(3):
        if ("PA", d) in used:
            E(f"const R PA{AX[d]} = R(p_[{3 + d}]);", 3)
        if ("WP", d) in used:
            E(f"const R WP{AX[d]} = R(-q * ippq * PQ{AX[d]});", 3)
        if ("WQ", d) in used:
            E(f"const R WQ{AX[d]} = R(p * ippq * PQ{AX[d]});", 3)
    if need("i2p"):
        E("const R i2p = R(0.5 * p_[8]);", 3)
    if need("i2p_r"):
        E("const R i2p_r = R(0.5 * p_[8] * rho * p_[8]);", 3)
    if need("i2q"):
        E("const R i2q = R(0.5 * iq);", 3)
    if need("i2q_r"):
        E("const R i2q_r = R(0.5 * iq * rho * iq);", 3)
    if need("i2pq"):
        E("const R i2pq = R(0.5 * ippq);", 3)
    # VRR, in build order
    for n in v.order:
        r = v.terms[n]
        if not r:
            E(f"const R {_vname(n)} = R(S * Fm[{n[2]}]);", 3)
            continue
        d, ts = r
        parts = []
        for t in ts:
            c = t.coef
            ch = _vname(t.child)
            if c in ("PA", "WP", "QC", "WQ"):
                parts.append(f"{c}{AX[d]} * {ch}")
            else:
                m = t.mult
                sign = "-" if m < 0 else "+"
                mag = abs(m)
                f = c if mag == 1 else f"(R({mag}) * {c})"
                parts.append(f"{sign} {f} * {ch}")
        expr = parts[0]
        for p in parts[1:]:
            expr += (" " + p) if p[0] in "+-" else (" + " + p)
        E(f"const R {_vname(n)} = {expr};", 3)
    for ef, cname in acc.items():
        E(f"{cname} += {_vname((ef[0], ef[1], 0))};", 3)
    E("}", 2)
    E("}")
    # HRR (bra first in build order, then ket)
    for n in order:
        r = terms[n]
        if not r:
            E(f"const double {_hname(n)} = double({acc[(n[0], n[2])]});")
            continue
        side, k, moved, lowered = r
        sep = ("AB" if side == "bra" else "CD") + AX[k]
        E(f"const double {_hname(n)} = {_hname(moved)} + {sep} * {_hname(lowered)};")
    # digestion: five ket-dependent blocks here, f_ab into the caller's buffer
    ca, cb, cc, cd = (cartesians(x) for x in cls)
    blocks = {"cd": (2, 3), "ac": (0, 2), "ad": (0, 3), "bc": (1, 2), "bd": (1, 3)}
    comp = {"ab": (2, 3), "cd": (0, 1), "ac": (1, 3), "ad": (1, 2), "bc": (0, 3), "bd": (0, 2)}
    sizes = (na, nb, nc, nd)
    aov = ("oa", "ob", "oc", "od")
    # weighted densities: w_cd (for F_ab) and the four exchange ones; w_ab comes in
    E("// the weighted densities (weight x degeneracy folded in once)")
    for blk in ("cd", "ac", "ad", "bc", "bd"):
        p0, p1 = blocks[blk]
        sc = _dw(split, blk)
        for i in range(sizes[p0]):
            for j in range(sizes[p1]):
                E(f"const double w{blk}{i}_{j} = {sc} * "
                  f"D[{_didx(split, f'{aov[p0]} + {i}', f'{aov[p1]} + {j}')}];")
    for blk in ("cd", "ac", "ad", "bc", "bd"):
        p0, p1 = blocks[blk]
        for i in range(sizes[p0]):
            for j in range(sizes[p1]):
                E(f"double f{blk}{i}_{j} = 0.0;")
    for t in targets:
        idx = tuple(cs.index(x) for cs, x in zip((ca, cb, cc, cd), t))
        E(f"{{ const double x = {_hname(t)};")
        for blk in ("ab", "cd", "ac", "ad", "bc", "bd"):
            p0, p1 = (0, 1) if blk == "ab" else blocks[blk]
            q0, q1 = comp[blk]
            if (q0, q1) == (0, 1):
                w = f"wab[{idx[0] * nb + idx[1]}]"
            else:
                w = f"w{_blkname(q0, q1)}{idx[q0]}_{idx[q1]}"
            if blk == "ab":
                E(f"  fab[{idx[0] * nb + idx[1]}] += {w} * x;")
            else:
                E(f"  f{blk}{idx[p0]}_{idx[p1]} += {w} * x;")
        E("}")
    for blk in ("cd", "ac", "ad", "bc", "bd"):
        p0, p1 = blocks[blk]
        if split:
            E(f"if (s_{blk} != 0.0) {{")
        for i in range(sizes[p0]):
            for j in range(sizes[p1]):
                E(f"unsafeAtomicAdd(F + {_fidx_out(split, f'{aov[p0]} + {i}', f'{aov[p1]} + {j}')}, "
                  f"f{blk}{i}_{j});", 2 if split else 1)
        if split:
            E("}")
    E("}", 0)
    return E.text()


def _blkname(p0: int, p1: int) -> str:
    return "abcd"[p0] + "abcd"[p1]


# Waves per SIMD to ask of the mixed-tier kernel (both instantiations in one
# kernel take the larger register count), for the classes whose FP64 kernel
# alone reaches more than the mixed one did unasked (hipcc's resource
# remarks for gfx950: e.g. (ss|dd) 2 vs 1, (sp|sp) 3 vs 2).
MIXED_MIN_WAVES = {(0, 0, 2, 2): 2, (0, 0, 1, 2): 3, (0, 1, 0, 1): 3, (0, 0, 1, 1): 4,
                   (0, 0, 0, 2): 4, (1, 0, 2, 0): 3, (0, 0, 0, 0): 8}


def emit_class(cls) -> str:
    la, lb, lc, ld = cls
    name = cls_name(cls)
    na, nb = n_cart(la), n_cart(lb)
    nab = na * nb
    out = [f"// generated by felizjueves_amd/gen_hip.py -- class ({la}{lb}|{lc}{ld})",
           "#pragma once", '#include "../fj_device.h"', ""]
    out.append(emit_quartet_staged(cls) if staged(cls) else emit_quartet(cls))
    abx, aby, abz = (("a.bra.sep[3 * bra]", "a.bra.sep[3 * bra + 1]", "a.bra.sep[3 * bra + 2]")
                     if lb else ("0.0",) * 3)
    cdx, cdy, cdz = (("a.ket.sep[3 * k]", "a.ket.sep[3 * k + 1]", "a.ket.sep[3 * k + 2]")
                     if ld else ("0.0",) * 3)
    W = ws_doubles(cls)
    if W:
        kernel_body = (
            "    // the Boys table columns of this class, staged once per workgroup\n"
            "    __shared__ double tb[FJ_BOYS_NROW * FJ_BOYS_LROW];\n"
            f"    fj_boys_stage<{sum(cls)}>(tb);\n"
            "    __syncthreads();\n"
            "    // persistent: one workspace slot per lane of the grid. The items are\n"
            "    // listed costliest first and dealt in snake order (round r takes\n"
            "    // item r G + b, or r G + G - 1 - b in odd rounds), so a workgroup\n"
            "    // that drew a costly item draws a cheap one next (greedy balance\n"
            "    // without atomics); the choice is uniform over the workgroup\n"
            "    const size_t st = (size_t)gridDim.x * blockDim.x;\n"
            "    double* ws = a.ws + (size_t)blockIdx.x * blockDim.x + threadIdx.x;\n"
            "    const int G = (int)gridDim.x, b = (int)blockIdx.x;\n"
            "    for (int r = 0; r * G < a.n_items; ++r) {\n"
            "        const int it = r * G + ((r & 1) ? G - 1 - b : b);\n"
            "        if (it >= a.n_items) continue;   // uniform: a partial last round\n"
            "        // an empty asm that 'redefines' the slot pointer and stride each\n"
            "        // item, so the compiler recomputes the workspace addresses where\n"
            "        // they are used instead of hoisting thousands of them out of the\n"
            "        // item loop (which spills)\n"
            "        double* wsi = ws;\n"
            "        size_t sti = st;\n"
            "        FJ_LAUNDER(wsi, sti);\n"
            f"        fj_item_{name}<R>(a, it % a.n_bx, it / a.n_bx, bsh, wsi, sti, tb);\n"
            "        __syncthreads();\n"
            "    }")
        launch_ws = (
            f"    // workspace: {W} doubles per lane; the grid is cut to the slots it holds\n"
            f"    {{\n"
            f"        const long long slots = a.ws ? a.ws_doubles / {W} : 0;\n"
            f"        const long long blocks = slots / threads;\n"
            f"        if (blocks < 1) return FJ_ERR_WORKSPACE;\n"
            f"        grid = dim3((unsigned)(blocks < (long long)a.n_items ? blocks : a.n_items));\n"
            f"    }}")
        f_ws = (f"    {{\n"
                f"        const long long need = (long long)8 * {W} * threads * FJ_LIB_WS_BLOCKS;\n"
                f"        a.ws = fj_library_workspace(s, need);\n"
                f"        if (!a.ws) return FJ_ERR_WORKSPACE;\n"
                f"        a.ws_doubles = need / 8;\n"
                f"    }}")
    else:
        kernel_body = "    fj_item_" + name + "<R>(a, (int)blockIdx.x, (int)blockIdx.y, bsh, nullptr, 0, nullptr);"
        launch_ws = ""
        f_ws = ""
    # the mixed kernel: the same body, each work item in its tier (the
    # per-item tier array, else items [0, n_fp64) FP64 and the rest FP32;
    # uniform over the workgroup)
    mixed_body = (kernel_body
                  .replace(f"        fj_item_{name}<R>(a, it % a.n_bx, it / a.n_bx, bsh, wsi, sti, tb);\n",
                           f"        if (fj_item_fp64(a, it % a.n_bx))\n"
                           f"            fj_item_{name}<double>(a, it % a.n_bx, it / a.n_bx, bsh, wsi, sti, tb);\n"
                           f"        else\n"
                           f"            fj_item_{name}<float>(a, it % a.n_bx, it / a.n_bx, bsh, wsi, sti, tb);\n")
                  .replace("    fj_item_" + name + "<R>(a, (int)blockIdx.x, (int)blockIdx.y, bsh, nullptr, 0, nullptr);",
                           "    if (fj_item_fp64(a, (int)blockIdx.x))\n"
                           "        fj_item_" + name + "<double>(a, (int)blockIdx.x, (int)blockIdx.y, bsh, nullptr, 0, nullptr);\n"
                           "    else\n"
                           "        fj_item_" + name + "<float>(a, (int)blockIdx.x, (int)blockIdx.y, bsh, nullptr, 0, nullptr);"))
    assert mixed_body != kernel_body
    w = MIXED_MIN_WAVES.get(tuple(cls))
    mixed_waves = f", {w}" if w else ""
    out.append(f"""
// One work item: workgroup-uniform early returns only (the persistent
// kernel calls it in a loop with a barrier after each item).
template <typename R>
__device__ __forceinline__ void fj_item_{name}(const FjFArgs& a, int bx, int by, double* bsh,
                                            double* __restrict__ ws, size_t st,
                                            const double* __restrict__ tb) {{
    // work item (bra pair, first ket, system, ket end) from the list, else
    // (x, y * blockDim, the single system, n_ket)
    // The third field packs the system (bits 0-15) and a chunk of the bra
    // primitive pairs: first (bits 16-23) and count (bits 24-31, 0 = all).
    // A ket-list entry packs the ket pair (bits 0-16) and a chunk of its
    // primitive pairs: first (bits 17-23) and count (bits 24-30, 0 = to the
    // end). The HRR and the digestion are linear in the contracted boundary,
    // so a quartet split over primitive chunks sums exactly through F's atomics.
    int bra, k0, k1, kb_lo = 0, kb_n = 0;
    FjSys S;
    if (a.work) {{
        const int4 w = a.work[bx];
        bra = w.x; k0 = w.y; k1 = w.w;
        S = a.sys[w.z & 0xffff];
        kb_lo = (w.z >> 16) & 0xff;
        kb_n = (w.z >> 24) & 0xff;
    }} else {{
        bra = bx;
        k0 = by * (int)blockDim.x;
        S = a.one;
        k1 = a.ket.n;
    }}
    const int kmin = a.diag * bra;
    if (k0 >= k1 || (!a.ket_list && k0 + (int)blockDim.x <= kmin)) return;   // nothing for this block
    int kb0, kab;
    fj_side_range(a.bra, bra, kb0, kab);
    if (kab > FJ_MAX_PRIM_PAIRS) return;   // the host refuses such a pair; never overrun the staging
    if (kb_lo >= kab) return;              // an empty bra chunk (the host emits none)
    kb0 += kb_lo;
    kab = kb_n ? min(kab - kb_lo, kb_n) : kab - kb_lo;
    for (int i = threadIdx.x; i < kab; i += blockDim.x) {{
        const double* __restrict__ r = a.bra.prim + 8 * (size_t)(kb0 + i);
        double* s = bsh + FJ_SREC * i;
#pragma unroll
        for (int f = 0; f < 8; ++f) s[f] = r[f];
        s[8] = 1.0 / r[6];
    }}
    __syncthreads();
    const int sa = a.bra.sh[2 * bra], sb = a.bra.sh[2 * bra + 1];
    const int oa = a.bra.ao[2 * bra], ob = a.bra.ao[2 * bra + 1];
    const double Gb = a.bra.G[bra];
    const double ABx = {abx}, ABy = {aby}, ABz = {abz};
    const int n_ao = S.n_ao;
    // with a ket list the work item names list[k0 .. k1), the live kets only
    const int kk = k0 + (int)threadIdx.x;
    int k = kk, kc_lo = 0, kc_n = 0;
    if (a.ket_list) {{
        const int v = kk < k1 ? a.ket_list[kk] : 0;
        k = v & 0x1ffff;
        kc_lo = (v >> 17) & 0x7f;
        kc_n = (v >> 24) & 0x7f;
    }}
    double fab[{nab}];
    double wab[{nab}];
#pragma unroll
    for (int i = 0; i < {nab}; ++i) fab[i] = 0.0;
    bool live = kk < k1 && k >= kmin;
    int sc = 0, sd = 0;
    if (live) {{
        sc = a.ket.sh[2 * k]; sd = a.ket.sh[2 * k + 1];
        const double* __restrict__ Dsh = S.Dsh;
        const int ns = S.n_sh;
        double dmax = 2.0 * Dsh[sa * ns + sb];
        dmax = fmax(dmax, 2.0 * Dsh[sc * ns + sd]);
        dmax = fmax(dmax, 0.5 * Dsh[sa * ns + sc]);
        dmax = fmax(dmax, 0.5 * Dsh[sa * ns + sd]);
        dmax = fmax(dmax, 0.5 * Dsh[sb * ns + sc]);
        dmax = fmax(dmax, 0.5 * Dsh[sb * ns + sd]);
        live = !(Gb * (a.ket.G[k] * dmax) < a.tau);
        int kc0_, kcd_;
        fj_side_range(a.ket, k, kc0_, kcd_);
        if (kcd_ <= kc_lo) live = false;   // this lane's ket has no primitives in the chunk
    }}
    if (live) {{
        double deg = 1.0;
        if (sa == sb) deg *= 0.5;
        if (sc == sd) deg *= 0.5;
        if (a.diag && k == bra) deg *= 0.5;
        const double s_j = 2.0 * deg, s_k = -0.5 * deg;
        const int oc = a.ket.ao[2 * k], od = a.ket.ao[2 * k + 1];
#pragma unroll
        for (int i = 0; i < {na}; ++i)
#pragma unroll
            for (int j = 0; j < {nb}; ++j) wab[i * {nb} + j] = s_j * S.D[(oa + i) * n_ao + ob + j];
        int kc0, kcd;
        fj_side_range(a.ket, k, kc0, kcd);
        kc0 += kc_lo;
        kcd = kc_n ? min(kcd - kc_lo, kc_n) : kcd - kc_lo;
        const double CDx = {cdx}, CDy = {cdy}, CDz = {cdz};
        fj_q_{name}<R>(bsh, kab, a.ket.prim + 8 * (size_t)kc0, kcd, ABx, ABy, ABz, CDx, CDy, CDz,
                        S.D, n_ao, oc, od, oa, ob, s_j, s_k, wab, fab, S.F, ws, st{", tb" if W else ""});
    }}
    // F_ab: one bra per workgroup, so every lane of a wavefront adds into the
    // same block; reduce across the wavefront, then one atomic per element.
#pragma unroll
    for (int i = 0; i < {nab}; ++i) {{
        const double s = fj_wave_sum(fab[i]);
        if ((threadIdx.x & 63) == 0 && s != 0.0)
            unsafeAtomicAdd(S.F + (oa + i / {nb}) * n_ao + ob + i % {nb}, s);
    }}
}}

template <typename R>
__global__ void __launch_bounds__(FJ_MAX_THREADS) fj_f_{name}_k(FjFArgs a) {{
    __shared__ double bsh[FJ_MAX_PRIM_PAIRS * FJ_SREC];
{kernel_body}
}}

// the FP32-tier instantiation and the mixed kernel live in their own
// translation units (f_{name}_r32.hip, f_{name}_mx.hip)
int fj_launch_{name}_r32(hipStream_t s, dim3 grid, dim3 block, const FjFArgs& a);
int fj_launch_{name}_mx(hipStream_t s, dim3 grid, dim3 block, const FjFArgs& a);
//@@MIXED@@
__global__ void __launch_bounds__(FJ_MAX_THREADS{mixed_waves}) fj_f_{name}_km(FjFArgs a) {{
    __shared__ double bsh[FJ_MAX_PRIM_PAIRS * FJ_SREC];
{mixed_body}
}}
//@@SPLIT@@
static int fj_launch_{name}(hipStream_t s, int threads, int precision, FjFArgs a, int n_blocks) {{
    if (threads <= 0) threads = FJ_DEFAULT_THREADS;
    if (threads % 64 || threads > FJ_MAX_THREADS) return FJ_ERR_THREADS;
    if (precision < 0 || precision > 2) return FJ_ERR_ARG;
    if (n_blocks <= 0 || a.ket.n <= 0) return 0;
    dim3 grid((unsigned)n_blocks, a.work ? 1u : (unsigned)((a.ket.n + threads - 1) / threads));
    a.n_bx = (int)grid.x;
    a.n_items = (int)(grid.x * grid.y);
{launch_ws}
    if (precision == 0) {{
        hipLaunchKernelGGL(fj_f_{name}_k<double>, grid, dim3(threads), 0, s, a);
        return (int)hipGetLastError();
    }}
    if (precision == 2) return fj_launch_{name}_mx(s, grid, dim3(threads), a);
    return fj_launch_{name}_r32(s, grid, dim3(threads), a);
}}

//@@UPSTREAM@@
// The upstream-ABI launch with every argument (AB / CD may be NULL when lb /
// ld is 0) and an optional workspace: ws == NULL takes the library-owned one
// of stream s (allocated on first use), else the persistent grid is cut to
// the slots ws holds and nothing is allocated.
int fj_f_any_{name}(hipStream_t s, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
                    int n_sh, int n_ao, int diag, double tau,
                    const double* bra_prim, const double* AB, const double* G_ab, const int* bra_sh,
                    const int* bra_ao, const double* ket_prim, const double* CD, const double* G_cd,
                    const int* ket_sh, const int* ket_ao, const double* Dsh, const double* D, double* F,
                    double* ws, long long ws_bytes) {{
    if (k_ab < 1 || k_ab > FJ_MAX_PRIM_PAIRS || k_cd < 1 || k_cd > FJ_MAX_PRIM_PAIRS) return FJ_ERR_DEGREE;
    if (n_bra < 0 || n_ket < 0 || (diag && n_bra != n_ket)) return FJ_ERR_SHAPE;
    FjFArgs a;
    a.bra = FjSide{{bra_prim, nullptr, k_ab, {"AB" if lb else "nullptr"}, G_ab, bra_sh, bra_ao, n_bra}};
    a.ket = FjSide{{ket_prim, nullptr, k_cd, {"CD" if ld else "nullptr"}, G_cd, ket_sh, ket_ao, n_ket}};
    a.work = nullptr;
    a.ket_list = nullptr;
    a.sys = nullptr;
    a.one = FjSys{{D, Dsh, F, n_ao, n_sh}};
    a.diag = diag; a.tau = tau;
    a.ws = ws; a.ws_doubles = ws ? ws_bytes / 8 : 0;
    a.n_fp64 = 0; a.tier = nullptr;
    if (!ws) {{
{f_ws}
    }}
    return fj_launch_{name}(s, threads, 0, a, n_bra);
}}

extern "C" int fj_f_{name}(hipStream_t s, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
                        int n_sh, int n_ao, int diag, double tau,
                        const double* bra_prim,{" const double* AB," if lb else ""} const double* G_ab, const int* bra_sh, const int* bra_ao,
                        const double* ket_prim,{" const double* CD," if ld else ""} const double* G_cd, const int* ket_sh, const int* ket_ao,
                        const double* Dsh, const double* D, double* F) {{
    return fj_f_any_{name}(s, threads, n_bra, n_ket, k_ab, k_cd, n_sh, n_ao, diag, tau, bra_prim,
                           {"AB" if lb else "nullptr"}, G_ab, bra_sh, bra_ao, ket_prim, {"CD" if ld else "nullptr"}, G_cd, ket_sh, ket_ao,
                           Dsh, D, F, nullptr, 0);
}}
//@@END_UPSTREAM@@

extern "C" int fj_fx_{name}(hipStream_t s, int threads, int precision,
                         const FjSide* bra, const FjSide* ket, const int* work, int n_work,
                         const int* ket_list, const FjSys* sys, int diag, double tau,
                         double* ws, long long ws_bytes) {{
    if (!bra || !ket || !work || !sys || ws_bytes < 0) return FJ_ERR_ARG;
    FjFArgs a;
    a.bra = *bra; a.ket = *ket;
    a.work = reinterpret_cast<const int4*>(work);
    a.ket_list = ket_list;
    a.sys = sys;
    a.one = FjSys{{nullptr, nullptr, nullptr, 0, 0}};
    a.diag = diag; a.tau = tau;
    a.ws = ws; a.ws_doubles = ws ? ws_bytes / 8 : 0;
    a.n_fp64 = 0; a.tier = nullptr;
    if (precision != 0 && precision != 1) return FJ_ERR_ARG;
    return fj_launch_{name}(s, threads, precision, a, n_work);
}}

extern "C" int fj_fxm_{name}(hipStream_t s, int threads, const FjSide* bra, const FjSide* ket,
                          const int* work, int n_work, int n_fp64, const int* tier,
                          const int* ket_list, const FjSys* sys, int diag, double tau, double* ws,
                          long long ws_bytes) {{
    if (!bra || !ket || !work || !sys || ws_bytes < 0) return FJ_ERR_ARG;
    if (n_fp64 < 0 || n_fp64 > n_work) return FJ_ERR_ARG;
    FjFArgs a;
    a.bra = *bra; a.ket = *ket;
    a.work = reinterpret_cast<const int4*>(work);
    a.ket_list = ket_list;
    a.sys = sys;
    a.one = FjSys{{nullptr, nullptr, nullptr, 0, 0}};
    a.diag = diag; a.tau = tau;
    a.ws = ws; a.ws_doubles = ws ? ws_bytes / 8 : 0;
    a.n_fp64 = n_fp64; a.tier = tier;
    return fj_launch_{name}(s, threads, 2, a, n_work);
}}
""")
    text = "\n".join(out)
    i, j = text.index("//@@UPSTREAM@@\n"), text.index("//@@END_UPSTREAM@@\n")
    if canonical(cls):
        # the upstream-ABI entry (walker launches use canonical classes only)
        return text[:i] + text[i + len("//@@UPSTREAM@@\n"):j] + text[j + len("//@@END_UPSTREAM@@\n"):]
    return text[:i] + text[j + len("//@@END_UPSTREAM@@\n"):]


# ---------------------------------------------------------------------------
# The split J and K kernels (upstream eri/hjk/kernel_hj.py, kernel_hk.py): one
# kernel per orbit of the 8-fold ERI symmetry, built on the orbit's
# representative (orbits.representative). A launch is one ordered batch pair
# of the split stream routed to the representative through its orientation
# `perm` (an index into orbits.GROUP): the kernel runs the representative's
# VRR and HRR (the fused kernels' bodies) and deposits ONE block, J_ab +=
# D_cd (ab|cd) or K_ac += D_bd (ab|cd) on the ordered quartet's positions,
# carried to the built quartet's by the orientation (fj_split_dig): no
# weights, no degeneracy, no canonical half, one-block density screen.

def emit_split_class(rep, mask: int) -> str:
    """hs_<rep>.hip: the split kernel of representative class `rep` and its
    entries fj_hj_<rep> and fj_hk_<rep>; `mask` has bit g set for every
    orientation the orbit's members reach (other perms deposit nothing)."""
    la, lb, lc, ld = rep
    name = cls_name(rep)
    na, nb = n_cart(la), n_cart(lb)
    nab = na * nb
    W = ws_doubles(rep)
    body = emit_quartet_staged(rep, split=True) if staged(rep) else emit_quartet(rep, split=True)
    abx, aby, abz = (("a.bra.sep[3 * bra]", "a.bra.sep[3 * bra + 1]", "a.bra.sep[3 * bra + 2]")
                     if lb else ("0.0",) * 3)
    cdx, cdy, cdz = (("a.ket.sep[3 * k]", "a.ket.sep[3 * k + 1]", "a.ket.sep[3 * k + 2]")
                     if ld else ("0.0",) * 3)
    if W:
        kernel_body = (
            "    __shared__ double tb[FJ_BOYS_NROW * FJ_BOYS_LROW];\n"
            f"    fj_boys_stage<{sum(rep)}>(tb);\n"
            "    __syncthreads();\n"
            "    const size_t st = (size_t)gridDim.x * blockDim.x;\n"
            "    double* ws = a.ws + (size_t)blockIdx.x * blockDim.x + threadIdx.x;\n"
            "    for (int it = (int)blockIdx.x; it < a.n_items; it += (int)gridDim.x) {\n"
            "        double* wsi = ws;\n"
            "        size_t sti = st;\n"
            "        FJ_LAUNDER(wsi, sti);\n"
            f"        fj_hsitem_{name}(a, dg, it % a.n_bx, it / a.n_bx, bsh, wsi, sti, tb);\n"
            "        __syncthreads();\n"
            "    }")
        launch_ws = (
            f"    {{\n"
            f"        const long long need = (long long)8 * {W} * threads * FJ_LIB_WS_BLOCKS;\n"
            f"        a.ws = fj_library_workspace(s, need);\n"
            f"        if (!a.ws) return FJ_ERR_WORKSPACE;\n"
            f"        a.ws_doubles = need / 8;\n"
            f"        const long long blocks = a.ws_doubles / {W} / threads;\n"
            f"        grid = dim3((unsigned)(blocks < (long long)a.n_items ? blocks : a.n_items));\n"
            f"    }}")
        call_tail = "S.D, n_ao, oc, od, oa, ob, dg, wab, fab, S.F, ws, st, tb"
    else:
        kernel_body = f"    fj_hsitem_{name}(a, dg, (int)blockIdx.x, (int)blockIdx.y, bsh, nullptr, 0, nullptr);"
        launch_ws = ""
        call_tail = "S.D, n_ao, oc, od, oa, ob, dg, wab, fab, S.F, ws, st"
    out = [f"// generated by felizjueves_amd/gen_hip.py -- split J / K kernel, orbit of ({la}{lb}|{lc}{ld})",
           '#include "../fj_device.h"', "", body]

    def entry(kind: str, screen: int) -> str:
        return f"""
extern "C" int fj_h{kind}_{name}(hipStream_t s, int threads, int n_bra, int n_ket, int k_ab, int k_cd,
                         int n_sh, int n_ao, int perm, double tau,
                         const double* bra_prim,{" const double* AB," if lb else ""} const double* G_ab, const int* bra_sh, const int* bra_ao,
                         const double* ket_prim,{" const double* CD," if ld else ""} const double* G_cd, const int* ket_sh, const int* ket_ao,
                         const double* Dsh, const double* D, double* {kind.upper()}) {{
    if (k_ab < 1 || k_ab > FJ_MAX_PRIM_PAIRS || k_cd < 1 || k_cd > FJ_MAX_PRIM_PAIRS) return FJ_ERR_DEGREE;
    if (n_bra < 0 || n_ket < 0) return FJ_ERR_SHAPE;
    if (perm < 0 || perm > 7) return FJ_ERR_ARG;
    if (!((FJ_HS_MASK_{name} >> perm) & 1)) return 0;   // an orientation outside the orbit: no deposit
    FjDig dg;
    const int rc = fj_split_dig({screen}, perm, n_ao, &dg);
    if (rc) return rc;
    FjFArgs a;
    a.bra = FjSide{{bra_prim, nullptr, k_ab, {"AB" if lb else "nullptr"}, G_ab, bra_sh, bra_ao, n_bra}};
    a.ket = FjSide{{ket_prim, nullptr, k_cd, {"CD" if ld else "nullptr"}, G_cd, ket_sh, ket_ao, n_ket}};
    a.work = nullptr; a.ket_list = nullptr; a.sys = nullptr;
    a.one = FjSys{{D, Dsh, {kind.upper()}, n_ao, n_sh}};
    a.diag = 0; a.tau = tau;
    a.ws = nullptr; a.ws_doubles = 0; a.n_fp64 = 0; a.tier = nullptr;
    return fj_hs_launch_{name}(s, threads, a, dg, n_bra);
}}
"""
    out.append(f"""
#define FJ_HS_MASK_{name} {mask}

// One ordered quartet per lane: the bra pair's records staged in LDS, the
// one-block density screen, the representative's quartet body with the
// launch's one-block digestion, the bra block reduced across the wavefront.
__device__ __forceinline__ void fj_hsitem_{name}(const FjFArgs& a, const FjDig& dg, int bx, int by,
                                              double* bsh, double* __restrict__ ws, size_t st,
                                              const double* __restrict__ tb) {{
    const int bra = bx;
    const int k0 = by * (int)blockDim.x;
    const FjSys S = a.one;
    const int k1 = a.ket.n;
    if (k0 >= k1) return;   // uniform
    int kb0, kab;
    fj_side_range(a.bra, bra, kb0, kab);
    if (kab > FJ_MAX_PRIM_PAIRS) return;
    for (int i = threadIdx.x; i < kab; i += blockDim.x) {{
        const double* __restrict__ r = a.bra.prim + 8 * (size_t)(kb0 + i);
        double* s = bsh + FJ_SREC * i;
#pragma unroll
        for (int f = 0; f < 8; ++f) s[f] = r[f];
        s[8] = 1.0 / r[6];
    }}
    __syncthreads();
    const int sa = a.bra.sh[2 * bra], sb = a.bra.sh[2 * bra + 1];
    const int oa = a.bra.ao[2 * bra], ob = a.bra.ao[2 * bra + 1];
    const double Gb = a.bra.G[bra];
    const double ABx = {abx}, ABy = {aby}, ABz = {abz};
    const int n_ao = S.n_ao;
    const int k = k0 + (int)threadIdx.x;
    double fab[{nab}];
    double wab[{nab}];
#pragma unroll
    for (int i = 0; i < {nab}; ++i) fab[i] = 0.0;
    bool live = k < k1;
    if (live) {{
        const int sc = a.ket.sh[2 * k], sd = a.ket.sh[2 * k + 1];
        const int s0 = dg.q0 == 0 ? sa : dg.q0 == 1 ? sb : dg.q0 == 2 ? sc : sd;
        const int s1 = dg.q1 == 0 ? sa : dg.q1 == 1 ? sb : dg.q1 == 2 ? sc : sd;
        live = !(Gb * (a.ket.G[k] * S.Dsh[s0 * S.n_sh + s1]) < a.tau);
    }}
    if (live) {{
        const int oc = a.ket.ao[2 * k], od = a.ket.ao[2 * k + 1];
#pragma unroll
        for (int i = 0; i < {na}; ++i)
#pragma unroll
            for (int j = 0; j < {nb}; ++j)
                wab[i * {nb} + j] = dg.w[1] * S.D[(oa + i) * dg.dr + (ob + j) * dg.dc];
        int kc0, kcd;
        fj_side_range(a.ket, k, kc0, kcd);
        const double CDx = {cdx}, CDy = {cdy}, CDz = {cdz};
        fj_qs_{name}<double>(bsh, kab, a.ket.prim + 8 * (size_t)kc0, kcd, ABx, ABy, ABz, CDx, CDy, CDz,
                         {call_tail});
    }}
    if (dg.w[0] != 0.0) {{   // uniform
#pragma unroll
        for (int i = 0; i < {nab}; ++i) {{
            const double s = fj_wave_sum(fab[i]);
            if ((threadIdx.x & 63) == 0 && s != 0.0)
                unsafeAtomicAdd(S.F + (oa + i / {nb}) * dg.fr + (ob + i % {nb}) * dg.fc, s);
        }}
    }}
}}

__global__ void __launch_bounds__(FJ_MAX_THREADS) fj_hs_{name}_k(FjFArgs a, FjDig dg) {{
    __shared__ double bsh[FJ_MAX_PRIM_PAIRS * FJ_SREC];
{kernel_body}
}}

static int fj_hs_launch_{name}(hipStream_t s, int threads, FjFArgs a, FjDig dg, int n_blocks) {{
    if (threads <= 0) threads = FJ_DEFAULT_THREADS;
    if (threads % 64 || threads > FJ_MAX_THREADS) return FJ_ERR_THREADS;
    if (n_blocks <= 0 || a.ket.n <= 0) return 0;
    dim3 grid((unsigned)n_blocks, (unsigned)((a.ket.n + threads - 1) / threads));
    a.n_bx = (int)grid.x;
    a.n_items = (int)(grid.x * grid.y);
{launch_ws}
    hipLaunchKernelGGL(fj_hs_{name}_k, grid, dim3(threads), 0, s, a, dg);
    return (int)hipGetLastError();
}}
""")
    out.append(entry("j", 0))
    out.append(entry("k", 1))
    return "\n".join(out)


def split_header_decls(reps) -> str:
    """The split J / K declarations of include/fj_mi355x.h."""
    out = []
    for rep in reps:
        la, lb, lc, ld = rep
        n = cls_name(rep)
        for kind in ("j", "k"):
            out.append(
                f"int fj_h{kind}_{n}(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab,"
                f" int k_cd,\n    int n_sh, int n_ao, int perm, double tau,\n"
                f"    const double* bra_prim,{' const double* AB,' if lb else ''} const double* G_ab,"
                f" const int* bra_sh, const int* bra_ao,\n"
                f"    const double* ket_prim,{' const double* CD,' if ld else ''} const double* G_cd,"
                f" const int* ket_sh, const int* ket_ao,\n"
                f"    const double* Dsh, const double* D, double* {kind.upper()});")
    return "\n".join(out) + "\n"


def generate_split(reps_masks, out: Path = GEN) -> list[Path]:
    """hs_<rep>.hip for every (representative, orientation mask)."""
    out.mkdir(parents=True, exist_ok=True)
    files = []
    for rep, mask in reps_masks:
        p = out / f"hs_{cls_name(rep)}.hip"
        text = emit_split_class(rep, mask)
        if not p.exists() or p.read_text() != text:
            p.write_text(text)
        files.append(p)
    return files


# ---------------------------------------------------------------------------
# The density-fitting classes (upstream eri/density_fit/kernel_metric.py,
# kernel_eri3.py): per-class entries with the upstream parameter lists onto
# the loop-form kernel (csrc/fj_dfloop.hip).

DF_AUX_LMAX = 6      # aux shells through i
DF_AO_LMAX = 3       # AO shells through f


def df_classes():
    """(metric classes (lP <= lQ), three-center classes (la <= lb, lP))."""
    met = [(p, q) for p in range(DF_AUX_LMAX + 1) for q in range(p, DF_AUX_LMAX + 1)]
    e3 = [(a, b, p) for a in range(DF_AO_LMAX + 1) for b in range(a, DF_AO_LMAX + 1)
          for p in range(DF_AUX_LMAX + 1)]
    return met, e3


METRIC_SIG = ("int threads, int n_pair, int n_rec, int n_aux, const double* A, const double* alpha,"
              " const double* s_aux, const int* p_rec0, const int* p_rec1, const int* q_rec0,"
              " const int* q_rec1, const int* off_p, const int* off_q, double* metric")
ERI3_SIG = ("int threads, int n_cell, int n_bra_rec, int n_aux_rec, int n_ao, int n_aux,"
            " const double* zeta_ab, const double* P, const double* XPA, const double* XPB,"
            " const double* s_ab, const double* A, const double* alpha, const double* s_aux,"
            " const int* ab_rec0, const int* ab_rec1, const int* p_rec0, const int* p_rec1,"
            " const int* off_a, const int* off_b, const int* off_p, double* eri3")


def emit_df_entries() -> str:
    met, e3 = df_classes()
    L = ["// generated by felizjueves_amd/gen_hip.py -- density-fitting classes on the loop-form kernel",
         '#include "../fj_device.h"', "",
         'extern "C" int fj_metric_loop(hipStream_t s, int lP, int lQ, ' + METRIC_SIG + ");",
         'extern "C" int fj_eri3_loop(hipStream_t s, int la, int lb, int lP, ' + ERI3_SIG + ");", ""]
    for p, q in met:
        L += [f'extern "C" int fj_metric_{p}{q}(hipStream_t s, {METRIC_SIG}) {{',
              f"    return fj_metric_loop(s, {p}, {q}, threads, n_pair, n_rec, n_aux, A, alpha, s_aux,"
              " p_rec0, p_rec1, q_rec0, q_rec1, off_p, off_q, metric);", "}", ""]
    for a, b, p in e3:
        L += [f'extern "C" int fj_eri3_{a}{b}{p}(hipStream_t s, {ERI3_SIG}) {{',
              f"    return fj_eri3_loop(s, {a}, {b}, {p}, threads, n_cell, n_bra_rec, n_aux_rec, n_ao,"
              " n_aux, zeta_ab, P, XPA, XPB, s_ab, A, alpha, s_aux, ab_rec0, ab_rec1, p_rec0, p_rec1,"
              " off_a, off_b, off_p, eri3);", "}", ""]
    return "\n".join(L)


def df_header_decls() -> str:
    met, e3 = df_classes()
    out = [f"int fj_metric_{p}{q}(fj_stream_t stream, {METRIC_SIG});" for p, q in met]
    out += [f"int fj_eri3_{a}{b}{p}(fj_stream_t stream, {ERI3_SIG});" for a, b, p in e3]
    return "\n".join(out) + "\n"


DF_BEGIN, DF_END = ("/* BEGIN DF CLASSES (written by felizjueves_amd/gen_hip.py) */",
                    "/* END DF CLASSES */")


def write_df(out: Path = GEN) -> Path:
    """gen/df_entries.hip and the header's DF section."""
    out.mkdir(parents=True, exist_ok=True)
    p = out / "df_entries.hip"
    text = emit_df_entries()
    if not p.exists() or p.read_text() != text:
        p.write_text(text)
    cur = HEADER.read_text()
    i, j = cur.index(DF_BEGIN), cur.index(DF_END)
    new = cur[:i + len(DF_BEGIN)] + "\n" + df_header_decls() + cur[j:]
    if new != cur:
        HEADER.write_text(new)
    return p


# ---------------------------------------------------------------------------
# The fused two-electron gradient: one entry per class of the gradient
# stream ((la lb | lc ld), la <= lb, lc <= ld, (la, lb) <= (lc, ld)) onto the
# loop-form kernel fj_gf_loop, the upstream gf_ parameter list.

GF_LMAX = 2

GF_SIG = ("int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_atom, int n_ao, int diag,"
          " const double* bra_prim, const double* AB, const int* bra_sh, const int* bra_ao,"
          " const int* bra_at, const double* ket_prim, const double* CD, const int* ket_sh,"
          " const int* ket_ao, const int* ket_at, const double* D, double* G")
GF_ARGS = ("threads, n_bra, n_ket, k_ab, k_cd, n_atom, n_ao, diag, bra_prim, AB, bra_sh, bra_ao,"
           " bra_at, ket_prim, CD, ket_sh, ket_ao, ket_at, D, G")


def gf_classes(lmax: int = GF_LMAX):
    prs = [(a, b) for a in range(lmax + 1) for b in range(a, lmax + 1)]
    return [(*p, *q) for i, p in enumerate(prs) for q in prs[i:]]


def emit_gf_entries() -> str:
    L = ["// generated by felizjueves_amd/gen_hip.py -- fused-gradient classes on the loop-form kernel",
         '#include "../fj_device.h"', "",
         'extern "C" int fj_gf_loop(hipStream_t s, int la, int lb, int lc, int ld, ' + GF_SIG + ");", ""]
    for c in gf_classes():
        name = "".join(map(str, c))
        L += [f'extern "C" int fj_gf_{name}(hipStream_t s, {GF_SIG}) {{',
              f"    return fj_gf_loop(s, {', '.join(map(str, c))}, {GF_ARGS});", "}", ""]
    return "\n".join(L)


def gf_header_decls() -> str:
    return "\n".join(f"int fj_gf_{''.join(map(str, c))}(fj_stream_t stream, {GF_SIG});"
                     for c in gf_classes()) + "\n"


GF_BEGIN, GF_END = ("/* BEGIN GRADIENT CLASSES (written by felizjueves_amd/gen_hip.py) */",
                    "/* END GRADIENT CLASSES */")


def write_gf(out: Path = GEN) -> Path:
    """gen/gf_entries.hip and the header's gradient section."""
    out.mkdir(parents=True, exist_ok=True)
    p = out / "gf_entries.hip"
    text = emit_gf_entries()
    if not p.exists() or p.read_text() != text:
        p.write_text(text)
    cur = HEADER.read_text()
    i, j = cur.index(GF_BEGIN), cur.index(GF_END)
    new = cur[:i + len(GF_BEGIN)] + "\n" + gf_header_decls() + cur[j:]
    if new != cur:
        HEADER.write_text(new)
    return p


SPLIT_BEGIN, SPLIT_END = ("/* BEGIN SPLIT CLASSES (written by felizjueves_amd/gen_hip.py) */",
                          "/* END SPLIT CLASSES */")


def write_split_header(reps) -> None:
    cur = HEADER.read_text()
    i, j = cur.index(SPLIT_BEGIN), cur.index(SPLIT_END)
    new = cur[:i + len(SPLIT_BEGIN)] + "\n" + split_header_decls(reps) + cur[j:]
    if new != cur:
        HEADER.write_text(new)


# ---------------------------------------------------------------------------
# Boys table: rows T_i = i * BOYS_H, i = 0..N; row = (e^-T_i, F_0(T_i) ..
# F_MTAB(T_i), 0 padding), every value the correctly rounded double of the
# Kummer series summed in 40-digit decimal arithmetic:
#   F_m(T) = e^-T sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)).

def boys_table() -> list[list[float]]:
    from decimal import Decimal, getcontext
    getcontext().prec = 40
    n_rows = int(round(BOYS_TMAX / BOYS_H)) + 1
    rows = []
    for i in range(n_rows):
        T = Decimal(i) / Decimal(8)
        eT = (-T).exp()
        row = [float(eT)]
        M = BOYS_MTAB
        term = Decimal(1) / Decimal(2 * M + 1)
        s = term
        k = 1
        while True:
            term = term * 2 * T / Decimal(2 * M + 2 * k + 1)
            s += term
            if term < s * Decimal("1e-38"):
                break
            k += 1
        f = eT * s
        vals = [f]
        for m in range(M, 0, -1):
            f = (2 * T * f + eT) / Decimal(2 * m - 1)
            vals.append(f)
        vals = vals[::-1]  # F_0 .. F_M
        row += [float(x) for x in vals]
        row += [0.0] * (BOYS_ROW - len(row))
        rows.append(row)
    return rows


def emit_boys_table() -> str:
    rows = boys_table()
    body = ",\n".join("    " + ", ".join(float.hex(x) for x in r) for r in rows)
    return (f"// generated by felizjueves_amd/gen_hip.py: Boys table, {len(rows)} rows "
            f"x {BOYS_ROW}\n#pragma once\n"
            f"#define FJ_BOYS_NROW {len(rows)}\n"
            f"__device__ const double fj_boys_tab[{len(rows)} * {BOYS_ROW}] = {{\n{body}\n}};\n")


def generate(classes, out: Path = GEN) -> list[Path]:
    out.mkdir(parents=True, exist_ok=True)
    files = []
    tab = out / "fj_boys_table.h"
    text = emit_boys_table()
    if not tab.exists() or tab.read_text() != text:
        tab.write_text(text)
    for cls in classes:
        n = cls_name(cls)
        head, main = emit_class(cls).split("//@@SPLIT@@\n")
        head, mixed = head.split("//@@MIXED@@\n")
        mx = (f"// generated by felizjueves_amd/gen_hip.py -- mixed-tier kernel of ({n})\n"
              f'#include "f_{n}.h"\n\n' + mixed +
              f"\nint fj_launch_{n}_mx(hipStream_t s, dim3 grid, dim3 block, const FjFArgs& a) {{\n"
              f"    hipLaunchKernelGGL(fj_f_{n}_km, grid, block, 0, s, a);\n"
              f"    return (int)hipGetLastError();\n}}\n")
        r32 = (f"// generated by felizjueves_amd/gen_hip.py -- FP32-tier kernel of ({n})\n"
               f'#include "f_{n}.h"\n\n'
               f"int fj_launch_{n}_r32(hipStream_t s, dim3 grid, dim3 block, const FjFArgs& a) {{\n"
               f"    hipLaunchKernelGGL(fj_f_{n}_k<float>, grid, block, 0, s, a);\n"
               f"    return (int)hipGetLastError();\n}}\n")
        texts = {f"f_{n}.h": head,
                 f"f_{n}.hip": f'// generated by felizjueves_amd/gen_hip.py\n#include "f_{n}.h"\n' + main,
                 f"f_{n}_r32.hip": r32,
                 f"f_{n}_mx.hip": mx}
        for fname, text in texts.items():
            p = out / fname
            if not p.exists() or p.read_text() != text:
                p.write_text(text)
        files += [out / f"f_{n}.hip", out / f"f_{n}_r32.hip", out / f"f_{n}_mx.hip"]
    wt = out / "fj_ws_table.hip"
    text = emit_ws_table(classes)
    if not wt.exists() or wt.read_text() != text:
        wt.write_text(text)
    files.append(wt)
    return files


def emit_ws_table(classes) -> str:
    """fj_fx_workspace_bytes: the workspace each built class asks for."""
    lines = ["// generated by felizjueves_amd/gen_hip.py -- launch workspace per class",
             '#include "../fj_device.h"', "",
             'extern "C" long long fj_fx_workspace_bytes(int la, int lb, int lc, int ld, int threads,',
             "                                           int n_work) {",
             "    if (threads <= 0) threads = FJ_DEFAULT_THREADS;",
             "    if (n_work <= 0) return 0;",
             "    long long per_lane = 0;",
             "    switch (((la * 4 + lb) * 4 + lc) * 4 + ld) {"]
    for cls in classes:
        w = ws_doubles(cls)
        if w:
            la, lb, lc, ld = cls
            lines.append(f"        case {((la * 4 + lb) * 4 + lc) * 4 + ld}: per_lane = {w}; break;")
    lines += ["        default: return 0;", "    }",
              "    const long long blocks = n_work < FJ_WS_MAX_BLOCKS ? n_work : FJ_WS_MAX_BLOCKS;",
              "    return 8 * per_lane * threads * blocks;", "}", ""]
    return "\n".join(lines)


def header_decls(classes) -> str:
    """The F-class declarations of include/fj_mi355x.h."""
    out = []
    for cls in classes:
        la, lb, lc, ld = cls
        n = cls_name(cls)
        up = (
            f"int fj_f_{n}(fj_stream_t stream, int threads, int n_bra, int n_ket, int k_ab, int k_cd,\n"
            f"    int n_sh, int n_ao, int diag, double tau,\n"
            f"    const double* bra_prim,{' const double* AB,' if lb else ''} const double* G_ab,"
            f" const int* bra_sh, const int* bra_ao,\n"
            f"    const double* ket_prim,{' const double* CD,' if ld else ''} const double* G_cd,"
            f" const int* ket_sh, const int* ket_ao,\n"
            f"    const double* Dsh, const double* D, double* F);\n") if canonical(cls) else ""
        out.append(
            up +
            f"int fj_fx_{n}(fj_stream_t stream, int threads, int precision, const FjSide* bra,"
            f" const FjSide* ket,\n"
            f"    const int* work, int n_work, const int* ket_list, const FjSys* sys, int diag,\n"
            f"    double tau, double* ws, long long ws_bytes);\n"
            f"int fj_fxm_{n}(fj_stream_t stream, int threads, const FjSide* bra, const FjSide* ket,\n"
            f"    const int* work, int n_work, int n_fp64, const int* tier, const int* ket_list,\n"
            f"    const FjSys* sys, int diag, double tau, double* ws, long long ws_bytes);")
    return "\n".join(out) + "\n"


def emit_loop_entries(classes) -> str:
    """fj_f_ (canonical classes), fj_fx_ and fj_fxm_ of the classes run by the
    loop-form fused kernel (csrc/fj_floop.h): the generated kernels' entry
    checks and argument records, one FP64 kernel for every tier."""
    L = ["// generated by felizjueves_amd/gen_hip.py -- F classes on the loop-form kernel",
         '#include "../fj_device.h"', "",
         "int fj_floop_launch(hipStream_t s, int la, int lb, int lc, int ld, int threads, FjFArgs a,",
         "                    int n_blocks);", ""]
    for cls in classes:
        la, lb, lc, ld = cls
        n = cls_name(cls)
        call = f"fj_floop_launch(s, {la}, {lb}, {lc}, {ld}, threads, a"
        if canonical(cls):
            L += [
                "// the upstream-ABI launch with every argument; the loop-form kernel",
                "// uses no workspace, so ws is ignored",
                f"int fj_f_any_{n}(hipStream_t s, int threads, int n_bra, int n_ket, int k_ab, int k_cd,",
                "                    int n_sh, int n_ao, int diag, double tau,",
                "                    const double* bra_prim, const double* AB, const double* G_ab,",
                "                    const int* bra_sh, const int* bra_ao, const double* ket_prim,",
                "                    const double* CD, const double* G_cd, const int* ket_sh,",
                "                    const int* ket_ao, const double* Dsh, const double* D, double* F,",
                "                    double* ws, long long ws_bytes) {",
                "    (void)ws; (void)ws_bytes;",
                "    if (k_ab < 1 || k_ab > FJ_MAX_PRIM_PAIRS || k_cd < 1 || k_cd > FJ_MAX_PRIM_PAIRS) return FJ_ERR_DEGREE;",
                "    if (n_bra < 0 || n_ket < 0 || (diag && n_bra != n_ket)) return FJ_ERR_SHAPE;",
                "    FjFArgs a;",
                f"    a.bra = FjSide{{bra_prim, nullptr, k_ab, {'AB' if lb else 'nullptr'}, G_ab, bra_sh, bra_ao, n_bra}};",
                f"    a.ket = FjSide{{ket_prim, nullptr, k_cd, {'CD' if ld else 'nullptr'}, G_cd, ket_sh, ket_ao, n_ket}};",
                "    a.work = nullptr; a.ket_list = nullptr; a.sys = nullptr;",
                "    a.one = FjSys{D, Dsh, F, n_ao, n_sh};",
                "    a.diag = diag; a.tau = tau;",
                "    a.ws = nullptr; a.ws_doubles = 0; a.n_fp64 = 0; a.tier = nullptr;",
                f"    return {call}, n_bra);", "}", "",
                f'extern "C" int fj_f_{n}(hipStream_t s, int threads, int n_bra, int n_ket, int k_ab, int k_cd,',
                "                        int n_sh, int n_ao, int diag, double tau,",
                f"                        const double* bra_prim,{' const double* AB,' if lb else ''}"
                " const double* G_ab, const int* bra_sh, const int* bra_ao,",
                f"                        const double* ket_prim,{' const double* CD,' if ld else ''}"
                " const double* G_cd, const int* ket_sh, const int* ket_ao,",
                "                        const double* Dsh, const double* D, double* F) {",
                f"    return fj_f_any_{n}(s, threads, n_bra, n_ket, k_ab, k_cd, n_sh, n_ao, diag, tau, bra_prim,",
                f"                        {'AB' if lb else 'nullptr'}, G_ab, bra_sh, bra_ao, ket_prim,"
                f" {'CD' if ld else 'nullptr'}, G_cd, ket_sh, ket_ao,",
                "                        Dsh, D, F, nullptr, 0);", "}", ""]
        L += [
            f'extern "C" int fj_fx_{n}(hipStream_t s, int threads, int precision,',
            "                         const FjSide* bra, const FjSide* ket, const int* work, int n_work,",
            "                         const int* ket_list, const FjSys* sys, int diag, double tau,",
            "                         double* ws, long long ws_bytes) {",
            "    if (!bra || !ket || !work || !sys || ws_bytes < 0) return FJ_ERR_ARG;",
            "    if (precision != 0 && precision != 1) return FJ_ERR_ARG;",
            "    FjFArgs a;",
            "    a.bra = *bra; a.ket = *ket;",
            "    a.work = reinterpret_cast<const int4*>(work);",
            "    a.ket_list = ket_list; a.sys = sys;",
            "    a.one = FjSys{nullptr, nullptr, nullptr, 0, 0};",
            "    a.diag = diag; a.tau = tau;",
            "    a.ws = nullptr; a.ws_doubles = 0; a.n_fp64 = 0; a.tier = nullptr;",
            "    // the loop-form kernel computes in FP64 at either precision",
            f"    return {call}, n_work);", "}", "",
            f'extern "C" int fj_fxm_{n}(hipStream_t s, int threads, const FjSide* bra, const FjSide* ket,',
            "                          const int* work, int n_work, int n_fp64, const int* tier,",
            "                          const int* ket_list, const FjSys* sys, int diag, double tau,",
            "                          double* ws, long long ws_bytes) {",
            "    if (!bra || !ket || !work || !sys || ws_bytes < 0) return FJ_ERR_ARG;",
            "    if (n_fp64 < 0 || n_fp64 > n_work) return FJ_ERR_ARG;",
            "    FjFArgs a;",
            "    a.bra = *bra; a.ket = *ket;",
            "    a.work = reinterpret_cast<const int4*>(work);",
            "    a.ket_list = ket_list; a.sys = sys;",
            "    a.one = FjSys{nullptr, nullptr, nullptr, 0, 0};",
            "    a.diag = diag; a.tau = tau;",
            "    a.ws = nullptr; a.ws_doubles = 0; a.n_fp64 = n_fp64; a.tier = tier;",
            f"    return {call}, n_work);", "}", ""]
    return "\n".join(L)


def emit_any_table(classes) -> str:
    """gen/f_any_table.h: the declarations of fj_f_any_<cls> of the built
    canonical classes and fj_f_any_for(la, lb, lc, ld), their dispatch."""
    sig = ("(hipStream_t s, int threads, int n_bra, int n_ket, int k_ab, int k_cd, int n_sh, "
           "int n_ao, int diag, double tau, const double* bra_prim, const double* AB, "
           "const double* G_ab, const int* bra_sh, const int* bra_ao, const double* ket_prim, "
           "const double* CD, const double* G_cd, const int* ket_sh, const int* ket_ao, "
           "const double* Dsh, const double* D, double* F, double* ws, long long ws_bytes)")
    can = [c for c in classes if canonical(c)]
    L = ["// generated by felizjueves_amd/gen_hip.py -- the upstream-ABI launches by class",
         "#pragma once", '#include "../fj_device.h"', "",
         f"typedef int (*fj_f_any_t){sig};"]
    L += [f"int fj_f_any_{cls_name(c)}{sig};" for c in can]
    L += ["", "static inline fj_f_any_t fj_f_any_for(int la, int lb, int lc, int ld) {",
          "    if (la < 0 || lb < 0 || lc < 0 || ld < 0 || la > 3 || lb > 3 || lc > 3 || ld > 3)",
          "        return nullptr;",
          "    switch (((la * 4 + lb) * 4 + lc) * 4 + ld) {"]
    for c in can:
        la, lb, lc, ld = c
        L.append(f"        case {((la * 4 + lb) * 4 + lc) * 4 + ld}: return fj_f_any_{cls_name(c)};")
    L += ["        default: return nullptr;", "    }", "}", ""]
    return "\n".join(L)


HEADER = HERE.parent / "include" / "fj_mi355x.h"
BEGIN, END = ("/* BEGIN F CLASSES (written by felizjueves_amd/gen_hip.py) */",
              "/* END F CLASSES */")


def header_text(classes, current: str) -> str:
    i, j = current.index(BEGIN), current.index(END)
    return current[:i + len(BEGIN)] + "\n" + header_decls(classes) + current[j:]


def write_header(classes) -> None:
    cur = HEADER.read_text()
    new = header_text(classes, cur)
    if new != cur:
        HEADER.write_text(new)


# ---------------------------------------------------------------------------
# The staged body, for classes whose fused body would not fit the register
# file (FUSED_LIVE_MAX): phase A runs the primitive nest once per group of
# bra components e (hgp.e_groups), each group holding only its own
# accumulators and the closure of its targets, and stores the contracted
# boundary (e0|f0) to a private array; phase B is the bra HRR, one loop trip
# per ket component f; phase C is the ket HRR per bra pair component ab,
# with that row's digestion (F_ab and F_cd in registers, the four exchange
# blocks flushed per ab).

# classes whose whole VRR keeps at most this many values live are emitted
# fused (all in registers); measured on MI355X (40-atom build): 120 -> 91.5
# ms, 200 -> 88.5 ms (e.g. (sd|sd) 6.2 -> 4.1 ms), 400 -> 116 ms (spills)
FUSED_LIVE_MAX = 200

# A staged class keeps its contracted boundary, half-transformed block and
# F_cd accumulators in a launch workspace in HBM instead of private scratch:
# the runtime sizes a queue's scratch for every wavefront the chip can hold,
# and ~17 KB per lane on four queues at once (or ~100 KB on one) is more
# than it will allocate. Its kernel is persistent over the work items, one
# workspace slot per resident lane, the slots interleaved so that a
# wavefront's accesses coalesce, and each of its (e, f) VRR groups is a
# separate non-inlined function so registers are allocated per group.
EF_LIVE_MAX = 120   # live values per (e, f) VRR group of an f class
# live values per bra-component group of a d class: measured on MI355X
# (40-atom build, eager): 100 -> 98.2 ms, 220 -> 92.0 ms, 300 -> 121 ms
# (pairs of components halve the Boys evaluations and cut the recompute
# from 1.75x to 1.62x at (dd|dd); larger groups spill)
E_LIVE_MAX = 220


def ws_doubles(cls) -> int:
    """Workspace doubles per lane of class cls (0: a fused class, all in
    registers)."""
    if not staged(cls):
        return 0
    la, lb, lc, ld = cls
    ne = sum(n_cart(n) for n in range(la, la + lb + 1))
    nf = sum(n_cart(n) for n in range(lc, lc + ld + 1))
    w = ne * nf + n_cart(la) * n_cart(lb) * nf
    return w + n_cart(lc) * n_cart(ld)


def staged(cls) -> bool:
    from .hgp import peak_live
    v = vrr_dag(tuple(cls))
    return peak_live(v, set(v.order)) + len(v.targets) > FUSED_LIVE_MAX


def _hrr_side(la: int, lb: int):
    """One side's HRR as a DAG over (a, b) with the other side fixed:
    (a, b) = (a + 1_k, b - 1_k) + sep_k (a, b - 1_k), k the first nonzero
    axis of b; leaves (e, 0), la <= |e| <= la + lb. Returns (order, terms,
    targets, leaves) with targets in cartesians(la) x cartesians(lb) order."""
    from .hgp import _add, _sub
    targets = [(a, b) for a in cartesians(la) for b in cartesians(lb)]
    terms: dict = {}
    todo = list(targets)
    while todo:
        n = todo.pop()
        if n in terms:
            continue
        a, b = n
        if b != ZERO:
            k = next(i for i, x in enumerate(b) if x)
            r = (k, (_add(a, k), _sub(b, k)), (a, _sub(b, k)))
            terms[n] = r
            todo += [r[1], r[2]]
        else:
            terms[n] = ()
    order = sorted(terms, key=lambda n: (sum(n[1]), -sum(n[0]), tuple(-x for x in n[0]),
                                         tuple(-x for x in n[1])))
    leaves = [e for n in range(la, la + lb + 1) for e in cartesians(n)]
    return order, terms, targets, leaves


def _emit_side_hrr_dfs(E: _Emit, la: int, lb: int, sep: str, src, dst, ind: int) -> None:
    """One side's HRR emitted depth-first per target, each target handed to
    dst(j, expr) right after its statement (targets feed no other node), so
    only the memoised intermediates that later targets still need stay
    live."""
    order, terms, targets, leaves = _hrr_side(la, lb)
    li = {e: i for i, e in enumerate(leaves)}
    name: dict = {}

    def need(n):
        if n in name:
            return name[n]
        r = terms[n]
        nm = f"t{_fidx(n[0])}_{_fidx(n[1])}"
        if not r:
            E(f"const double {nm} = {src(li[n[0]])};", ind)
        else:
            k, moved, lowered = r
            x, y = need(moved), need(lowered)
            E(f"const double {nm} = {x} + {sep}{AX[k]} * {y};", ind)
        name[n] = nm
        return nm
    for j, t in enumerate(targets):
        dst(j, need(t))


def _emit_side_hrr(E: _Emit, la: int, lb: int, sep: str, src, dst, ind: int) -> None:
    """Emit one side's HRR: leaves read by src(i) (i over the e range),
    targets written by dst(j, expr) (j over cartesians(la) x cartesians(lb))."""
    order, terms, targets, leaves = _hrr_side(la, lb)
    li = {e: i for i, e in enumerate(leaves)}
    name = {}
    for n in order:
        nm = f"t{_fidx(n[0])}_{_fidx(n[1])}"
        name[n] = nm
        r = terms[n]
        if not r:
            E(f"const double {nm} = {src(li[n[0]])};", ind)
        else:
            k, moved, lowered = r
            E(f"const double {nm} = {name[moved]} + {sep}{AX[k]} * {name[lowered]};", ind)
    for j, t in enumerate(targets):
        dst(j, name[t])


def emit_quartet_staged(cls, split: bool = False) -> str:
    from .hgp import closure, e_groups
    la, lb, lc, ld = cls
    L = sum(cls)
    v = vrr_dag(tuple(cls))
    name = cls_name(cls)
    na, nb, nc, nd = (n_cart(x) for x in cls)
    eR = [e for n in range(la, la + lb + 1) for e in cartesians(n)]
    fR = [f for n in range(lc, lc + ld + 1) for f in cartesians(n)]
    NE, NF = len(eR), len(fR)
    NAB, NCD = na * nb, nc * nd
    ei = {e: i for i, e in enumerate(eR)}
    fi = {f: i for i, f in enumerate(fR)}
    groups = e_groups(tuple(cls), E_LIVE_MAX)

    E = _Emit()
    W = ws_doubles(cls)
    # f classes: (e, f) groups, each its own non-inlined function (registers
    # allocated per group); d classes: groups of bra components inline, with
    # the primitive records read straight from LDS
    fns = max(cls) >= 3
    if W:
        from .hgp import ef_groups
        tgroups = ([list(g) for g in ef_groups(tuple(cls), EF_LIVE_MAX)] if fns else
                   [[(e, f) for e, f in v.targets if e in g] for g in groups])
        bnd_ref = lambda i: f"ws[(size_t)({i}) * st]"  # noqa: E731
        mid_ref = lambda i: f"ws[(size_t)({NF * NE} + {i}) * st]"  # noqa: E731
    else:
        tgroups = [[(e, f) for e, f in v.targets if e in g] for g in groups]
        bnd_ref = lambda i: f"bnd[{i}]"  # noqa: E731
        mid_ref = lambda i: f"mid[{i}]"  # noqa: E731
    E(f"// ({la}{lb}|{lc}{ld}), staged: {len(tgroups)} VRR groups, "
      f"{len(v.targets)} boundary values", 0)

    def group_body(tg, ind):
        """The primitive nest of one group: its accumulators, the closure of
        its targets, the contracted values stored to the boundary."""
        nodes = closure(v, [(e, f, 0) for e, f in tg])
        used = set()
        for n in nodes:
            r = v.terms[n]
            if r:
                d, ts = r
                for t in ts:
                    used.add((t.coef, d if t.coef in ("PA", "WP", "QC", "WQ") else None))
        need = lambda c: any(u[0] == c for u in used)  # noqa: E731
        for i in range(len(tg)):
            E(f"R c{i} = R(0);", ind)
        E("for (int kc = 0; kc < k_cd; ++kc) {", ind)
        E("const double* __restrict__ q_ = kp + 8 * kc;", ind + 1)
        E("const double Qx = q_[0], Qy = q_[1], Qz = q_[2], q = q_[6], Uq = q_[7];", ind + 1)
        for d in range(3):
            if ("QC", d) in used:
                E(f"const R QC{AX[d]} = R(q_[{3 + d}]);", ind + 1)
        if need("i2q") or need("i2q_r"):
            E("const double iq = 1.0 / q;", ind + 1)
        E("for (int kb = 0; kb < k_ab; ++kb) {", ind + 1)
        i2 = ind + 2
        E("const double* __restrict__ p_ = bp + FJ_SREC * kb;", i2)
        E("const double PQx = p_[0] - Qx, PQy = p_[1] - Qy, PQz = p_[2] - Qz;", i2)
        E("const double p = p_[6];", i2)
        E("const double ippq = 1.0 / (p + q);", i2)
        E("const double rho = p * q * ippq;", i2)
        E("const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);", i2)
        E(f"double Fm[{L + 1}];", i2)
        E(f"fj_boys_lds<{L}>(T, Fm, tb);" if W else f"fj_boys<{L}>(T, Fm);", i2)
        E("const double S = p_[7] * Uq * (FJ_TWO_OVER_SQRT_PI * sqrt(rho));", i2)
        for d in range(3):
            if ("PA", d) in used:
                E(f"const R PA{AX[d]} = R(p_[{3 + d}]);", i2)
            if ("WP", d) in used:
                E(f"const R WP{AX[d]} = R(-q * ippq * PQ{AX[d]});", i2)
            if ("WQ", d) in used:
                E(f"const R WQ{AX[d]} = R(p * ippq * PQ{AX[d]});", i2)
        if need("i2p"):
            E("const R i2p = R(0.5 * p_[8]);", i2)
        if need("i2p_r"):
            E("const R i2p_r = R(0.5 * p_[8] * rho * p_[8]);", i2)
        if need("i2q"):
            E("const R i2q = R(0.5 * iq);", i2)
        if need("i2q_r"):
            E("const R i2q_r = R(0.5 * iq * rho * iq);", i2)
        if need("i2pq"):
            E("const R i2pq = R(0.5 * ippq);", i2)
        for n in v.order:
            if n not in nodes:
                continue
            r = v.terms[n]
            if not r:
                E(f"const R {_vname(n)} = R(S * Fm[{n[2]}]);", i2)
                continue
            d, ts = r
            parts = []
            for t in ts:
                c = t.coef
                ch = _vname(t.child)
                if c in ("PA", "WP", "QC", "WQ"):
                    parts.append(f"{c}{AX[d]} * {ch}")
                else:
                    m = t.mult
                    sign = "-" if m < 0 else "+"
                    f = c if abs(m) == 1 else f"(R({abs(m)}) * {c})"
                    parts.append(f"{sign} {f} * {ch}")
            expr = parts[0]
            for q_ in parts[1:]:
                expr += (" " + q_) if q_[0] in "+-" else (" + " + q_)
            E(f"const R {_vname(n)} = {expr};", i2)
        for i, (e, f) in enumerate(tg):
            E(f"c{i} += {_vname((e, f, 0))};", i2)
        E("}", ind + 1)
        E("}", ind)
        for i, (e, f) in enumerate(tg):
            E(f"{bnd_ref(fi[f] * NE + ei[e])} = double(c{i});", ind)

    if W and fns:
        for gi, tg in enumerate(tgroups):
            E("template <typename R>", 0)
            E(f"__device__ __noinline__ void fj_g{gi}_{name}(const double* __restrict__ bp, int k_ab,", 0)
            E("const double* __restrict__ kp, int k_cd, double* __restrict__ ws, size_t st,", 1)
            E("const double* __restrict__ tb) {", 1)
            group_body(tg, 1)
            E("}", 0)
    E("template <typename R>", 0)
    E(f"__device__ __forceinline__ void fj_q{'s' if split else ''}_{name}(", 0)
    E("const double* __restrict__ bp, int k_ab,")
    E("const double* __restrict__ kp, int k_cd,")
    E("double ABx, double ABy, double ABz, double CDx, double CDy, double CDz,")
    E("const double* __restrict__ D, int n_ao, int oc, int od, int oa, int ob,")
    E("const FjDig& dg, const double* __restrict__ wab," if split else
      "double s_j, double s_k, const double* __restrict__ wab,")
    E("double* __restrict__ fab, double* __restrict__ F, double* __restrict__ ws, size_t st,", 1)
    E("const double* __restrict__ tb) {", 1)
    if split:
        E(SPLIT_WEIGHTS, 0)
    if W:
        E(f"// (e0|f0) [f][e], (ab|f0) [ab][f] and F_cd in the workspace ({W} doubles per lane)")
    if W and fns:
        for gi in range(len(tgroups)):
            E(f"fj_g{gi}_{name}<R>(bp, k_ab, kp, k_cd, ws, st, tb);")
    else:
        if not W:
            E(f"double bnd[{NF * NE}];   // contracted (e0|f0), [f][e]")
        for gi, tg in enumerate(tgroups):
            E(f"{{   // group {gi}: e = {sorted(set(e for e, _ in tg))}")
            group_body(tg, 2)
            E("}")
    # phase B: bra HRR per f
    if not W:
        E(f"double mid[{NAB * NF}];   // (ab|f0), [ab][f]")
    if lb == 0:
        E("#pragma unroll 1", 0)
        E(f"for (int i = 0; i < {NAB * NF}; ++i) {mid_ref(f'(i % {NE}) * {NF} + i / {NE}')} = {bnd_ref('i')};")
    else:
        E("#pragma unroll 1", 0)
        E(f"for (int f = 0; f < {NF}; ++f) {{")
        hrr = _emit_side_hrr_dfs if fns else _emit_side_hrr
        hrr(E, la, lb, "AB", lambda i: bnd_ref(f"f * {NE} + {i}"),
            lambda j, nm: E(f"{mid_ref(f'{j} * {NF} + f')} = {nm};", 2), 2)
        E("}")
    # phase C: ket HRR per ab + digestion, each ERI digested as it is made.
    # F_cd accumulates over ab: in registers, or (workspace classes) in the
    # lane's workspace slice with its density weights read from D.
    NCD = nc * nd
    fcd_ws = bool(W) and fns   # f classes: F_cd too large for registers
    fcd_ref = lambda j: f"ws[(size_t)({NF * NE + NAB * NF} + {j}) * st]"  # noqa: E731
    if fcd_ws:
        E("#pragma unroll 1", 0)
        E(f"for (int j = 0; j < {NCD}; ++j) {fcd_ref('j')} = 0.0;")
    else:
        for c in range(nc):
            for d in range(nd):
                E(f"const double wcd{c}_{d} = {_dw(split, 'cd')} * "
                  f"D[{_didx(split, f'oc + {c}', f'od + {d}')}];")
                E(f"double fcd{c}_{d} = 0.0;")
    E("#pragma unroll 1", 0)
    E(f"for (int ab = 0; ab < {NAB}; ++ab) {{")
    E(f"const int a = ab / {nb}, b = ab % {nb};", 2)
    E(f"const double w_ab = {_dw(split, 'ab')} * D[{_didx(split, 'oa + a', 'ob + b')}];", 2)
    if split:
        for c in range(nc):
            E(f"const double dbc{c} = s_ad * D[{_didx(True, 'ob + b', f'oc + {c}')}], "
              f"dac{c} = s_bd * D[{_didx(True, 'oa + a', f'oc + {c}')}];", 2)
        for d in range(nd):
            E(f"const double dbd{d} = s_ac * D[{_didx(True, 'ob + b', f'od + {d}')}], "
              f"dad{d} = s_bc * D[{_didx(True, 'oa + a', f'od + {d}')}];", 2)
    else:
        E("const double* __restrict__ Db = D + (size_t)(ob + b) * n_ao;", 2)
        E("const double* __restrict__ Da = D + (size_t)(oa + a) * n_ao;", 2)
        for c in range(nc):
            E(f"const double dbc{c} = s_k * Db[oc + {c}], dac{c} = s_k * Da[oc + {c}];", 2)
        for d in range(nd):
            E(f"const double dbd{d} = s_k * Db[od + {d}], dad{d} = s_k * Da[od + {d}];", 2)
    E("double sab = 0.0;", 2)
    for c in range(nc):
        E(f"double fac{c} = 0.0, fbc{c} = 0.0;", 2)
    for d in range(nd):
        E(f"double fad{d} = 0.0, fbd{d} = 0.0;", 2)

    def digest(j, x):
        c, d = divmod(j, nd)
        if fcd_ws:
            E(f"sab += ({_dw(split, 'cd')} * D[{_didx(split, f'oc + {c}', f'od + {d}')}]) * {x}; "
              f"{fcd_ref(j)} += w_ab * {x};", 2)
        else:
            E(f"sab += wcd{c}_{d} * {x}; fcd{c}_{d} += w_ab * {x};", 2)
        E(f"fac{c} += dbd{d} * {x}; fad{d} += dbc{c} * {x}; "
          f"fbc{c} += dad{d} * {x}; fbd{d} += dac{c} * {x};", 2)
    (_emit_side_hrr_dfs if fns else _emit_side_hrr)(
        E, lc, ld, "CD", lambda i: mid_ref(f"ab * {NF} + {i}"), digest, 2)
    E("fab[ab] += sab;", 2)
    for c in range(nc):
        if split:
            E(f"if (s_ac != 0.0) unsafeAtomicAdd(F + {_fidx_out(True, 'oa + a', f'oc + {c}')}, fac{c});", 2)
            E(f"if (s_bc != 0.0) unsafeAtomicAdd(F + {_fidx_out(True, 'ob + b', f'oc + {c}')}, fbc{c});", 2)
        else:
            E(f"unsafeAtomicAdd(F + (size_t)(oa + a) * n_ao + oc + {c}, fac{c});", 2)
            E(f"unsafeAtomicAdd(F + (size_t)(ob + b) * n_ao + oc + {c}, fbc{c});", 2)
    for d in range(nd):
        if split:
            E(f"if (s_ad != 0.0) unsafeAtomicAdd(F + {_fidx_out(True, 'oa + a', f'od + {d}')}, fad{d});", 2)
            E(f"if (s_bd != 0.0) unsafeAtomicAdd(F + {_fidx_out(True, 'ob + b', f'od + {d}')}, fbd{d});", 2)
        else:
            E(f"unsafeAtomicAdd(F + (size_t)(oa + a) * n_ao + od + {d}, fad{d});", 2)
            E(f"unsafeAtomicAdd(F + (size_t)(ob + b) * n_ao + od + {d}, fbd{d});", 2)
    E("}")
    if fcd_ws:
        E("#pragma unroll 1", 0)
        E(f"for (int j = 0; j < {NCD}; ++j)")
        if split:
            E(f"if (s_cd != 0.0) unsafeAtomicAdd(F + {_fidx_out(True, f'oc + j / {nd}', f'od + j % {nd}')}, "
              f"{fcd_ref('j')});", 2)
        else:
            E(f"unsafeAtomicAdd(F + (size_t)(oc + j / {nd}) * n_ao + od + j % {nd}, {fcd_ref('j')});", 2)
    else:
        if split:
            E("if (s_cd != 0.0) {")
        for c in range(nc):
            for d in range(nd):
                E(f"unsafeAtomicAdd(F + {_fidx_out(True, f'oc + {c}', f'od + {d}')}, fcd{c}_{d});"
                  if split else f"unsafeAtomicAdd(F + (size_t)(oc + {c}) * n_ao + od + {d}, fcd{c}_{d});",
                  2 if split else 1)
        if split:
            E("}")
    E("}", 0)
    return E.text()
