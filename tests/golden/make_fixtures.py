"""Write the libcint golden vectors this repository's tests read, as .npz.

Run once, in a container that has the upstream FelizJueves checkout at
/root/reference (it is not present where the tests run):

    python tests/golden/make_fixtures.py

Source: the upstream tracked fixtures `test/oracle/fixtures/*.json`, produced
there by PySCF 2.14.0 / libcint 6.1.3 (each file's `provenance`). They are
DATA -- the molecule, the basis already converted to the unnormalised-
Cartesian-primitive convention (`converted.shells`), the recorded density and
libcint's arrays -- read here with `json` only; nothing of the upstream code
is imported or executed.

Each output `<name>.npz` holds:

    shell_l [n_sh] i4, shell_centre [n_sh,3], shell_nprim [n_sh] i4,
    exps [sum nprim], coefs [sum nprim]        the converted basis, AO order
    density [nao,nao]                          the fixture's D (Hilbert matrix)
    vj, vk [nao,nao]          (if present)     PySCF get_jk J[D], K[D]
    int2e_diag [nao,nao]      (if present)     libcint (mn|mn)
    eri_keys [nblk,4] i4, eri_off [nblk+1] i8, eri_vals   int2e blocks, one
                                               shell quartet each, row-major
                                               (n_i, n_j, n_k, n_l)

and each density-fitting set `df_<name>.npz` (DF_SETS) holds the auxiliary
basis as aux_shell_l, aux_shell_centre, aux_shell_nprim, aux_exps,
aux_coefs, with int2c2e and int1e_ovlp [n_aux, n_aux] (metric sets) or the
AO basis and int3c2e [n_ao, n_ao, n_aux] (three-center sets).

Each gradient set `grad_<name>.npz` (GRAD_SETS) holds the converted basis,
shell_atom [n_sh] i4 (the atom each shell sits on), atom_coord [n_atom,3],
the converged density PySCF differentiated at, and its two-electron
gradient terms grad_j, grad_k [n_atom,3] (PySCF's grad:j, grad:k), whose
sum is dE_2e/dR for E_2e = 1/2 tr(D J) - 1/4 tr(D K).
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/test/oracle/fixtures")
OUT = Path(__file__).resolve().parent

# output name -> source fixture files merged into it
SETS = {
    "h2_sto3g": ["h2_sto3g_j", "h2_sto3g_k", "h2_sto3g_schwarz"],
    "h2o_sto3g": ["h2o_sto3g_j", "h2o_sto3g_k", "h2o_sto3g_schwarz"],
    "h2o_631g": ["h2o_631g_j", "h2o_631g_k", "h2o_631g_schwarz", "h2o_631g_eri"],
    "h2o_ccpvdz": ["h2o_ccpvdz_j", "h2o_ccpvdz_k"],
    "h10_chain_ccpvdz": ["h10_chain_ccpvdz_j", "h10_chain_ccpvdz_k"],
    "co_ccpvtz": ["co_ccpvtz_j", "co_ccpvtz_k"],
    "hf_ccpvtz": ["hf_ccpvtz_eri", "hf_ccpvtz_eri_l9", "hf_ccpvtz_eri_l10",
                  "hf_ccpvtz_eri_l11", "hf_ccpvtz_eri_l12"],
}


def _arr(a: dict) -> np.ndarray:
    return np.asarray(a["values"], dtype=np.float64).reshape(a["shape"])


def _basis(conv: dict) -> dict[str, np.ndarray]:
    sh = conv["shells"]
    return {
        "shell_l": np.array([s["l"] for s in sh], dtype=np.int32),
        "shell_centre": np.array([s["centre"] for s in sh], dtype=np.float64),
        "shell_nprim": np.array([len(s["exponents"]) for s in sh], dtype=np.int32),
        "exps": np.array([x for s in sh for x in s["exponents"]], dtype=np.float64),
        "coefs": np.array([x for s in sh for x in s["coefficients"]], dtype=np.float64),
    }


def convert(name: str, sources: list[str]) -> dict[str, np.ndarray]:
    out: dict[str, np.ndarray] = {}
    blocks: dict[tuple[int, ...], np.ndarray] = {}
    for src in sources:
        fx = json.loads((SRC / f"{src}.json").read_text())
        basis = _basis(fx["converted"])
        for k, v in basis.items():
            if k in out and not np.array_equal(out[k], v):
                raise ValueError(f"{src}: basis differs from the set's")
            out[k] = v
        arrays = fx["arrays"]
        for key in ("vj", "vk", "int2e_diag"):
            if key in arrays:
                out[key] = _arr(arrays[key])
        if "vj" in arrays or "vk" in arrays or "density" not in out:
            out["density"] = _arr(arrays["density"])
        for k, a in arrays.items():
            if k.startswith("int2e:"):
                q = tuple(int(x) for x in k[len("int2e:"):].split(","))
                blocks[q] = _arr(a)
    if blocks:
        keys = sorted(blocks)
        out["eri_keys"] = np.array(keys, dtype=np.int32)
        off = np.cumsum([0] + [blocks[q].size for q in keys]).astype(np.int64)
        out["eri_off"] = off
        out["eri_vals"] = np.concatenate([blocks[q].ravel() for q in keys])
    return out


# density-fitting sets: output name -> (kind, source chunk files, the AO
# basis's source fixture or None). Each chunk holds a block of rows (metric)
# or aux columns (three-center) of libcint's int2c2e_cart / int3c2e_cart.
DF_SETS = {
    "df_h2o_sto3g_jkfit_3c": ("3c", ["h2o_sto3g_jkfit_3c_cols0_166"], "h2o_sto3g_j"),
    "df_prim_spd_f_auxi_3c": ("3c", ["prim_spd_f_auxi_3c_cols0_35", "prim_spd_f_auxi_3c_cols35_84",
                                     "prim_spd_f_auxi_3c_cols84_119",
                                     "prim_spd_f_auxi_3c_cols119_140",
                                     "prim_spd_f_auxi_3c_cols140_168"], "prim_spd_f"),
    "df_prim_spd_f_auxi_metric": ("metric", ["prim_spd_f_auxi_metric_rows0_56",
                                             "prim_spd_f_auxi_metric_rows56_104",
                                             "prim_spd_f_auxi_metric_rows104_140",
                                             "prim_spd_f_auxi_metric_rows140_168"], None),
    "df_h2o_ccpvtz_jkfit_metric": ("metric", ["h2o_ccpvtz_jkfit_metric_rows0_43",
                                              "h2o_ccpvtz_jkfit_metric_rows43_81",
                                              "h2o_ccpvtz_jkfit_metric_rows81_121",
                                              "h2o_ccpvtz_jkfit_metric_rows121_166"], None),
}


def convert_df(kind: str, sources: list[str], parent: str | None) -> dict[str, np.ndarray]:
    """The auxiliary basis (aux_* arrays, the same converted convention as
    the AO basis), the AO basis of the parent fixture (three-center sets)
    and libcint's array assembled from its chunks: int2c2e [n_aux, n_aux]
    and int1e_ovlp [n_aux, n_aux], or int3c2e [n_ao, n_ao, n_aux]."""
    out: dict[str, np.ndarray] = {}
    parts, ovl = [], []
    for src in sources:
        fx = json.loads((SRC / f"{src}.json").read_text())
        sec = fx["metric" if kind == "metric" else "three_center"]
        aux = {"aux_" + k: v for k, v in _basis(sec["converted"]).items()}
        for k, v in aux.items():
            if k in out and not np.array_equal(out[k], v):
                raise ValueError(f"{src}: aux basis differs from the set's")
            out[k] = v
        arrays = fx["arrays"]
        if kind == "metric":
            parts.append(_arr(arrays["int2c2e_cart:rows"]))
            ovl.append(_arr(arrays["int1e_ovlp:rows"]))
        else:
            (key,) = [k for k in arrays if k.startswith("int3c2e_cart")]
            parts.append(_arr(arrays[key]))
    if kind == "metric":
        out["int2c2e"] = np.concatenate(parts, axis=0)
        out["int1e_ovlp"] = np.concatenate(ovl, axis=0)
    else:
        out["int3c2e"] = np.concatenate(parts, axis=2)
        fx = json.loads((SRC / f"{parent}.json").read_text())
        out.update(_basis(fx["converted"]))
    return out


# gradient sets: output name -> the source fixture (PySCF 2.14.0 rhf.Gradients
# at the fixture's converged density)
GRAD_SETS = {
    "grad_h2o_sto3g": "h2o_sto3g_grad",
    "grad_h2o_sto3g_d": "h2o_sto3g_d_grad",
    "grad_h2o_631g": "h2o_631g_grad",
    "grad_h2o_ccpvdz": "h2o_ccpvdz_grad",
    "grad_nh2f_631g": "nh2f_631g_grad",
}


def convert_grad(src: str) -> dict[str, np.ndarray]:
    fx = json.loads((SRC / f"{src}.json").read_text())
    out = _basis(fx["converted"])
    out["shell_atom"] = np.array([s["atom"] for s in fx["converted"]["shells"]], dtype=np.int32)
    out["atom_coord"] = np.array([a["coord"] for a in fx["atoms"]], dtype=np.float64)
    arrays = fx["arrays"]
    out["density"] = _arr(arrays["density"])
    out["grad_j"] = _arr(arrays["grad:j"])
    out["grad_k"] = _arr(arrays["grad:k"])
    return out


def main() -> None:
    for name, sources in SETS.items():
        data = convert(name, sources)
        np.savez_compressed(OUT / f"{name}.npz", **data)
        print(name, {k: v.shape for k, v in data.items()})
    for name, (kind, sources, parent) in DF_SETS.items():
        data = convert_df(kind, sources, parent)
        np.savez_compressed(OUT / f"{name}.npz", **data)
        print(name, {k: v.shape for k, v in data.items()})
    for name, src in GRAD_SETS.items():
        data = convert_grad(src)
        np.savez_compressed(OUT / f"{name}.npz", **data)
        print(name, {k: v.shape for k, v in data.items()})


if __name__ == "__main__":
    main()
