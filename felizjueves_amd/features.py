"""Per-tile features of the precision selector (shared by the planner, the
error probe that makes training data and the network's inference).

A tile is one bra pair and a run of ket pairs of one class pair. Its 14
features are [la, lb, lc, ld, log10 min/max bra exponent sum, log10 min/max
ket exponent sum, |AB|, max |CD|, min |PQ| (zeta-weighted pair centres),
log10 Q_ab, log10 max Q_cd, log10 max dmax], where dmax is the six-block
density weight of the screen and Q the pair bounds.
"""

from __future__ import annotations

import numpy as np

N_FEATURES = 14
_TINY = 1e-300


def side_arrays(basis, pairs) -> dict:
    """Per-pair arrays of one side table, in table order."""
    zmin = np.array([p.rec[:, 6].min() for p in pairs])
    zmax = np.array([p.rec[:, 6].max() for p in pairs])
    sep = np.array([np.linalg.norm(basis.shell_centre[p.a] - basis.shell_centre[p.b]) for p in pairs])
    cen = np.array([(p.rec[:, :3] * p.rec[:, 6:7]).sum(0) / p.rec[:, 6].sum() for p in pairs])
    G = np.array([p.G for p in pairs])
    return {"lzmin": np.log10(zmin), "lzmax": np.log10(zmax), "sep": sep, "cen": cen.reshape(-1, 3),
            "lG": np.log10(np.maximum(G, _TINY))}


def tile_features(cls, bf: dict, kf: dict, b: int, ks: np.ndarray, starts: np.ndarray,
                  dmax: np.ndarray) -> np.ndarray:
    """Features [n_tiles, N_FEATURES] of the tiles ks[starts[i]:starts[i+1]]
    of bra pair b (dmax: the six-block weight of each ket in ks)."""
    n = len(starts)
    out = np.empty((n, N_FEATURES))
    out[:, 0:4] = cls
    out[:, 4] = bf["lzmin"][b]
    out[:, 5] = bf["lzmax"][b]
    out[:, 6] = np.minimum.reduceat(kf["lzmin"][ks], starts)
    out[:, 7] = np.maximum.reduceat(kf["lzmax"][ks], starts)
    out[:, 8] = bf["sep"][b]
    out[:, 9] = np.maximum.reduceat(kf["sep"][ks], starts)
    d = np.linalg.norm(kf["cen"][ks] - bf["cen"][b], axis=1)
    out[:, 10] = np.minimum.reduceat(d, starts)
    out[:, 11] = bf["lG"][b]
    out[:, 12] = np.maximum.reduceat(kf["lG"][ks], starts)
    out[:, 13] = np.log10(np.maximum(np.maximum.reduceat(dmax, starts), _TINY))
    return out


def item_kets(work: np.ndarray, ket_list: np.ndarray):
    """The ket pairs of every work item (bra, first, system|chunk, end) of
    one launch, flattened: (kets, starts, lengths) with item i's kets
    kets[starts[i]:starts[i] + lengths[i]] (a ket split into primitive
    chunks appears once per chunk)."""
    k0 = work[:, 1].astype(np.int64)
    lens = work[:, 3].astype(np.int64) - k0
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]]) if len(lens) else np.zeros(0, np.int64)
    idx = np.repeat(k0 - starts, lens) + np.arange(int(lens.sum()))
    return ket_list[idx].astype(np.int64) & 0x1ffff, starts, lens


def items_features(cls, bf: dict, kf: dict, work: np.ndarray, ket_list: np.ndarray) -> np.ndarray:
    """tile_features of every work item of one launch (its bra pair and the
    kets its lanes take), the density column left at zero (the device fills
    it per build from the current density)."""
    n = len(work)
    out = np.zeros((n, N_FEATURES))
    if n == 0:
        return out
    ks, starts, lens = item_kets(work, ket_list)
    b = work[:, 0].astype(np.int64)
    out[:, 0:4] = cls
    out[:, 4] = bf["lzmin"][b]
    out[:, 5] = bf["lzmax"][b]
    out[:, 6] = np.minimum.reduceat(kf["lzmin"][ks], starts)
    out[:, 7] = np.maximum.reduceat(kf["lzmax"][ks], starts)
    out[:, 8] = bf["sep"][b]
    out[:, 9] = np.maximum.reduceat(kf["sep"][ks], starts)
    d = np.linalg.norm(kf["cen"][ks] - bf["cen"][np.repeat(b, lens)], axis=1)
    out[:, 10] = np.minimum.reduceat(d, starts)
    out[:, 11] = bf["lG"][b]
    out[:, 12] = np.maximum.reduceat(kf["lG"][ks], starts)
    return out
