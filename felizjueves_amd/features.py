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
