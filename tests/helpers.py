"""Shared test helpers: the oracle as an injected launcher, tolerances."""

from __future__ import annotations

import numpy as np

from oracle import oracle as O


class OracleRunner:
    """walk_f's injected `run` backed by the CPU oracle's f_launch."""

    def zeros(self, n):
        return np.zeros((n, n))

    def put(self, args):
        return args

    def __call__(self, cls, grid, block, args, F):
        O.f_launch(cls, args, F)

    def host(self, F):
        return F


def scale_G(basis, D, Q=None):
    """A magnitude scale for G: sum_{ls} |D_ls| times max |(mn|mn)|^(1/2)
    products, simplified to max|vj| style: used only for relative tolerances."""
    return float(np.abs(D).sum())
