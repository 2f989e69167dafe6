"""Shared set-up for the BK parity tests: small synthetic workloads and the
oracle-side RK driver (numpy over the C oracle's right-hand side)."""

from __future__ import annotations

import math

import numpy as np

from oracle import oracle as O

TABLEAUX = O.TABLEAUX
rk_step = O.rk_step


def workload(R: int = 16, minr: float = 1e-6, maxr: float = 1e2, qs0sqr: float = 0.2):
    r = O.r_grid(minr, maxr, R - 1)
    N = O.mv(r, qs0sqr, 1.0, 1.0, 0.241)
    return r, N


def random_edges(R: int, bins: int, seed: int = 0) -> np.ndarray:
    """Non-uniform but valid grids (strictly increasing, 0 and 1 pinned)."""

    rng = np.random.default_rng(seed)
    w = rng.uniform(0.2, 1.0, (R, 4, bins))
    e = np.concatenate([np.zeros((R, 4, 1)), np.cumsum(w, axis=2)], axis=2)
    e /= e[:, :, -1:]
    e[:, :, 0] = 0.0
    e[:, :, -1] = 1.0
    return e


oracle_rhs = O.rhs
oracle_evolve = O.evolve


def bk_params(stable: bool = False):
    return O.params(minr=1e-6, stable=stable)


__all__ = ["TABLEAUX", "bk_params", "math", "oracle_evolve", "oracle_rhs", "random_edges", "rk_step", "workload"]
