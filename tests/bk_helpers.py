"""Shared set-up for the BK parity tests: small synthetic workloads and the
oracle-side RK driver (numpy over the C oracle's right-hand side)."""

from __future__ import annotations

import math

import numpy as np

from oracle import oracle as O

# Butcher tableaux of the rapidity steps (S/bk/evolution.py:190-236 for
# Kutta's third-order step, :265-276 for the midpoint rule).
TABLEAUX = {
    "euler": ([[]], [1.0]),
    "rk2": ([[], [0.5]], [0.0, 1.0]),
    "kutta3": ([[], [0.5], [-1.0, 2.0]], [1 / 6, 4 / 6, 1 / 6]),
    "rk4": ([[], [0.5], [0.0, 0.5], [0.0, 0.0, 1.0]], [1 / 6, 2 / 6, 2 / 6, 1 / 6]),
}


def rk_step(N: np.ndarray, h: float, method: str, derivative) -> np.ndarray:
    a, b = TABLEAUX[method]
    K = []
    for s in range(len(b)):
        stage = N + h * sum(a[s][j] * K[j] for j in range(len(a[s]))) if s else N
        K.append(derivative(stage))
    return N + h * sum(bs * k for bs, k in zip(b, K))


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


def oracle_rhs(r, N, edges, table, n, p, *, first=0, nr=8, nt=24):
    """K1 and K2 + Kf per r node from the oracle (active set of S/bk/evolution.py:501)."""

    sN, sS = O.evolution_splines(r, N)
    active = np.nonzero((np.arange(r.size) >= 1) & (N <= 0.99999))[0]
    k1 = np.zeros(r.size)
    k2kf = np.zeros(r.size)
    if active.size:
        k1[active] = O.k1_fixed(r[active], sN, nr=nr, nt=nt, p=p)
        lo, width = O.nlo_bounds(r)
        spline = sN if p.stable else sS
        sums, _, _ = O.bk_nlo_sums(r[active], spline, edges[active], lo, width, table, first, n, p)
        k2kf[active] = sums[:, 0] / n
    return k1, k2kf


def oracle_evolve(r, N0, edges, table, n, p, *, steps, h, method, force_positive=True):
    N = N0.copy()
    out = [N.copy()]

    def derivative(stage):
        k1, k2kf = oracle_rhs(r, stage, edges, table, n, p)
        d = k1 + k2kf
        return np.where(np.isfinite(d), d, 0.0)

    for _ in range(steps):
        N = rk_step(N, h, method, derivative)
        if force_positive:
            N = np.clip(N, 0.0, 1.0)
        out.append(N.copy())
    return np.array(out)


def bk_params(stable: bool = False):
    return O.params(minr=1e-6, stable=stable)


__all__ = ["TABLEAUX", "bk_params", "math", "oracle_evolve", "oracle_rhs", "random_edges", "rk_step", "workload"]
