"""Adaptive RK2 step-size controller of the BK rapidity driver, restated in numpy.

TEST INFRASTRUCTURE ONLY: the reference the native controller is tested
against.  Only ``tests/`` may import this module; the product package never
does.

``adaptive_rk2_to`` restates ``_adaptive_rk2_to`` (S/bk/evolution.py:239-326,
S/ = torch_nlo_clean/src/nlo_torch/): midpoint RK2 with step doubling, the
error |two half steps - one full step| / 3 against DE_SOLVER_ABSERR +
DE_SOLVER_RELERR |N|, growth 0.9 ratio^(-1/3) capped at 5 (5 for a zero
error), shrink floored at 0.1, and a step-underflow error.  Pinned by the
reference's own controller driven by a closed-form right-hand side
(tests/golden/bk_reference_vectors.npz, ``ark2_*``).
"""

from __future__ import annotations

from collections.abc import Callable

import numpy as np


def adaptive_rk2_to(f: Callable[[np.ndarray, float], np.ndarray], N: np.ndarray, y: float, nexty: float,
                    h: float, abserr: float, relerr: float, eps: float, stats: dict | None = None):
    """Advance N from y to nexty; returns (N, y, h).  ``f(N, y)`` is dN/dy."""

    while y < nexty:
        step = min(h, nexty - y)
        k1 = f(N, y)
        k2 = f(N + 0.5 * step * k1, y + 0.5 * step)
        full = N + step * k2
        half = step / 2
        first = f(N + 0.5 * half * k1, y + 0.5 * half)
        half_N = N + half * first
        second_k1 = f(half_N, y + half)
        second_k2 = f(half_N + 0.5 * half * second_k1, y + half + 0.5 * half)
        two_half = half_N + half * second_k2
        error = np.abs(two_half - full) / 3
        tolerance = abserr + relerr * np.abs(two_half)
        ratio = float(np.max(error / tolerance))
        if ratio <= 1:
            N = two_half
            y += step
            factor = 5.0 if ratio == 0 else min(5.0, 0.9 * ratio ** (-1 / 3))
            h = step * factor
            if stats is not None:
                stats["accepted"] = stats.get("accepted", 0) + 1
        else:
            h = step * max(0.1, 0.9 * ratio ** (-1 / 3))
            if stats is not None:
                stats["rejected"] = stats.get("rejected", 0) + 1
            if h <= eps * max(1.0, abs(y)):
                raise RuntimeError(f"BK RK2 step underflow at y={y}")
    return N, y, h
