"""Integration drivers for the DIS contributions on the device
(S/numerics/integration.py:12-18, 95-117, 120-589).

``vegas_device`` keeps the control flow of the reference's ``vegas()``: per
iteration one batch of samples through the current grid, the estimate and
its variance from sum w and sum w^2, grid adaptation after every iteration,
the inverse-variance combination and the convergence test.  Each iteration
is one ``nlo_dis_vegas`` launch (points, integrand, FP64 sums and histogram on
the GPU) plus one ``nlo_vegas_adapt``; the host reads two doubles per
iteration for the convergence test.  Sampling differs from the reference in
one respect: points are scrambled Sobol points (SobolEngine(dim, scramble=True,
seed), iteration k taking the next block of the sequence) instead of the
reference's torch.Generator draws, so results are deterministic per seed.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch


@dataclass(frozen=True, slots=True)
class IntegralResult:
    value: torch.Tensor
    error: torch.Tensor
    n_eval: int
    converged: bool
    seed: int | None


def vegas_device(desc, kind: str, *, samples_per_iteration: int, max_iterations: int = 8,
                 min_iterations: int = 3, bins: int = 32, epsrel: float = 0.2, epsabs: float = 0.0,
                 seed: int = 0, precision: str = "fp32", first_index: int = 0,
                 device: torch.device | str = "cuda") -> IntegralResult:
    """VEGAS over the unit cube of one DIS contribution (see the module docstring)."""

    from nlo_bk_optimized_amd import _lib
    from nlo_bk_optimized_amd import device as dev

    if samples_per_iteration < 2 or max_iterations < 1:
        raise ValueError("Vegas requires at least two samples and one iteration")
    if min_iterations < 1 or min_iterations > max_iterations:
        raise ValueError("min_iterations must lie between one and max_iterations")
    dim = _lib.DIS_DIMENSIONS[kind]
    table = dev.sobol_table(dim, seed)
    edges = torch.linspace(0, 1, bins + 1, dtype=torch.float64, device=device).expand(dim, -1).contiguous()
    stride = (samples_per_iteration + 63) // 64 * 64
    if first_index + stride * max_iterations > 1 << _lib.SOBOL_BITS:
        raise ValueError("the Sobol sequence (2^30 points) is too short for this budget")
    eps = float(torch.finfo(torch.float32 if precision == "fp32" else torch.float64).eps)
    n = samples_per_iteration
    estimates: list[float] = []
    variances: list[float] = []
    tiny = float(torch.finfo(torch.float32 if precision == "fp32" else torch.float64).tiny)
    value = error = None
    converged = False
    n_eval = 0
    for it in range(max_iterations):
        sums, hist, count = dev.dis_vegas(desc, kind, edges, table, n, first_index=first_index + it * stride,
                                          precision=precision)
        edges = dev.vegas_adapt(edges, hist, count, eps)
        n_eval += n
        total, total_sq = (float(v) for v in sums.cpu())
        estimate = total / n
        variance = max((max(total_sq - total * total / n, 0.0) / (n - 1)) / n, tiny)
        estimates.append(estimate)
        variances.append(variance)
        weights = [1.0 / v for v in variances]
        wsum = math.fsum(weights)
        value = math.fsum(w * e for w, e in zip(weights, estimates)) / wsum
        error = math.sqrt(1.0 / wsum)
        converged = len(estimates) >= min_iterations and error <= max(epsabs, epsrel * abs(value))
        if converged:
            break
    return IntegralResult(torch.tensor(value, dtype=torch.float64), torch.tensor(error, dtype=torch.float64),
                          n_eval, converged, seed)


def gauss_legendre_device(desc, kind: str, *, points: int = 24, epsrel: float = 1e-6, epsabs: float = 0.0,
                          precision: str = "fp32", device: torch.device | str = "cuda") -> IntegralResult:
    """Tensor Gauss-Legendre on the unit square with the half-order error check
    (tensor_gauss_legendre, S/numerics/integration.py:95-117)."""

    from nlo_bk_optimized_amd import device as dev

    out = dev.dis_gauss_legendre(desc, kind, points, precision=precision, device=device).cpu()
    value, coarse = float(out[0]), float(out[1])
    error = abs(value - coarse)
    coarse_points = max(2, points // 2)
    n_eval = points**2 + coarse_points**2
    return IntegralResult(torch.tensor(value, dtype=torch.float64), torch.tensor(error, dtype=torch.float64),
                          n_eval, error <= max(epsabs, epsrel * abs(value)), None)


__all__ = ["IntegralResult", "gauss_legendre_device", "vegas_device"]
