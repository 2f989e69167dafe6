"""numpy restatement of the reference's GBW fixed-Sobol qqbar-g estimator.

TEST INFRASTRUCTURE ONLY: the reference the HIP fixed-Sobol path is tested
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py`` may
import this module; the product package never imports it.

Each function restates the reference function it names (S/ =
torch_nlo_clean/src/nlo_torch/), dtype-generic as the reference's torch code:
float32 inputs are computed in float32 operation by operation (numpy has no
fused multiply-add), float64 in float64.  Two steps are run in FP64 whatever
the dtype, as the HIP path runs them: the grid refinement (the C restatement
``oracle.vegas_adapt`` with the dtype's epsilon; the reference refines float32
grids in float32) and the integrand values (``dis_oracle.qg_integrand``; the
reference's float32 kernels evaluate I2 / I3 in FP32 and I1 in split
precision).

Parity status:

* ``map_importance_grid``: pinned by the uniform-grid identity of the
  reference's own test (T/tests/test_dis_production.py:119-150) and by brute
  force over the bins of a non-uniform grid;
* ``paired_endpoint_inputs``: pinned by the measure identity the split must
  satisfy, E[w1 g(first) + w2 g(second)] = integral of g over the unit cube for
  every g even in the angle, checked against tensor Gauss-Legendre values of
  the right-hand side, and by the reference's mean-weight check
  (T/tests/test_dis_production.py:144-150);
* ``learn_triple_importance_grid``: pinned by golden vectors of the reference's
  own function on a synthetic integrand (tests/golden/fixed_sobol_reference_vectors.npz,
  written by tests/golden/make_fixed_sobol_golden.py);
* ``integrate_triple_fixed_sobol``: pinned by the reference's known answer
  (T/tests/test_dis_production.py:153-192);
* ``gbw_qg_fixed_sobol``: the composition of the above with the integrands of
  ``dis_oracle`` (pinned there by the reference's vectors); its end-to-end values
  are checked statistically against the reference's C++ values
  (T/benchmarks/dis_fixed_sobol_all_qg.py:69-76) in the GPU tests.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from oracle import dis_oracle as D
from oracle import oracle as O

TRAINING_SAMPLES = 65_536
TRAINING_ROUNDS = 3
SAMPLES_PER_REPLICATE = 32_768
REPLICATES = 7
GRID_BINS = 32


@dataclass(frozen=True)
class Estimate:
    value: float
    error: float
    n_eval: int
    converged: bool


def sobol_points(seed: int, count: int, dtype, dim: int = 5) -> np.ndarray:
    """SobolEngine(dim, scramble=True, seed).draw_base2(log2 count, dtype)
    (S/dis/fixed_sobol.py:87-93): float32 values are float32(q) 2^-30 (the
    engine converts the integer state, then scales), float64 q 2^-30; point 0 is
    the engine's first point."""

    t = O.sobol_table(dim, seed)
    q = O.sobol_u32(t, 0, count)
    if count > 0:
        q = q.copy()
        q[0] = t.first_q
    if np.dtype(dtype) == np.float32:
        return q.astype(np.float32) * np.float32(2.0**-30)
    return q.astype(np.float64) * 2.0**-30


def map_importance_grid(points: np.ndarray, edges: np.ndarray):
    """S/dis/fixed_sobol.py:100-113: (source, inverse density, bin index)."""

    dt = points.dtype.type
    bins = edges.shape[1] - 1
    scaled = points * dt(bins)
    bin_index = np.minimum(np.floor(scaled).astype(np.int64), bins - 1)
    fraction = scaled - bin_index.astype(points.dtype)
    rows = np.arange(points.shape[1])[None, :]
    left = edges[rows, bin_index]
    right = edges[rows, bin_index + 1]
    width = right - left
    source = left + width * fraction
    inverse = np.ones(points.shape[0], dtype=points.dtype)
    for d in range(points.shape[1]):
        inverse = inverse * (dt(bins) * width[:, d])
    return source, inverse, bin_index


def paired_endpoint_inputs(source: np.ndarray, maxr: float, *, fold_angle: bool = False):
    """S/dis/fixed_sobol.py:116-158: (first, first weight, second, second
    weight) without the grid's inverse density."""

    dt = source.dtype.type
    tiny = np.finfo(source.dtype).tiny
    two_pi = dt(2 * math.pi)
    m = dt(maxr)
    parent = m * source[:, 2]
    radial = source[:, 3]
    angle_fraction = dt(0.5) * source[:, 4] if fold_angle else source[:, 4]
    angle = two_pi * angle_fraction
    cosine = np.cos(angle)
    sine = np.sin(angle)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        closer = np.where(cosine > 0, parent / (dt(2) * np.maximum(cosine, tiny)), dt(np.inf))
    first_limit = np.minimum(m, closer)
    first = source.copy()
    first[:, 3] = radial * first_limit / m
    first[:, 4] = angle_fraction
    first_weight = first_limit / m
    disk = parent * cosine + np.sqrt(np.maximum(dt(maxr * maxr) - (parent * parent) * (sine * sine), dt(0)))
    second_limit = np.maximum(np.minimum(disk, closer), dt(0))
    second_radius = radial * second_limit
    gx = parent - second_radius * cosine
    gy = second_radius * sine
    gr = np.sqrt(gx * gx + gy * gy)
    ga = np.arctan2(gy, gx)
    ga = np.where(ga < 0, ga + two_pi, ga)
    second = source.copy()
    second[:, 3] = gr / m
    second[:, 4] = ga / two_pi
    second_weight = second_radius * second_limit / (np.maximum(gr, tiny) * m)
    return first, first_weight, second, second_weight


def endpoint_map(points: np.ndarray, edges: np.ndarray, maxr: float, *, fold_angle: bool = True):
    """dis_sobol_endpoint_map (K/dis/sampling/kernel.cu:15-81): the map and the
    split, the weights times the inverse density."""

    source, inverse, _ = map_importance_grid(points, edges)
    first, w1, second, w2 = paired_endpoint_inputs(source, maxr, fold_angle=fold_angle)
    return first, inverse * w1, second, inverse * w2


def _adapt(edges: np.ndarray, histogram: np.ndarray, count: np.ndarray, dtype) -> np.ndarray:
    eps = float(np.finfo(dtype).eps)
    return O.vegas_adapt(edges.astype(np.float64), histogram, count, eps)


def learn_triple_importance_grid(function, training_points, *, maxr: float, bins: int = GRID_BINS,
                                 fold_angle: bool = True, i1_weight: float = 1.0) -> np.ndarray:
    """S/dis/fixed_sobol.py:161-210.  ``function(x)`` returns (I1, I2, I3) at
    the points x; the grid is returned in float64 (used in the points'
    dtype)."""

    if not training_points:
        raise ValueError("at least one independent training point set is required")
    dt = training_points[0].dtype
    dims = training_points[0].shape[1]
    edges = np.tile(np.arange(bins + 1, dtype=np.float64) / bins, (dims, 1))
    tiny = np.finfo(dt).tiny
    for points in training_points:
        source, inverse, bin_index = map_importance_grid(points, edges.astype(dt))
        first, w1, second, w2 = paired_endpoint_inputs(source, maxr, fold_angle=fold_angle)
        fv = function(first)
        sv = function(second)
        weighted = [(np.asarray(a) * w1 + np.asarray(b) * w2) * inverse for a, b in zip(fv, sv)]
        weights = (i1_weight, 1.0, 1.0)
        score = 0
        for wt, v in zip(weights, weighted):
            score = score + wt * np.abs(v) / max(float(np.abs(v).mean()), tiny)
        score = np.asarray(score, dtype=np.float64)
        histogram = np.zeros((dims, bins))
        bin_count = np.zeros((dims, bins))
        for d in range(dims):
            histogram[d] = np.bincount(bin_index[:, d], weights=score, minlength=bins)
            bin_count[d] = np.bincount(bin_index[:, d], minlength=bins)
        edges = _adapt(edges, histogram, bin_count, dt)
    return edges


def integrate_triple_fixed_sobol(i1_sum_function, nested_function, edges, replicate_points, *, maxr: float,
                                 epsrel: float, training_evaluations: int = 0, endpoint_map_function=None):
    """S/dis/fixed_sobol.py:213-271: (contributions I1, I2, I3, total) as
    Estimates.  ``i1_sum_function(x, weights)`` returns (regular sums,
    sensitive sums); ``nested_function(x)`` returns (I2, I3) values;
    ``endpoint_map_function(points, edges, maxr)`` defaults to endpoint_map."""

    if len(replicate_points) < 2:
        raise ValueError("at least two scrambled Sobol replicates are required")
    mapper = endpoint_map_function or (lambda p, e, m: endpoint_map(p, e, m, fold_angle=True))
    estimates = []
    for points in replicate_points:
        first, w1, second, w2 = mapper(points, edges, maxr)
        fr, fs = i1_sum_function(first, w1)
        sr, ss = i1_sum_function(second, w2)
        f2, f3 = nested_function(first)
        s2, s3 = nested_function(second)
        n = points.shape[0]
        regular = (np.sum(fr) + np.sum(sr)) / n
        sensitive = (np.sum(fs) + np.sum(ss)) / n
        estimates.append([float(regular + sensitive),
                          float(np.mean((np.asarray(f2) * w1 + np.asarray(s2) * w2).astype(np.float64))),
                          float(np.mean((np.asarray(f3) * w1 + np.asarray(s3) * w2).astype(np.float64)))])
    stacked = np.asarray(estimates, dtype=np.float64)
    samples = training_evaluations + 2 * sum(p.shape[0] for p in replicate_points)

    def result(values: np.ndarray) -> Estimate:
        value = float(values.mean())
        error = float(values.std(ddof=1) / math.sqrt(values.size))
        return Estimate(value, error, samples, error <= epsrel * abs(value))

    contributions = tuple(result(stacked[:, k]) for k in range(3))
    return contributions, result(stacked.sum(axis=1))


def gbw_qg_fixed_sobol(Q2: float, xbj: float, transverse: bool, quark: D.QuarkSpec, gbw: D.GBWParams,
                       setup: D.DisSetup, *, seed: int, dtype=np.float32, training_samples: int = TRAINING_SAMPLES,
                       training_rounds: int = TRAINING_ROUNDS, samples_per_replicate: int = SAMPLES_PER_REPLICATE,
                       replicates: int = REPLICATES, bins: int = GRID_BINS, nested_points: int = 48,
                       i1_weight: float = 1.0, epsrel: float = 1e-3, return_edges: bool = False):
    """S/dis/fixed_sobol.py:274-368 with the integrands of dis_oracle:
    (contributions, total) and, with return_edges, the learned grid."""

    def amp(r, Y):
        return D.gbw(r, Y, gbw)

    def i1(x):
        return D.qg_integrand(x.astype(np.float64), Q2, xbj, transverse, quark, "I1", amp, setup)

    def nested(x):
        x64 = x.astype(np.float64)
        return tuple(D.qg_integrand(x64, Q2, xbj, transverse, quark, c, amp, setup, nested_points=nested_points)
                     for c in ("I2", "I3"))

    def values(x):
        return (i1(x),) + nested(x)

    training = [sobol_points(seed + k, training_samples, dtype) for k in range(training_rounds)]
    reps = [sobol_points(seed + 10_000 + k, samples_per_replicate, dtype) for k in range(replicates)]
    edges = learn_triple_importance_grid(values, training, maxr=setup.maxr, bins=bins, i1_weight=i1_weight)
    grid = edges.astype(dtype)

    def i1_sums(x, w):
        return i1(x) * w, np.zeros(1)

    contributions, total = integrate_triple_fixed_sobol(
        i1_sums, nested, grid, reps, maxr=setup.maxr, epsrel=epsrel,
        training_evaluations=2 * training_rounds * training_samples)
    return (contributions, total, edges) if return_edges else (contributions, total)
