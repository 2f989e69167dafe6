"""The fixed-Sobol oracle (oracle/fixed_sobol_oracle.py) pinned on the CPU.

* Sobol point sets: bit-exact against torch.quasirandom.SobolEngine.draw_base2
  in float32 and float64 (S/dis/fixed_sobol.py:87-93).
* map_importance_grid: the uniform-grid identity and mean endpoint weight of
  the reference's own test (T/tests/test_dis_production.py:119-150), and brute
  force over the bins of a non-uniform grid.
* paired_endpoint_inputs: the measure identity E[w1 g(first) + w2 g(second)]
  = integral of g over the unit cube, for g even in the angle, with and without
  the fold, against tensor Gauss-Legendre values; the same check run on a copy
  with a wrong second-sector Jacobian fails, so it pins that factor.
* learn_triple_importance_grid: the reference's own function's grids on a
  synthetic integrand (tests/golden/fixed_sobol_reference_vectors.npz).
* integrate_triple_fixed_sobol: the reference's known answer
  (T/tests/test_dis_production.py:153-192).
"""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import dis_oracle as D
from oracle import fixed_sobol_oracle as F

GOLDEN = Path(__file__).resolve().parent / "golden" / "fixed_sobol_reference_vectors.npz"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sobol_point_sets_equal_sobol_engine(dtype):
    tdtype = torch.float32 if dtype == np.float32 else torch.float64
    for seed in (3, 10_011):
        ours = F.sobol_points(seed, 256, dtype)
        ref = torch.quasirandom.SobolEngine(5, scramble=True, seed=seed).draw_base2(8, dtype=tdtype).numpy()
        assert ours.dtype == ref.dtype
        np.testing.assert_array_equal(ours, ref)


def test_uniform_grid_identity_and_mean_endpoint_weight():
    """T/tests/test_dis_production.py:119-150."""

    points = F.sobol_points(11 + 10_000, 512, np.float32)
    edges = np.tile(np.linspace(0, 1, 33, dtype=np.float32), (5, 1))
    mapped, inverse, _ = F.map_importance_grid(points, edges)
    np.testing.assert_array_equal(mapped, points)
    np.testing.assert_array_equal(inverse, np.ones_like(inverse))
    _, w1, _, w2 = F.paired_endpoint_inputs(points, 5.0)
    assert abs(float(np.mean(w1 + w2)) - 1.0) <= 1e-2


def test_map_matches_brute_force_bins():
    rng = np.random.default_rng(4)
    bins = 12
    inner = np.sort(rng.uniform(0, 1, (5, bins - 1)), axis=1)
    edges = np.concatenate([np.zeros((5, 1)), inner, np.ones((5, 1))], axis=1)
    points = F.sobol_points(8, 1024, np.float64)
    source, inverse, b = F.map_importance_grid(points, edges)
    for d in range(5):
        for i in range(0, 1024, 37):
            u = points[i, d]
            k = min(int(math.floor(u * bins)), bins - 1)
            assert b[i, d] == k
            frac = u * bins - k
            assert source[i, d] == edges[d, k] + (edges[d, k + 1] - edges[d, k]) * frac
    expect = np.prod(bins * (edges[np.arange(5)[None, :], b + 1] - edges[np.arange(5)[None, :], b]), axis=1)
    np.testing.assert_allclose(inverse, expect, rtol=1e-14)


def _gl_cube(g, n=48):
    x, w = np.polynomial.legendre.leggauss(n)
    x = (x + 1) / 2
    w = w / 2
    s, r, a = np.meshgrid(x, x, x, indexing="ij")
    ws = w[:, None, None] * w[None, :, None] * w[None, None, :]
    return float(np.sum(ws * g(s, r, a)))


# functions of (s2 = parent / maxr, x3 = |x02| / maxr, x4 = angle / 2 pi), even in the angle
_MEASURE_FUNCTIONS = (
    lambda s, r, a: np.ones_like(s),
    lambda s, r, a: r * (1 + s),
    lambda s, r, a: r**2 * (1 + np.cos(2 * np.pi * a)) * np.exp(-s),
    lambda s, r, a: (r - s) ** 2 * np.cos(2 * np.pi * a) ** 2 + s * r * np.cos(4 * np.pi * a),
)


def _split_mean(points, g, split, fold):
    first, w1, second, w2 = split(points, 5.0, fold_angle=fold)
    v = w1 * g(first[:, 2], first[:, 3], first[:, 4]) + w2 * g(second[:, 2], second[:, 3], second[:, 4])
    return float(np.mean(v))


@pytest.mark.parametrize("fold", [False, True])
def test_endpoint_split_preserves_the_measure(fold):
    points = F.sobol_points(3, 1 << 18, np.float64)
    for g in _MEASURE_FUNCTIONS:
        exact = _gl_cube(g)
        got = _split_mean(points, g, F.paired_endpoint_inputs, fold)
        assert abs(got - exact) <= 2e-4 * abs(exact), (got, exact)


def test_endpoint_measure_check_catches_a_wrong_jacobian():
    def wrong(source, maxr, fold_angle=False):
        first, w1, second, w2 = F.paired_endpoint_inputs(source, maxr, fold_angle=fold_angle)
        # a second-sector weight with the change-of-pole factor r1 / |x02| altered
        return first, w1, second, w2 * second[:, 3] * maxr / np.maximum(source[:, 3] * maxr, 1e-300)

    points = F.sobol_points(3, 1 << 16, np.float64)
    g = _MEASURE_FUNCTIONS[1]
    exact = _gl_cube(g)
    assert abs(_split_mean(points, g, wrong, True) - exact) > 1e-2 * abs(exact)


def _synthetic(x):
    """tests/golden/make_fixed_sobol_golden.py synthetic(), as the reference
    combines it: I1 = regular + float32(sensitive)."""

    phi = 2 * math.pi * x[:, 4]
    regular = x[:, 0] * (1 + x[:, 3]) * np.cos(phi) + 0.1
    sensitive = (0.01 * x[:, 1] * x[:, 2]).astype(np.float32).astype(np.float64)
    i2 = np.exp(-3 * x[:, 1]) * x[:, 2] - 0.05
    i3 = (x[:, 3] - 0.4) * x[:, 0] ** 2 * (1 + np.cos(phi) ** 2)
    return regular + sensitive, i2, i3


def test_learned_grids_match_reference_vectors():
    golden = np.load(GOLDEN)
    k = 0
    while f"case{k}_params" in golden:
        seed, sets, count, bins, fold, i1_weight, maxr = golden[f"case{k}_params"]
        training = [F.sobol_points(int(seed) + j, int(count), np.float64) for j in range(int(sets))]
        edges = F.learn_triple_importance_grid(_synthetic, training, maxr=float(maxr), bins=int(bins),
                                               fold_angle=bool(fold), i1_weight=float(i1_weight))
        np.testing.assert_allclose(edges, golden[f"case{k}_edges"], rtol=0, atol=1e-12)
        k += 1
    assert k == 3


def test_integrate_known_answer():
    """T/tests/test_dis_production.py:153-192."""

    def mapper(points, edges, maxr):
        return points, np.ones(points.shape[0]), points, np.zeros(points.shape[0])

    def i1_sums(x, w):
        return x[:, 0] * w, np.zeros(1)

    def i2_i3(x):
        return x[:, 1], -0.5 * x[:, 0]

    replicates = (np.array([[0.1, 0.8], [0.3, 0.4]]), np.array([[0.5, 0.2], [0.7, 0.6]]),
                  np.array([[0.2, 0.9], [0.4, 0.1]]))
    _, total = F.integrate_triple_fixed_sobol(i1_sums, i2_i3, None, replicates, maxr=5.0, epsrel=1.0,
                                              training_evaluations=10, endpoint_map_function=mapper)
    totals = np.array([0.7, 0.7, 0.65])
    assert total.value == pytest.approx(totals.mean(), abs=1e-15)
    assert total.error == pytest.approx(totals.std(ddof=1) / math.sqrt(3), rel=1e-12)
    assert total.n_eval == 22 and total.converged


def test_small_oracle_run_is_additive_and_finite():
    cfg = D.DisSetup(maxr=5.0, nf_alphas=1)
    contributions, total = F.gbw_qg_fixed_sobol(
        9.0, 1e-3, True, D.QuarkSpec(1.4, 2 / 3), D.GBWParams(qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0), cfg,
        seed=7, dtype=np.float64, training_samples=256, training_rounds=2, samples_per_replicate=256,
        replicates=3, bins=8)
    assert all(math.isfinite(c.value) and c.error >= 0 for c in contributions)
    assert total.value == pytest.approx(sum(c.value for c in contributions), rel=1e-12)
    assert total.n_eval == 2 * 2 * 256 + 2 * 3 * 256
