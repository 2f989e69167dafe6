"""The flops-per-sample constant of the benchmark's roofline: the counted
formula of oracle/flop_count.py equals the C oracle's cancellation-free K2 + Kf
integrand by value (so the counted operations are the right ones), and its
count is the constant bench.py uses."""

from __future__ import annotations

import math

import numpy as np

import bench
from oracle import flop_count as F
from oracle import oracle as O

NC, NF = 3.0, 3.0
PHYS = dict(nf_over_nc=NF / NC, nc=NC, b0=(11 * NC - 2 * NF) / (12 * math.pi), log_mu0=10 * math.log(2.5),
            log_scale0=math.log(4.0 / 0.241**2))


def _workload():
    r = O.r_grid(1e-6, 1e2, 255)
    return r, O.mv(r, 0.2, 1.0, 1.0, 0.241)


def test_counted_formula_equals_oracle_integrand():
    r, N = _workload()
    sN, _ = O.evolution_splines(r, N)
    sp = F.LocalSpline(sN.x, sN.coef)
    lo = math.log(0.5 * r[0])
    w = math.log(2 * r[-1]) - lo
    rng = np.random.default_rng(11)
    got, ref = [], []
    for _ in range(400):
        p = int(rng.integers(1, 255))
        x = (lo + w * rng.random(), lo + w * rng.random(), 2 * math.pi * rng.random(), 2 * math.pi * rng.random())
        tot, _, _ = F.integrand(float(r[p]), math.log(r[p]), *map(F.Num, x), sp, **PHYS)
        got.append(tot.v)
        ref.append(O.nlo_integrand_stable([r[p]], [x[0]], [x[1]], [x[2]], [x[3]], sN,
                                          O.params(minr=1e-6, stable=True))[0][0])
    got, ref = np.array(got), np.array(ref)
    big = np.abs(ref) >= 1e-10 * np.abs(ref).max()
    assert big.sum() > 100
    np.testing.assert_allclose(got[big], ref[big], rtol=1e-9, atol=0)
    # the rest are below 1e-10 of the largest sample (differences of nearly equal N)
    assert np.all(np.abs(got - ref) <= 1e-8 * np.abs(ref).max())


def test_map_part_equals_oracle_map():
    bins = 50
    r, N = _workload()
    q = O.sobol_u32(O.sobol_table(4, 0), 0, 128).reshape(-1, 4)
    lo_r = math.log(0.5 * r[0])
    w_r = math.log(2.0 * r[-1]) - lo_r
    e = np.sort(np.concatenate([[0.0, 1.0], np.random.default_rng(3).random(bins - 1)]))
    X = np.empty((4, bins))
    W = np.empty((4, bins))
    lo = np.array([lo_r, lo_r, 0.0, 0.0])
    width = np.array([w_r, w_r, 2 * math.pi, 2 * math.pi])
    for d in range(4):
        X[d] = lo[d] + width[d] * e[:-1]
        W[d] = width[d] * (e[1:] - e[:-1])
    x_ref, jac_ref, _ = O.vegas_map(q, np.tile(e, (4, 1)), lo, width)
    for k in range(q.shape[0]):
        x, jac = F.vegas_map(q[k], bins, X, W)
        np.testing.assert_allclose([v.v for v in x], x_ref[k], rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(jac.v, jac_ref[k], rtol=1e-12)


def test_bench_constant_is_the_count():
    r, N = _workload()
    q = O.sobol_u32(O.sobol_table(4, 0), 0, 256).reshape(-1, 4)
    count = F.flops_per_sample(r, N, [5, 100, 200], q, **PHYS)
    assert bench.FLOPS_PER_SAMPLE == round(count), count
