"""The oracle against mathematics and third-party routines: scipy Bessel and
splines, 50-digit mpmath kernels, torch's SobolEngine, closed-form VEGAS
identities, adaptive quadrature for the fixed K1 rule, and RK known answers
(including the reference's own constant-derivative test,
T/tests/test_bk_production.py:22-49)."""

from __future__ import annotations

import math

import mpmath as mp
import numpy as np
import pytest
import scipy.integrate
import scipy.interpolate
import scipy.special
import torch

from oracle import oracle as O
from tests.bk_helpers import TABLEAUX, rk_step, workload


def test_bessel_against_scipy():
    x = np.concatenate([np.linspace(1e-6, 70, 3000), [1e-300, 1e-10, 0.5]])
    np.testing.assert_allclose(O.bessel_j1(x), scipy.special.j1(x), rtol=0, atol=2e-15)
    np.testing.assert_allclose(O.bessel_i1(x), scipy.special.i1(x), rtol=5e-15, atol=1e-300)


def _kernels_mp(r, z, z2, t1, t2):
    X = mp.sqrt((r - z) ** 2 + 4 * r * z * mp.sin(t1 / 2) ** 2)
    X2 = mp.sqrt((r - z2) ** 2 + 4 * r * z2 * mp.sin(t2 / 2) ** 2)
    Z = mp.sqrt((z - z2) ** 2 + 4 * z * z2 * mp.sin((t1 - t2) / 2) ** 2)

    def k2(X, Y, X2, Y2):
        a, b = X * Y2, X2 * Y
        L = mp.log(a / b)
        return -2 / Z**4 + ((a * a + b * b - 4 * (r * Z) ** 2) / (Z**4 * (a * a - b * b))
                            + r**4 / (a * a * (a * a - b * b)) + r * r / (a * Z) ** 2) * 2 * L

    def kf(X, Y, X2, Y2):
        a, b = X * Y2, X2 * Y
        return 2 / Z**4 - (a * a + b * b - (r * Z) ** 2) / (Z**4 * (a * a - b * b)) * 2 * mp.log(a / b)

    J = (z * z2) ** 2
    return [k2(X, z, X2, z2) * J, k2(X2, z2, X, z) * J, kf(X, z, X2, z2) * J, kf(X2, z2, X, z) * J]


def test_stable_kernels_against_mpmath():
    import ctypes

    mp.mp.dps = 50
    fn = O.lib().orc_nlo_kernels_stable
    fn.argtypes = [ctypes.c_double] * 7 + [ctypes.POINTER(ctypes.c_double)]
    rng = np.random.default_rng(11)
    worst = np.zeros(4)
    for _ in range(150):
        r, z, z2 = np.exp(rng.uniform(-13, 4)), np.exp(rng.uniform(-14, 5)), np.exp(rng.uniform(-14, 5))
        t1, t2 = rng.uniform(0, 2 * math.pi, 2)
        exact = _kernels_mp(*(mp.mpf(float(v)) for v in (r, z, z2, t1, t2)))
        out = (ctypes.c_double * 4)()
        fn(r, z, t1, z2, t2, 3.0, 3.0, out)
        for k in range(4):
            worst[k] = max(worst[k], abs(out[k] - float(exact[k])) / abs(float(exact[k])))
    # K2 J keeps a genuine sum of four terms (some cancel near its zeros)
    assert worst[0] < 1e-6 and worst[1] < 1e-6, worst
    assert worst[2] < 1e-12 and worst[3] < 1e-12, worst


def test_reference_formula_loses_accuracy_where_stable_does_not():
    # far region z, z' >> r: the reference's K2 cancels -2/Z^4 against its
    # log term; the stable form stays within 1e-6 of 50-digit arithmetic
    import ctypes

    mp.mp.dps = 50
    fn = O.lib().orc_nlo_kernels_stable
    fn.argtypes = [ctypes.c_double] * 7 + [ctypes.POINTER(ctypes.c_double)]
    r, z, z2, t1, t2 = 1e-6, 0.3, 0.7, 1.1, 2.3
    exact = float(_kernels_mp(*(mp.mpf(v) for v in (r, z, z2, t1, t2)))[0])
    out = (ctypes.c_double * 4)()
    fn(r, z, t1, z2, t2, 3.0, 3.0, out)
    X = math.sqrt((r - z) ** 2 + 4 * r * z * math.sin(t1 / 2) ** 2)
    X2 = math.sqrt((r - z2) ** 2 + 4 * r * z2 * math.sin(t2 / 2) ** 2)
    Z = math.sqrt((z - z2) ** 2 + 4 * z * z2 * math.sin((t1 - t2) / 2) ** 2)
    direct = O.kernel_nlo([r], [X], [z], [X2], [z2], [Z])[0] * (z * z2) ** 2
    assert abs(out[0] - exact) <= 1e-6 * abs(exact)
    assert abs(direct - exact) > 1e-3 * abs(exact)


@pytest.mark.parametrize("dim,seed", [(4, 0), (5, 123)])
def test_sobol_gray_code_matches_sobolengine(dim, seed):
    t = O.sobol_table(dim, seed)
    n = 5000
    q = O.sobol_u32(t, 0, n)
    f64 = torch.quasirandom.SobolEngine(dim, scramble=True, seed=seed).draw(n, dtype=torch.float64).numpy()
    f32 = torch.quasirandom.SobolEngine(dim, scramble=True, seed=seed).draw(n, dtype=torch.float32).numpy()
    np.testing.assert_array_equal(q[1:] / 2.0**30, f64[1:])
    np.testing.assert_array_equal(O.sobol_f64(t, 0, n), f64)  # including the first-point rounding
    np.testing.assert_array_equal(q.astype(np.float32) * np.float32(2.0**-30), f32)
    # offsets continue the same sequence
    np.testing.assert_array_equal(O.sobol_u32(t, 1234, 100), q[1234:1334])


def test_vegas_map_closed_forms():
    rng = np.random.default_rng(0)
    dim, bins, n = 4, 50, 3000
    q = rng.integers(0, 2**30, (n, dim), dtype=np.int64).astype(np.uint32)
    q[0] = 2**30 - 1
    edges = np.sort(rng.uniform(0, 1, (dim, bins + 1)), axis=1)
    edges[:, 0], edges[:, -1] = 0.0, 1.0
    lo = rng.uniform(-2, 2, dim)
    width = rng.uniform(0.5, 3, dim)
    x, jac, b = O.vegas_map(q, edges, lo, width)
    u = q / 2.0**30
    np.testing.assert_array_equal(b, np.minimum(np.floor(u * bins), bins - 1).astype(np.int32))
    frac = u * bins - b
    left = np.take_along_axis(edges.T, b, 0) if False else edges[np.arange(dim), b]
    right = edges[np.arange(dim), b + 1]
    np.testing.assert_allclose(x, lo + width * (left + (right - left) * frac), rtol=1e-15)
    np.testing.assert_allclose(jac, np.prod(width) * np.prod(bins * (right - left), axis=1), rtol=1e-14)
    # uniform grid: identity map, unit density (T/tests/test_dis_production.py:139-143)
    uni = np.tile(np.linspace(0, 1, bins + 1), (dim, 1))
    x, jac, _ = O.vegas_map(q, uni, np.zeros(dim), np.ones(dim))
    np.testing.assert_allclose(x, u, rtol=0, atol=1e-15)
    np.testing.assert_allclose(jac, 1.0, rtol=1e-13)


def test_vegas_adapt_keeps_uniform_grid_for_flat_importance():
    bins = 50
    edges = np.tile(np.linspace(0, 1, bins + 1), (4, 1))
    hist = np.full((4, bins), 3.0)
    count = np.full((4, bins), 3.0)
    new = O.vegas_adapt(edges, hist, count, float(np.finfo(np.float64).eps))
    np.testing.assert_allclose(new, edges, rtol=0, atol=1e-15)


def test_vegas_adapt_moves_edges_toward_importance():
    bins = 32
    edges = np.tile(np.linspace(0, 1, bins + 1), (1, 1))
    hist = np.ones((1, bins))
    hist[0, :4] = 100.0  # importance piled in the first bins
    new = O.vegas_adapt(edges, hist, np.ones((1, bins)), float(np.finfo(np.float64).eps))
    assert new[0, 0] == 0.0 and new[0, -1] == 1.0
    assert np.all(np.diff(new[0]) > 0)
    assert new[0, bins // 2] < 4.0 / bins  # half of the new bins inside the first 4 old ones


def test_natural_spline_matches_scipy():
    rng = np.random.default_rng(1)
    x = np.sort(rng.uniform(-5, 5, 40))
    y = np.sin(x) + 0.1 * x**2
    coef = O.natural_spline(x, y)
    cs = scipy.interpolate.CubicSpline(x, y, bc_type="natural")
    np.testing.assert_allclose(coef[0], y[:-1], rtol=0, atol=0)
    np.testing.assert_allclose(coef[1], cs.c[2], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(coef[2], cs.c[1], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(coef[3], cs.c[0], rtol=1e-10, atol=1e-11)


def test_natural_spline_exact_on_lines():
    x = np.linspace(0, 3, 7)
    coef = O.natural_spline(x, 2.5 * x - 1.0)
    np.testing.assert_allclose(coef[1], 2.5, rtol=1e-15)
    np.testing.assert_allclose(coef[2:], 0.0, atol=1e-15)


def test_loglog_spline_range_rules():
    r, N = workload(R=16)
    sN, sS = O.evolution_splines(r, N)
    # below / above the grid: N -> 0 / 1 and S -> 1 / 0 (S/bk/derivatives.py:761-772)
    assert O.loglog_eval(sN, [r[0] * 0.5], 0.0, 1.0)[0] == 0.0
    assert O.loglog_eval(sN, [r[-1] * 2.0], 0.0, 1.0)[0] == 1.0
    assert O.loglog_eval(sS, [r[0] * 0.5], 1.0, 0.0)[0] == 1.0
    # at the nodes the spline reproduces N
    np.testing.assert_allclose(O.loglog_eval(sN, r[1:-1], 0.0, 1.0), N[1:-1], rtol=1e-13)


# production; BEUF with parent resummation and Nf = 5; FAC with the LO order
K1_RULE_CASES = ({}, {"rc_lo": 5, "resum_rc": 1, "nf": 5.0}, {"rc_lo": 4, "order": 5})


@pytest.mark.parametrize("options", K1_RULE_CASES)
def test_k1_fixed_rule_against_adaptive_quadrature(options):
    r, N = workload(R=24)
    sN, _ = O.evolution_splines(r, N)
    p = O.params(minr=1e-6, **options)
    for idx in ((8, 16) if not options else (16,)):
        rp = r[idx]
        fixed = O.k1_fixed(np.array([rp]), sN, p=p)[0]

        def inner(log_z):
            z = math.exp(log_z)
            f = lambda th: O.k1_integrand([rp], [z], [th], sN, p)[0]  # noqa: E731
            # split theta at the X -> 0 corner
            val = scipy.integrate.quad(f, 0.0, math.pi, limit=200, epsabs=0, epsrel=1e-9, points=[1e-3])[0]
            return val * 2 * z * z

        pts = [math.log(rp)]
        ref = scipy.integrate.quad(inner, math.log(r[0]), math.log(r[-1]), limit=200, epsabs=0,
                                   epsrel=1e-7, points=pts)[0]
        assert fixed == pytest.approx(ref, rel=2e-3), (idx, fixed, ref)


def test_rk_known_answers():
    # T/tests/test_bk_production.py:22-49: constant derivative 2, h = 0.2 -> N + 0.4
    N = np.array([0.1, 0.5])
    for method in TABLEAUX:
        np.testing.assert_allclose(rk_step(N, 0.2, method, lambda n: np.full_like(n, 2.0)), N + 0.4)
    # dN/dy = N: one step reproduces the Taylor series to the method's order
    h = 0.3
    taylor = [1, h, h**2 / 2, h**3 / 6, h**4 / 24]
    for method, order in (("euler", 1), ("rk2", 2), ("kutta3", 3), ("rk4", 4)):
        got = rk_step(np.array([1.0]), h, method, lambda n: n)[0]
        assert got == pytest.approx(sum(taylor[: order + 1]), rel=1e-15), method
