"""The oracle's adaptive Gauss-Kronrod rule and adaptive K1, pinned without
the reference: the single-interval rule against scipy's QUADPACK dqk21
(scipy.integrate.quad with limit=1), the Kronrod and Gauss exactness degrees,
closed-form integrals, and the adaptive K1 against the fixed Gauss-Legendre
rule (the same integral by an independent quadrature)."""

import math

import numpy as np
import pytest
import scipy.integrate as si

from oracle import oracle as O


@pytest.mark.parametrize(
    "f, a, b",
    [
        (np.exp, 0.0, 1.0),
        (lambda x: 1.0 / (1.0 + x * x), -3.0, 2.0),
        (np.sqrt, 0.0, 1.0),
        (lambda x: x**20, -1.0, 1.0),
        (lambda x: math.cos(30.0 * x), 0.0, 1.0),
    ],
)
def test_single_interval_matches_quadpack_dqk21(f, a, b):
    """One interval (max_intervals = 1): estimate and error of
    _gauss_kronrod_intervals (S/numerics/integration.py:624-746) equal
    QUADPACK's dqk21, which scipy.integrate.quad returns for limit=1."""

    v, e, _, n = O.adaptive_gk21(f, a, b, 1e-15, 1)
    ref, ref_err, *_ = si.quad(f, a, b, limit=1, epsabs=1e-300, epsrel=1e-300, full_output=1)
    assert n == 21
    assert abs(v - ref) <= 4e-16 * max(1.0, abs(ref))
    assert abs(e - ref_err) <= 1e-10 * ref_err


def test_kronrod_and_gauss_exactness():
    """GK21 integrates polynomials to degree 31 exactly; the embedded 10-point
    Gauss rule to degree 19, so |K - G| vanishes below degree 20 only."""

    for k in range(32):
        v, e, _, _ = O.adaptive_gk21(lambda x, k=k: x**k, -1.0, 1.0, 1e-15, 1)
        exact = (1.0 - (-1.0) ** (k + 1)) / (k + 1)
        assert abs(v - exact) <= 1e-15, k
        if k < 20:
            assert e <= 1e-13, k
    _, e20, _, _ = O.adaptive_gk21(lambda x: x**20, -1.0, 1.0, 1e-15, 1)
    assert e20 > 1e-6


@pytest.mark.parametrize(
    "f, a, b, exact",
    [
        (math.sin, 0.0, math.pi, 2.0),
        (lambda x: 1.0 / math.sqrt(x) if x > 0 else 0.0, 0.0, 1.0, 2.0),
        (lambda x: math.log(x) if x > 0 else 0.0, 0.0, 1.0, -1.0),
        (lambda x: 1.0 / (1e-4 + x * x), -1.0, 1.0, 2.0 * math.atan(100.0) * 100.0),
    ],
)
def test_adaptive_rule_converges_to_closed_forms(f, a, b, exact):
    """Bisection of the largest-error interval (adaptive_gauss_kronrod_1d,
    S/numerics/integration.py:43-92): the result meets its own tolerance and
    lies within its error estimate; evaluations = 21 + 42 per bisection."""

    v, e, conv, n = O.adaptive_gk21(f, a, b, 1e-8, 85)
    assert conv
    assert e <= 1e-8 * abs(v)
    assert abs(v - exact) <= e + 1e-14 * abs(exact)
    assert (n - 21) % 42 == 0


def test_adaptive_rule_interval_limit():
    """With the interval limit reached the result is returned unconverged."""

    f = lambda x: 1.0 / math.sqrt(x) if x > 0 else 0.0  # noqa: E731
    v, e, conv, n = O.adaptive_gk21(f, 0.0, 1.0, 1e-14, 5)
    assert not conv
    assert n == 21 + 4 * 42
    assert abs(v - 2.0) < 0.1


def _mv_setup(R, minr, maxr):
    r = np.exp(np.linspace(np.log(minr), np.log(maxr), R))
    N = np.array([O.lib().orc_mv(x, 0.2, 1.0, 1.0, 0.241) for x in r])
    sN, _ = O.evolution_splines(r, N)
    return r, N, sN


def test_adaptive_k1_agrees_with_fixed_rule():
    """rapidity_derivative_lo's nested rule (S/bk/derivatives.py:43-110) and the
    fixed 8 x 24 Gauss-Legendre rule are two quadratures of one integral: at
    the reference's tolerance (INTACCURACY = 1e-3) they agree to it.  Tightened,
    the adaptive value is stable to its error estimates (1e-8 vs 1e-10) and
    meets a 32 x 64 fixed rule to 5e-6 (the fixed rule converges only
    algebraically: N is a C2 spline in ln X, so theta sees kinks)."""

    r, N, sN = _mv_setup(101, 1e-6, 30.0)
    p = O.params(minr=1e-6)
    idx = np.array([3, 20, 45, 70, 85, 95])
    ka, ea, ev = O.k1_adaptive(r[idx], sN, p=p)
    kf = O.k1_fixed(r[idx], sN, p=p)
    assert np.all(np.abs(ka - kf) <= 1e-3 * np.abs(kf))
    assert np.all(ea <= 1e-3 * np.abs(ka))
    assert np.all((ev > 0))
    idx = np.array([20, 70, 85])
    k8, e8, _ = O.k1_adaptive(r[idx], sN, rint=100, tint=100, epsrel=1e-8, p=p)
    k10, e10, _ = O.k1_adaptive(r[idx], sN, rint=100, tint=100, epsrel=1e-10, p=p)
    assert np.all(np.abs(k8 - k10) <= e8 + e10)
    kf2 = O.k1_fixed(r[idx], sN, nr=32, nt=64, p=p)
    assert np.all(np.abs(k10 - kf2) <= 5e-6 * np.abs(kf2)), (k10, kf2)


def test_adaptive_k1_other_physics_agrees_with_fixed_rule():
    """The same for a non-production choice (parent coupling, LO with the
    double-log resummation, Nf = 5 thresholds)."""

    r, N, sN = _mv_setup(64, 1e-6, 1e2)
    p = O.params(minr=1e-6, rc_lo=2, order=4, nf=5.0)
    idx = np.array([5, 30, 50])
    ka, _, _ = O.k1_adaptive(r[idx], sN, p=p)
    kf = O.k1_fixed(r[idx], sN, p=p)
    assert np.all(np.abs(ka - kf) <= 1e-3 * np.abs(kf))


@pytest.mark.parametrize("tag,accuracy", [("e3", 1e-3), ("e6", 1e-6)])
def test_adaptive_k1_matches_reference_rapidity_derivative_lo(tag, accuracy):
    """The oracle's nested adaptive K1 against the reference's own
    rapidity_derivative_lo (S/bk/derivatives.py:43-110) on the 64-node MV
    workload (tests/golden/make_bk_rhs_golden.py): within 5 % of the
    reference's error estimate and 2e-8 relative, with the accurate J1
    (accj1) and with torch's J1 (torchj1, off by up to 4.7e-7 absolute).
    A dropped term or a wrong sign moves K1 by far more than the 1e-3 / 1e-6
    error estimates."""

    from pathlib import Path

    g = np.load(Path(__file__).with_name("golden") / "bk_rhs_reference_vectors.npz")
    r, N, P = g["r_grid"], g["N0"], g["parents"]
    np.testing.assert_allclose(r, O.r_grid(1e-6, 1e2, 63), rtol=4e-16)
    sN, _ = O.evolution_splines(r, N)
    value, error, _ = O.k1_adaptive(r[P], sN, 85, 85, accuracy, O.params(minr=1e-6))
    for j1 in ("accj1", "torchj1"):
        ref, ref_err = g[f"k1_{tag}_{j1}"], g[f"k1err_{tag}_{j1}"]
        diff = np.abs(value - ref)
        assert np.all(diff <= 0.05 * ref_err), (j1, diff / ref_err)
        assert np.all(diff <= 2e-8 * np.abs(ref)), (j1, diff / np.abs(ref))
        # both report an error estimate below the requested accuracy
        assert np.all(ref_err <= accuracy * np.abs(ref)) and np.all(error <= accuracy * np.abs(value))
