"""The oracle for the non-production physics choices, the kinematical
constraint and the adaptive RK2 controller, against vectors from the
reference's own FP64 code (tests/golden/make_bk_golden.py: ``var_*``,
``alpha_*``, ``kin_*``, ``ark2_*``)."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O
from oracle.rk_oracle import adaptive_rk2_to

VARIANTS = ("parent_dls", "smallest_lod", "fac_nlod", "beuf_dls", "fixed_dls", "nf5_dls", "nf0_lo",
            "bal_nlo", "bal_lodls_nf5")


def variant_params(golden, k, **extra):
    lo, nlo, resum, order, nf, fixed_as, c2, ksub = golden["var_table"][k]
    return O.params(minr=1e-6, rc_lo=int(lo), rc_nlo=int(nlo), resum_rc=int(resum), order=int(order),
                    nf=float(nf), fixed_as=float(fixed_as), c2=float(c2), ksub=float(ksub), **extra)


@pytest.fixture(scope="module")
def splines(golden):
    return O.evolution_splines(golden["r_grid"], golden["mv_N"])


def test_coupling_flavour_thresholds_and_nf0(golden):
    r = golden["alpha_r"]
    np.testing.assert_allclose(O.alpha_s(r, O.params(nf=5.0)), golden["alpha_nf5"], rtol=2e-15, atol=0)
    np.testing.assert_allclose(O.alpha_s(r, O.params(nf=5.0, c2=2.0)), golden["alpha_nf5_c2"], rtol=2e-15,
                               atol=0)
    np.testing.assert_allclose(O.alpha_s(r, O.params(nf=0.0, c2=2.0)), golden["alpha_nf0_c2"], rtol=2e-15,
                               atol=0)
    # the cap: alpha = 1 where the one-loop value exceeds it
    assert np.any(golden["alpha_nf5"] == 1.0) and np.all(golden["alpha_nf5"] <= 1.0)
    # fixed coupling: FIXED_AS everywhere
    fixed = O.alpha_s(r, O.params(rc_lo=1, rc_nlo=1, fixed_as=0.17))
    assert np.all(fixed == 0.17)


@pytest.mark.parametrize("k,name", list(enumerate(VARIANTS)))
def test_kernel_lo_and_lo_integrand_variants(golden, splines, k, name):
    rr, zz, th = golden["var_klo_in"]
    p = variant_params(golden, k, geometry_cos=True)
    got = O.kernel_lo(rr, zz, th, p)
    mag = O.kernel_lo_magnitude(rr, zz, th, p)
    ref = golden[f"var_{name}_klo"]
    assert np.all(np.abs(got - ref) <= 1e-13 * mag + 1e-300), np.max(np.abs(got - ref) / (mag + 1e-300))
    sN, _ = splines
    got = O.k1_integrand(rr, zz, th, sN, p)
    mag = O.k1_integrand_magnitude(rr, zz, th, sN, p)
    ref = golden[f"var_{name}_lo_integrand"]
    assert np.all(np.abs(got - ref) <= 1e-11 * mag + 1e-300)
    assert np.any(ref != 0)


def test_orders_without_resummation_are_the_bare_kernel(golden):
    # Order LO / NLO: Kernel_lo returns the running-coupling kernel alone
    # (S/bk_kernels/k1.py:100-101), so the values equal the resummed order's
    # kernel with alpha -> 0 in the resummation; check one point by hand.
    rr, zz, th = golden["var_klo_in"]
    k = VARIANTS.index("bal_nlo")
    ref = golden["var_bal_nlo_klo"]
    p = variant_params(golden, k, geometry_cos=True)
    X = np.sqrt(rr**2 + zz**2 - 2 * rr * zz * np.cos(th))
    a = O.alpha_s(np.stack((rr, X, zz)).ravel(), p).reshape(3, -1)
    bare = 3 / (2 * math.pi**2) * a[0] * (rr**2 / (X**2 * zz**2) + (a[2] / a[1] - 1) / zz**2 +
                                          (a[1] / a[2] - 1) / X**2)
    scale = 3 / (2 * math.pi**2) * a[0] * (rr**2 / (X**2 * zz**2) + (a[2] / a[1] + 1) / zz**2 +
                                           (a[1] / a[2] + 1) / X**2)
    assert np.all(np.abs(ref - bare) <= 1e-12 * scale)


@pytest.mark.parametrize("k,name", list(enumerate(VARIANTS)))
def test_nlo_integrand_variants(golden, splines, k, name):
    _, sS = splines
    p = variant_params(golden, k, geometry_cos=True)
    got, _, _ = O.nlo_integrand_log(*golden["var_nlo_in"], sS, p)
    mag = O.nlo_integrand_magnitude(*golden["var_nlo_in"], sS, p)
    ref = golden[f"var_{name}_nlo"]
    assert np.all(np.abs(got - ref) <= 1e-14 * mag + 1e-300)


def test_stable_integrand_follows_the_coupling_choice(golden):
    # The cancellation-free integrand (N spline) against the reference
    # formula's values (S spline) for every choice, on the points where the
    # two spline choices agree for the production coupling (they differ where
    # N ~ 1, by the absolute precision of 1 - N).  A wrong coupling or Nf
    # shifts the values by percents; the spline choice by ~1e-5.
    sN, sS = O.evolution_splines(golden["r_grid"], golden["mv_N"])
    inp = golden["var_nlo_in"]
    prod, _, _ = O.nlo_integrand_stable(*inp, sN, O.params(minr=1e-6, geometry_cos=True))
    ref_prod = golden["var_bal_nlo_nlo"]  # production NLO coupling, Nf = 3
    mask = (ref_prod != 0) & (np.abs(prod - ref_prod) <= 1e-3 * np.abs(ref_prod))
    assert mask.sum() > 100
    for k, name in enumerate(VARIANTS):
        p = variant_params(golden, k, geometry_cos=True)
        got, _, _ = O.nlo_integrand_stable(*inp, sN, p)
        ref = golden[f"var_{name}_nlo"]
        # and where the reference formula keeps its digits (r << z, z' loses them)
        keep = mask & (O.nlo_integrand_magnitude(*inp, sS, p) <= 1e4 * np.abs(ref))
        assert keep.sum() > 40, name
        assert np.all(np.abs(got - ref)[keep] <= 1e-3 * np.abs(ref)[keep]), name
        if p.rc_nlo != 0 or p.nf != 3.0 or p.rc_lo == 1:
            # the production coupling would not pass
            assert np.median(np.abs(prod - ref)[keep] / np.abs(ref)[keep]) > 1e-2, name


def test_history_lookup(golden):
    r, hy, hN = golden["r_grid"], golden["kin_hist_y"], golden["kin_hist_N"]
    sN, _ = O.evolution_splines(r, hN[-1])
    h = O.History(r, hy, hN, float(hy[-1]))
    got = O.history_N(golden["kin_q_r"], golden["kin_q_y"], sN, h)
    np.testing.assert_allclose(got, golden["kin_q_N"], rtol=1e-12, atol=1e-15)
    # rows reproduce themselves on the grid (a natural spline interpolates)
    inner = r[3:-3]
    for k in range(hy.size - 1):
        np.testing.assert_allclose(O.history_N(inner, np.full(inner.size, hy[k]), sN, h),
                                   np.clip(hN[k][3:-3], 0, 1), rtol=1e-13, atol=1e-16)
    assert math.isnan(O.history_N([0.1], [hy[-1] + 0.1], sN, h)[0])


def test_kinematical_lo_integrand(golden):
    r, hy, hN = golden["r_grid"], golden["kin_hist_y"], golden["kin_hist_N"]
    sN, _ = O.evolution_splines(r, hN[-1])
    h = O.History(r, hy, hN, float(hy[-1]))
    rr, zz, th = golden["kin_lo_in"]
    p = O.params(minr=1e-6, geometry_cos=True)
    got = O.k1_integrand_history(rr, zz, th, sN, h, p)
    scale = 2 * O.kernel_lo_magnitude(rr, zz, th, p)
    ref = golden["kin_lo_value"]
    assert np.all(np.abs(got - ref) <= 1e-11 * scale + 1e-300)
    assert np.count_nonzero(ref == 0) > 10 and np.count_nonzero(ref) > 100


def test_kinematical_constraint_at_zero_rapidity_width(golden):
    # with the current rapidity at the first row every shifted rapidity
    # below it is cut, and delta = 0 (min(X, Y) <= r) reads the current
    # amplitude: the integrand is K_lo (1 - N_r - S_X S_Y) there
    r, hN = golden["r_grid"], golden["kin_hist_N"]
    sN, _ = O.evolution_splines(r, hN[0])
    h = O.History(r, np.array([0.0]), hN[:1], 0.0)
    rr = np.array([0.5, 0.5, 1e-3])
    zz = np.array([0.1, 2.0, 5e-4])
    th = np.array([1.0, 1.0, 0.3])
    p = O.params(minr=1e-6)
    got = O.k1_integrand_history(rr, zz, th, sN, h, p)
    X = np.sqrt((rr - zz) ** 2 + 4 * rr * zz * np.sin(th / 2) ** 2)
    nX = np.minimum(O.loglog_eval(sN, X, 0.0, 1.0), 1)
    nY = np.minimum(O.loglog_eval(sN, zz, 0.0, 1.0), 1)
    nr = np.minimum(O.loglog_eval(sN, rr, 0.0, 1.0), 1)
    keep = np.minimum(X, zz) <= rr
    want = np.where(keep, O.kernel_lo(rr, zz, th, p) * (1 - nr - (1 - nX) * (1 - nY)), 0.0)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)
    assert keep[0] and not keep[1]


@pytest.mark.parametrize("case", [0, 1, 2])
def test_adaptive_rk2_controller(golden, case):
    step, h0, abserr, relerr = golden[f"ark2_case{case}"]
    rate = golden["ark2_rate"]

    def f(N, y):
        return rate * N * (1 - N) * (1 + 0.5 * np.sin(3 * y))

    N, y, h = golden["ark2_N0"].copy(), 0.0, h0
    stats: dict = {}
    for row in golden[f"ark2_traj{case}"]:
        N, y, h = adaptive_rk2_to(f, N, y, y + step, h, abserr, relerr, float(np.finfo(np.float64).eps), stats)
        assert y == pytest.approx(row[0], rel=1e-15)
        assert h == pytest.approx(row[1], rel=1e-9)
        np.testing.assert_allclose(N, row[2:], rtol=1e-12, atol=0)
    assert stats["accepted"] >= 4
    if case == 1:  # h0 = 2 against tight tolerances: the first attempts are rejected
        assert stats.get("rejected", 0) > 0


def test_adaptive_rk2_underflow_raises():
    with pytest.raises(RuntimeError, match="underflow"):
        adaptive_rk2_to(lambda N, y: np.full_like(N, np.nan), np.ones(3), 0.0, 1.0, 0.5, 1e-6, 1e-4, 1e-7)
