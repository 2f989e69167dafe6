"""The oracle against vectors written by the reference's own FP64 CPU code
(tests/golden/make_bk_golden.py).

Tolerances are relative to each expression's rounding scale: the reference's
formulas cancel (1 - N at small distances, the K2 far-region -2/Z^4 against
the log term, X ~ r), so a 1-ulp libm difference is amplified there.  The
oracle's magnitude functions give that scale (every additive term in
absolute value, 1/Delta-divided terms amplified by the relative size of Delta).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture(scope="module")
def splines(golden):
    return O.evolution_splines(golden["r_grid"], golden["mv_N"])


def test_r_grid(golden):
    r = O.r_grid(float(golden["cfg_minr"]), float(golden["cfg_maxr"]), int(golden["cfg_rpoints"]))
    np.testing.assert_allclose(r, golden["r_grid"], rtol=4e-16, atol=0)


def test_mv_initial_condition(golden):
    N = O.mv(golden["r_grid"], 0.2, 1.0, 1.0, 0.241)
    # 1 - exp(-e) for e ~ 1e-5 loses digits to cancellation in both codes
    np.testing.assert_allclose(N, golden["mv_N"], rtol=2e-12, atol=0)
    q, g, ec, lam = golden["mv2_params"]
    N2 = O.mv(golden["mv2_r"], q, g, ec, lam)
    np.testing.assert_allclose(N2, golden["mv2_N"], rtol=2e-12, atol=0)


def test_mv_uses_truncated_e():
    # the reference's e is 2.7182818, not math.e: at r = 1/Lambda the log term
    # is ln(1 + e_c * 2.7182818) exactly
    r = np.array([1 / 0.241])
    e = 0.2 * r[0] ** 2 / 4 * np.log(1 + 2.7182818)
    assert O.mv(r, 0.2, 1.0, 1.0, 0.241)[0] == pytest.approx(1 - np.exp(-e), rel=1e-15)
    e_true = 0.2 * r[0] ** 2 / 4 * np.log(1 + np.e)
    assert abs(O.mv(r, 0.2, 1.0, 1.0, 0.241)[0] - (1 - np.exp(-e_true))) > 1e-9


def test_evolution_splines(golden, splines):
    sN, sS = splines
    np.testing.assert_allclose(sN.x, golden["splN_x"], rtol=1e-15)
    for s, key in ((sN, "splN_coef"), (sS, "splS_coef")):
        ref = golden[key]
        m = ref.size // 4
        for k in range(4):  # per coefficient family: scale by its largest value
            part = slice(k * m, (k + 1) * m)
            scale = np.abs(ref[part]).max()
            np.testing.assert_allclose(s.coef[part], ref[part], rtol=0, atol=1e-12 * scale)


def test_kernel_lo_with_accurate_j1(golden):
    # reference Kernel_lo with torch.special.bessel_j1 replaced by scipy's j1
    args = (golden["klo_r"], golden["klo_z"], golden["klo_theta"])
    for cos_form in (True, False):
        p = O.params(minr=1e-6, geometry_cos=cos_form)
        got = O.kernel_lo(*args, p)
        mag = O.kernel_lo_magnitude(*args, p)
        ref = golden["klo_value_accj1"]
        assert np.all(np.abs(got - ref) <= 1e-13 * mag + 1e-300), cos_form


def test_kernel_lo_raw_reference_differs_only_by_torch_j1(golden):
    # torch.special.bessel_j1 deviates from scipy.special.j1 by up to 4.7e-7
    # (absolute) on [5, 8]; the reference's raw values differ from the accurate
    # ones by at most that amount times the double-log term's weight
    args = (golden["klo_r"], golden["klo_z"], golden["klo_theta"])
    p = O.params(minr=1e-6, geometry_cos=True)
    got = O.kernel_lo(*args, p)
    mag = O.kernel_lo_magnitude(*args, p)
    assert np.all(np.abs(got - golden["klo_value"]) <= 5e-6 * mag + 1e-300)


def test_lo_integrand(golden, splines):
    sN, _ = splines
    args = (golden["klo_r"], golden["klo_z"], golden["klo_theta"])
    for cos_form in (True, False):
        p = O.params(minr=1e-6, geometry_cos=cos_form)
        got = O.k1_integrand(*args, sN, p)
        mag = O.k1_integrand_magnitude(*args, sN, p)
        # S/N spline values enter through exp(spline(log x)): the golden
        # spline coefficients differ at 1e-12, so the scale allows 1e-11
        assert np.all(np.abs(got - golden["lo_integrand_accj1"]) <= 1e-11 * mag + 1e-300), cos_form


def test_kernel_nlo_and_fermion(golden):
    Rr, X, Y, X2, Y2, ZZ = golden["knlo_in"]
    a2, b2 = (X * Y2) ** 2, (X2 * Y) ** 2
    amp = (a2 + b2) / np.abs(a2 - b2)  # conditioning of the 1/Delta terms
    k = O.kernel_nlo(Rr, X, Y, X2, Y2, ZZ)
    kf = O.kernel_nlo_fermion(Rr, X, Y, X2, Y2, ZZ)
    for got, ref in ((k, golden["knlo_value"]), (kf, golden["knlof_value"])):
        scale = np.abs(ref) + 2 / ZZ**4
        assert np.all(np.abs(got - ref) <= 1e-14 * scale * amp)


def test_nlo_integrand_reference_formula(golden, splines):
    _, sS = splines
    for cos_form in (True, False):
        p = O.params(minr=1e-6, geometry_cos=cos_form)
        got, _, _ = O.nlo_integrand_log(*golden["nlo_in"], sS, p)
        mag = O.nlo_integrand_magnitude(*golden["nlo_in"], sS, p)
        assert np.all(np.abs(got - golden["nlo_value"]) <= 1e-14 * mag + 1e-300), cos_form


def test_adapt_edges_bit_exact(golden):
    for bins in (50, 32):
        got = O.vegas_adapt(golden[f"adapt{bins}_edges"], golden[f"adapt{bins}_hist"],
                            golden[f"adapt{bins}_count"], float(np.finfo(np.float64).eps))
        np.testing.assert_array_equal(got, golden[f"adapt{bins}_new"])
