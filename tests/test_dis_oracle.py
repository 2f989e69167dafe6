"""The FP64 DIS oracle against the reference's own vectors and closed forms.

Golden vectors come from tests/golden/make_dis_golden.py (the reference's CPU
code in FP64).  Where the reference formula cancels (the light-quark mass
difference terms), agreement is limited by that formula's conditioning, not by
either implementation: those cases carry a tolerance computed from the size of
the cancelling terms (see _cancellation_tolerance).
"""

from __future__ import annotations

import math

import mpmath
import numpy as np
import pytest

from oracle import dis_oracle as D

GOLDEN = "tests/golden/dis_reference_vectors.npz"


@pytest.fixture(scope="module")
def g():
    return np.load(GOLDEN)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert np.array_equal(np.isfinite(a), np.isfinite(b))
    m = np.isfinite(b)
    return np.abs(a[m] - b[m]) / np.maximum(np.abs(b[m]), 1e-300)


LIGHT = D.QuarkSpec(0.002, math.sqrt(2 / 3))
CHARM = D.QuarkSpec(1.4, 2 / 3)
GBW_DIS = D.GBWParams(1.0, 0.3, 1.0, 1.0)
SETUPS = {
    "default": D.DisSetup(),
    "variant": D.DisSetup(maxr=5.0, finite_nc=False, parent_coupling=True, smooth_coupling=True,
                          c2_alpha=2.0, nf_alphas=1),
}


def amp(r, Y):
    return D.gbw(r, Y, GBW_DIS)


def test_li2_matches_mpmath_and_reference(g):
    x = g["li2_x"]
    ref = np.array([float(mpmath.polylog(2, mpmath.mpf(v))) for v in x])
    assert _rel(D.li2(x), ref).max() < 2e-15
    assert _rel(D.li2(x), g["li2"]).max() < 2e-15
    # Euler reflection and the duplication formula
    t = np.linspace(0.01, 0.99, 41)
    np.testing.assert_allclose(D.li2(t) + D.li2(1 - t), math.pi**2 / 6 - np.log(t) * np.log(1 - t), rtol=1e-14)
    np.testing.assert_allclose(D.li2(t) + D.li2(-t), 0.5 * D.li2(t * t), rtol=1e-13, atol=1e-16)


def test_bessel_k_matches_reference(g):
    assert _rel(D.k0(g["bessel_x"]), g["bessel_k0"]).max() < 5e-15
    assert _rel(D.k1(g["bessel_x"]), g["bessel_k1"]).max() < 5e-15
    assert _rel(D.k2(g["bessel_x"]), g["bessel_k2"]).max() < 5e-15


def test_couplings_and_gbw_match_reference(g):
    r = g["alpha_r"]
    assert _rel(D.alpha_freeze(r, 4), g["alpha_freeze_nf4"]).max() < 1e-15
    assert _rel(D.alpha_freeze(r, 1, 2.5, 0.5), g["alpha_freeze_nf1_c2"]).max() < 1e-15
    assert _rel(D.alpha_smooth(r, 4), g["alpha_smooth_nf4"]).max() < 1e-15
    p = D.GBWParams(0.24, 0.29, 1.1, 0.01)
    assert _rel(D.gbw(g["gbw_r"], g["gbw_Y"], p), g["gbw_N"]).max() < 1e-15
    # closed form: small-r limit r^2 Qs^2 / 4
    np.testing.assert_allclose(D.gbw(np.array([1e-5]), np.array([0.0]), GBW_DIS), [1e-10 / 4], rtol=1e-15)


def test_bk_dipole_matches_reference(g):
    from nlo_bk_optimized_amd.dipole.table import load_bk_table

    t = load_bk_table("tests/golden/xb_001_mc50w.dat")
    bk = D.BKDipoleOracle(t.r.numpy(), t.y.numpy(), t.N.numpy(), t.x0)
    assert _rel(bk.amplitude(g["bk_r"], g["bk_Y"]), g["bk_N_local"]).max() < 1e-14
    fn = bk.initialized(float(g["bk_init_Y"]))
    assert fn.maxr == pytest.approx(float(g["bk_maxr_interpolate"]), rel=1e-15)
    assert _rel(fn(g["bk_r"]), g["bk_N_init"]).max() < 1e-14
    lin = D.BKDipoleOracle(t.r.numpy(), t.y.numpy(), t.N.numpy(), t.x0, bilinear=True)
    assert _rel(lin.amplitude(g["bk_r"], g["bk_Y"]), g["bk_N_bilinear"]).max() < 1e-14
    fl = lin.initialized(float(g["bk_init_Y"]))
    assert fl.maxr == pytest.approx(float(g["bk_bilinear_maxr_interpolate"]), rel=1e-15)
    assert _rel(fl(g["bk_r"]), g["bk_N_bilinear_init"]).max() < 1e-14


def test_lo_integrand_matches_reference(g):
    for pol in ("T", "L"):
        r, z = g["lo_r"], g["lo_z"]
        v = sum(D.lo_wavefunction(r, z, 9.0, q, pol == "T") for q in (LIGHT, CHARM))
        v = v * amp(r, np.full_like(r, math.log(1 / 1e-3)))
        assert _rel(v, g[f"lo_{pol}"]).max() < 1e-14


def _cancellation_tolerance(fn, monkeypatch, rel=1e-13, bessel_agreement=2e-15):
    """Tolerance per sample: rel |value| plus the value's sensitivity to a
    relative change of the Bessel K values, times the agreement of two Bessel
    implementations.  The sensitivity is measured by scaling K0/K1 by 1 +- 1e-10,
    the sign of call c in trial t being bit t of c: any two calls get opposite
    signs in some trial, so a difference of two Bessel terms always shows."""

    base = fn()
    k0, k1 = D.k0, D.k1
    eps = 1e-10
    spread = np.zeros_like(base)
    for trial in range(6):
        calls = [0]

        def jit(f, trial=trial, calls=calls):
            def g(x):
                calls[0] += 1
                sign = 1.0 if (calls[0] >> trial) & 1 else -1.0
                return f(x) * (1 + sign * eps)
            return g

        monkeypatch.setattr(D, "k0", jit(k0))
        monkeypatch.setattr(D, "k1", jit(k1))
        spread = np.maximum(spread, np.abs(fn() - base) / eps)
    monkeypatch.setattr(D, "k0", k0)
    monkeypatch.setattr(D, "k1", k1)
    return base, rel * np.abs(base) + 10 * bessel_agreement * spread


def _check(value, tol, ref):
    """Per-sample tolerance, with a floor of 1e-13 of the mean |value| of the
    case: samples far below the integrand's scale carry rounding of the
    arguments (e.g. sqrt(kappa^2 + delta) with delta << kappa^2) that the
    Bessel jitter does not see and that cannot move an integral."""

    assert np.array_equal(np.isfinite(value), np.isfinite(ref))
    m = np.isfinite(ref)
    tol = np.maximum(tol, 1e-13 * np.mean(np.abs(ref[m])))
    bad = np.abs(value[m] - ref[m]) > tol[m]
    assert not bad.any(), (np.abs(value[m] - ref[m])[bad], tol[m][bad])


@pytest.mark.parametrize("dims,pol,c", [(2, "T", "T0"), (2, "L", "Omega_L_const"), (3, "T", "T1"),
                                        (3, "L", "ab"), (4, "T", "T2"), (4, "L", "cd")])
def test_dipole_integrands_match_reference(g, monkeypatch, dims, pol, c):
    x = g[f"dip_x{dims}"]
    for qn, q in (("light", LIGHT), ("charm", CHARM)):
        for cn, s in SETUPS.items():
            v, tol = _cancellation_tolerance(
                lambda: D.dipole_integrand(x, 9.0, 1e-3, pol == "T", q, c, amp, 1.0, s), monkeypatch)
            _check(v, tol, g[f"dip_{c}_{qn}_{cn}"])


@pytest.mark.parametrize("dims,c", [(5, "I1"), (6, "I2"), (7, "I3")])
def test_qg_integrands_match_reference(g, monkeypatch, dims, c):
    x = g[f"qg_x{dims}"]
    for pol in ("T", "L"):
        for qn, q in (("light", LIGHT), ("charm", CHARM)):
            for cn, s in SETUPS.items():
                v, tol = _cancellation_tolerance(
                    lambda: D.qg_integrand(x, 9.0, 1e-3, pol == "T", q, c, amp, s), monkeypatch)
                _check(v, tol, g[f"qg_{c}_{pol}_{qn}_{cn}"])


def test_nested_qg_equals_reference_quadrature(g, monkeypatch):
    x5 = g["nest_x5"]
    s = SETUPS["default"]
    for pol in ("T", "L"):
        for qn, q in (("light", LIGHT), ("charm", CHARM)):
            _, z1, z2, _, _, x01sq, x02sq, x21sq, _, _ = D.qg_geometry(x5, 9.0, 1e-3, s)
            for which in (0, 1):
                v, tol = _cancellation_tolerance(
                    lambda: D.qg_nested(9.0, q.mass, z1, z2, x01sq, x02sq, x21sq, pol == "T", 48)[which],
                    monkeypatch)
                _check(v, tol, g[f"nest_I{which + 2}_{pol}_{qn}"])


def test_lo_cross_section_matches_reference(g):
    factor = 4 * D.ALPHA_EM * D.NC / (2 * math.pi) ** 2
    for Q2, xbj, tflag, value, error in g["lo_sigma"]:
        def fn(x, Q2=Q2, xbj=xbj, tflag=tflag):
            return D.lo_integrand(x, Q2, xbj, bool(tflag), (LIGHT, CHARM), amp, 1.0, SETUPS["default"])

        fine, coarse = D.gl_tensor(fn, 2, 24)
        assert fine * factor == pytest.approx(value, rel=1e-13)
        assert abs(fine - coarse) * factor == pytest.approx(error, rel=1e-9, abs=1e-15)


def test_mapped_rule_integrates_polynomials_exactly():
    y, w = D.mapped_gl_rule(48)
    for k in range(0, 20):
        assert np.sum(w * y**k) == pytest.approx(1 / (k + 1), rel=1e-13)
