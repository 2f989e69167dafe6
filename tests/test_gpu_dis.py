"""DIS on the GPU against the FP64 oracle (oracle/dis_oracle.py) and the
reference's own numbers.

FP64 device values must match the oracle per point to the conditioning-aware
tolerance of the oracle tests (the reference formula's own sensitivity to
Bessel rounding); FP32 device values to the same sensitivity scaled to FP32
rounding.  VEGAS sums are compared with the oracle evaluated at the same
Sobol points; the LO cross sections with the reference's deterministic
Gauss-Legendre values.
"""

from __future__ import annotations

import math
import zlib

import numpy as np
import pytest
import torch

from oracle import dis_oracle as D
from oracle import oracle as O
from tests.test_dis_oracle import _cancellation_tolerance

pytestmark = pytest.mark.gpu

GOLDEN = "tests/golden/dis_reference_vectors.npz"
LIGHT = D.QuarkSpec(0.002, math.sqrt(2 / 3))
CHARM = D.QuarkSpec(1.4, 2 / 3)
GBW_ORACLE = D.GBWParams(1.0, 0.3, 1.0, 1.0)


@pytest.fixture(scope="module")
def g():
    return np.load(GOLDEN)


@pytest.fixture(scope="module")
def dev():
    from nlo_bk_optimized_amd import device

    return device


def _configs():
    from nlo_bk_optimized_amd.dis.config import (
        DISConfig,
        NcScheme,
        RunningCouplingIRScheme,
        RunningCouplingScheme,
    )

    return {
        "default": (DISConfig(), D.DisSetup()),
        "variant": (DISConfig(maxr=5.0, nc_scheme=NcScheme.LargeNC, rc_scheme=RunningCouplingScheme.PARENT,
                              rc_ir_scheme=RunningCouplingIRScheme.SMOOTH, C2_alpha=2.0, nf_alphas=1),
                    D.DisSetup(maxr=5.0, finite_nc=False, parent_coupling=True, smooth_coupling=True, c2_alpha=2.0,
                               nf_alphas=1)),
    }


def _quark(spec):
    from nlo_bk_optimized_amd.dis.config import Quark, QuarkType

    return Quark(QuarkType.LIGHT) if spec is LIGHT else Quark(QuarkType.C, 1.4)


def cuda(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _close(value, ref, tol, what, floor=1e-13):
    """Per-point check with a floor of `floor` x the case's mean |ref|."""

    value, ref = np.asarray(value), np.asarray(ref)
    m = np.isfinite(ref)
    assert np.array_equal(np.isfinite(value), m), what
    tol = np.maximum(tol, floor * np.mean(np.abs(ref[m])) if m.any() else 0)
    bad = np.abs(value[m] - ref[m]) > tol[m]
    assert not bad.any(), (what, np.abs(value[m] - ref[m])[bad][:5], tol[m][bad][:5])


def _fp32_input_sensitivity(f, x):
    """Sum over the coordinates of |f(x with coordinate d moved by one FP32 ulp)
    - f(x)|: what rounding the sample point to FP32 alone changes (a coordinate
    next to 1 whose complement sets an exponential's argument can carry a
    relative condition number of thousands)."""

    base = np.nan_to_num(f(x), nan=0.0, posinf=0.0, neginf=0.0)
    total = np.zeros_like(base)
    for d in range(x.shape[1]):
        xp = x.copy()
        xp[:, d] = x[:, d] + np.spacing(x[:, d].astype(np.float32)).astype(np.float64)
        total += np.abs(np.nan_to_num(f(xp), nan=0.0, posinf=0.0, neginf=0.0) - base)
    return total


def _fp32_tolerance(ref, tol, f, x):
    """FP32 path: 1e-4 relative, the FP64 Bessel sensitivity scaled to FP32
    rounding, and twice the sensitivity to FP32 rounding of the sample point."""

    return 1e-4 * np.abs(ref) + (tol - 1e-13 * np.abs(ref)) * (2e-7 / 2e-15) + 2.0 * _fp32_input_sensitivity(f, x)


# ------------------------------------------------------------ special functions


def test_special_functions_fp64_and_fp32(dev, g):
    x = g["bessel_x"]
    for fn, ref in (("K0", D.k0(x)), ("K1", D.k1(x)), ("K2", D.k2(x))):
        got = dev.dis_special(fn, cuda(x)).cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=4e-15, atol=0, err_msg=fn)
        got32 = dev.dis_special(fn, cuda(x), precision="fp32").cpu().numpy()
        fin = ref < 1e37
        np.testing.assert_allclose(got32[fin], ref[fin], rtol=2e-6, atol=0, err_msg=fn + " fp32")
    xl = g["li2_x"]
    ref = D.li2(xl)
    np.testing.assert_allclose(dev.dis_special("Li2", cuda(xl)).cpu().numpy(), ref, rtol=4e-15, atol=1e-300)
    got32 = dev.dis_special("Li2", cuda(xl), precision="fp32").cpu().numpy()
    np.testing.assert_allclose(got32, ref, rtol=2e-6, atol=1e-7)
    # special points: K(0) = inf, Li2(1) = pi^2 / 6, Li2(0) = 0
    z = dev.dis_special("K0", cuda([0.0])).cpu().numpy()
    assert np.isinf(z).all()
    assert dev.dis_special("Li2", cuda([1.0, 0.0])).cpu().numpy().tolist() == [math.pi**2 / 6, 0.0]


# ------------------------------------------------------------------ dipoles


def _bk_dipole(initialize_at=None):
    from nlo_bk_optimized_amd.dipole import BKDipole, load_bk_table

    d = BKDipole(load_bk_table("tests/golden/xb_001_mc50w.dat"))
    if initialize_at is not None:
        d.initialize_interpolation(initialize_at)
    return d


def test_dipole_amplitudes_match_oracle(dev, g):
    from nlo_bk_optimized_amd.dipole import GBW

    gb = GBW(Qs0sqr=0.24, lambda_=0.29, gamma=1.1, x0=0.01)
    got = gb.dipole_amplitude(cuda(g["gbw_r"]), cuda(g["gbw_Y"])).cpu().numpy()
    # the device forms 1 - e^-a as -expm1(-a); the reference's 1 - exp(-a) is
    # only absolutely accurate (~1e-16) for small a
    np.testing.assert_allclose(got, D.gbw(g["gbw_r"], g["gbw_Y"], D.GBWParams(0.24, 0.29, 1.1, 0.01)),
                               rtol=1e-14, atol=3e-16)
    bk = _bk_dipole()
    got = bk.dipole_amplitude(cuda(g["bk_r"]), cuda(g["bk_Y"])).cpu().numpy()
    np.testing.assert_allclose(got, g["bk_N_local"], rtol=1e-13, atol=1e-16)
    bk.initialize_interpolation(float(g["bk_init_Y"]))
    assert bk.maxr_interpolate == pytest.approx(float(g["bk_maxr_interpolate"]), rel=1e-14)
    got = bk.dipole_amplitude(cuda(g["bk_r"]), cuda(np.full_like(g["bk_r"], float(g["bk_init_Y"])))).cpu().numpy()
    np.testing.assert_allclose(got, g["bk_N_init"], rtol=1e-13, atol=1e-16)
    with pytest.raises(ValueError):
        bk.dipole_amplitude(cuda([1.0]), cuda([bk.max_Y + 1]))
    from nlo_bk_optimized_amd.dipole import AmplitudeInterpolation, BKDipole, load_bk_table

    lin = BKDipole(load_bk_table("tests/golden/xb_001_mc50w.dat"), AmplitudeInterpolation.LINEAR_LINEAR)
    got = lin.dipole_amplitude(cuda(g["bk_r"]), cuda(g["bk_Y"])).cpu().numpy()
    np.testing.assert_allclose(got, g["bk_N_bilinear"], rtol=1e-13, atol=1e-16)
    lin.initialize_interpolation(float(g["bk_init_Y"]))
    assert lin.maxr_interpolate == pytest.approx(float(g["bk_bilinear_maxr_interpolate"]), rel=1e-14)
    got = lin.dipole_amplitude(cuda(g["bk_r"]), cuda(np.full_like(g["bk_r"], float(g["bk_init_Y"])))).cpu().numpy()
    np.testing.assert_allclose(got, g["bk_N_bilinear_init"], rtol=1e-13, atol=1e-16)


def test_bk_local_spline_table_equals_per_query_solve(dev):
    """The precomputed local-spline table (every row and interval) against
    solving the 7- / 9-point natural spline of the Y-interpolated rows per
    query, at random (r, Y) over and beyond the table (clamped)."""

    bk = _bk_dipole()
    assert bk.device_local is not None
    rng = np.random.default_rng(8)
    r = np.exp(rng.uniform(np.log(bk.min_r) - 1, np.log(bk.max_r) + 1, 20000))
    Y = rng.uniform(0.0, bk.max_Y, 20000)
    r = np.concatenate([r, np.asarray(bk._r_host)])
    Y = np.concatenate([Y, np.full(bk._r_host.numel(), bk.max_Y * 0.5)])
    table = bk.dipole_amplitude(cuda(r), cuda(Y)).cpu().numpy()
    local = bk.device_local
    bk.device_local = None
    try:
        solve = bk.dipole_amplitude(cuda(r), cuda(Y)).cpu().numpy()
    finally:
        bk.device_local = local
    np.testing.assert_allclose(table, solve, rtol=1e-13, atol=1e-16)


# --------------------------------------------------------------- integrands

POINT_CASES = [("LO", "T"), ("LO", "L"), ("T0", "T"), ("T1", "T"), ("T2", "T"), ("Omega_L_const", "L"),
               ("ab", "L"), ("cd", "L"), ("I1", "T"), ("I1", "L"), ("I2", "T"), ("I2", "L"), ("I3", "T"),
               ("I3", "L"), ("I2_nested", "T"), ("I2_nested", "L"), ("I3_nested", "T"), ("I3_nested", "L")]


def _oracle_values(kind, pol, q, setup, x, amp, Q2=9.0, xbj=1e-3, r_range=(1e-30, 1e3), lo_quarks=(LIGHT, CHARM)):
    transverse = pol == "T"
    if kind == "LO":
        return D.lo_integrand(x, Q2, xbj, transverse, lo_quarks, amp, 1.0, setup)
    if kind in ("T0", "T1", "T2", "Omega_L_const", "ab", "cd"):
        return D.dipole_integrand(x, Q2, xbj, transverse, q, kind, amp, 1.0, setup, *r_range)
    if kind.endswith("_nested"):
        return D.qg_integrand(x, Q2, xbj, transverse, q, kind[:2], amp, setup, nested_points=48)
    return D.qg_integrand(x, Q2, xbj, transverse, q, kind, amp, setup)


@pytest.mark.parametrize("kind,pol", POINT_CASES)
def test_integrands_match_oracle_gbw(dev, monkeypatch, kind, pol):
    from nlo_bk_optimized_amd import _lib
    from nlo_bk_optimized_amd.dipole import GBW
    from nlo_bk_optimized_amd.dis.config import Quark, QuarkType

    dim = _lib.DIS_DIMENSIONS[kind]
    rng = np.random.default_rng(zlib.crc32(f"{kind}{pol}".encode()))
    n = 64 if kind.endswith("_nested") else 256
    x = rng.uniform(0, 1, (n, dim))
    amp = lambda r, Y: D.gbw(r, Y, GBW_ORACLE)  # noqa: E731
    gbw = GBW(Qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0)
    for q in (LIGHT, CHARM):
        for name, (cfg, setup) in _configs().items():
            quarks = (Quark(QuarkType.LIGHT), Quark(QuarkType.C, 1.4)) if kind == "LO" else (_quark(q),)
            desc = dev.dis_desc(9.0, 1e-3, pol == "T", quarks, gbw, cfg)
            ref, tol = _cancellation_tolerance(lambda: _oracle_values(kind, pol, q, setup, x, amp), monkeypatch)
            got = dev.dis_integrand(desc, kind, cuda(x)).cpu().numpy()
            _close(got, ref, tol, (kind, pol, q, name, "fp64"))
            got32 = dev.dis_integrand(desc, kind, cuda(x), precision="fp32").cpu().numpy()
            tol32 = _fp32_tolerance(ref, tol, lambda xx: _oracle_values(kind, pol, q, setup, xx, amp), x)
            _close(got32, ref, tol32, (kind, pol, q, name, "fp32"), floor=1e-5)


# the masses of the light and charm quarks of the 32 x 16 reduced cross-section grid
LIGHT_GRID = D.QuarkSpec(0.14, math.sqrt(2 / 3))
CHARM_GRID = D.QuarkSpec(1.27, 2 / 3)


@pytest.mark.parametrize("Q2,xbj", [(100.0, 1e-5), (1.0, 1e-2), (100.0, 1e-2), (1.0, 1e-5)])
@pytest.mark.parametrize("kind,pol", POINT_CASES)
def test_integrands_match_oracle_at_grid_corners(dev, monkeypatch, kind, pol, Q2, xbj):
    """The corners of the 32 x 16 (Q^2 in [1, 100], x in [1e-5, 1e-2]) grid with
    its masses m = 0.14 and 1.27, FP64 and FP32, GBW dipole, default setup."""

    from nlo_bk_optimized_amd import _lib
    from nlo_bk_optimized_amd.dipole import GBW
    from nlo_bk_optimized_amd.dis.config import Quark, QuarkType

    dim = _lib.DIS_DIMENSIONS[kind]
    rng = np.random.default_rng(zlib.crc32(f"{kind}{pol}{Q2}{xbj}".encode()))
    x = rng.uniform(0, 1, (32 if kind.endswith("_nested") else 128, dim))
    amp = lambda r, Y: D.gbw(r, Y, GBW_ORACLE)  # noqa: E731
    gbw = GBW(Qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0)
    cfg, setup = _configs()["default"]
    light, charm = Quark(QuarkType.LIGHT, 0.14), Quark(QuarkType.C, 1.27)
    for q, quark in ((LIGHT_GRID, light), (CHARM_GRID, charm)):
        quarks = (light, charm) if kind == "LO" else (quark,)
        desc = dev.dis_desc(Q2, xbj, pol == "T", quarks, gbw, cfg)
        def f(xx):
            return _oracle_values(kind, pol, q, setup, xx, amp, Q2=Q2, xbj=xbj, lo_quarks=(LIGHT_GRID, CHARM_GRID))

        ref, tol = _cancellation_tolerance(lambda: f(x), monkeypatch)
        assert np.isfinite(ref).any()
        got = dev.dis_integrand(desc, kind, cuda(x)).cpu().numpy()
        _close(got, ref, tol, (kind, pol, q, Q2, xbj, "fp64"))
        got32 = dev.dis_integrand(desc, kind, cuda(x), precision="fp32").cpu().numpy()
        _close(got32, ref, _fp32_tolerance(ref, tol, f, x), (kind, pol, q, Q2, xbj, "fp32"), floor=1e-5)


@pytest.mark.parametrize("kind,pol", [("T1", "T"), ("cd", "L"), ("I1", "T"), ("I2_nested", "L"),
                                      ("I3", "T")])
def test_integrands_match_oracle_bk_table(dev, monkeypatch, kind, pol):
    from nlo_bk_optimized_amd import _lib
    from nlo_bk_optimized_amd.dipole import load_bk_table

    t = load_bk_table("tests/golden/xb_001_mc50w.dat")
    bko = D.BKDipoleOracle(t.r.numpy(), t.y.numpy(), t.N.numpy(), t.x0)
    bk = _bk_dipole()
    dim = _lib.DIS_DIMENSIONS[kind]
    x = np.random.default_rng(5).uniform(0, 1, (48, dim))
    cfg, setup = _configs()["default"]
    for q in (LIGHT, CHARM):
        desc = dev.dis_desc(9.0, 1e-3, pol == "T", (_quark(q),), bk, cfg)
        # the table spans r in [1e-5, 30] and Y in [ln 100, ln 100 + 25]
        ref, tol = _cancellation_tolerance(
            lambda: _oracle_values(kind, pol, q, setup, x, bko.amplitude, Q2=9.0, xbj=1e-3,
                                   r_range=(float(t.r[0]), float(t.r[-1]))), monkeypatch)
        got = dev.dis_integrand(desc, kind, cuda(x)).cpu().numpy()
        _close(got, ref, tol * 10 + 1e-12 * np.abs(ref), (kind, pol, q, "bk"))


# ------------------------------------------------------------ integrations


def test_vegas_pass_matches_oracle_at_the_same_points(dev):
    from nlo_bk_optimized_amd import _lib
    from nlo_bk_optimized_amd.dipole import GBW
    from nlo_bk_optimized_amd.dis.config import Quark, QuarkType

    gbw = GBW(Qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0)
    amp = lambda r, Y: D.gbw(r, Y, GBW_ORACLE)  # noqa: E731
    cfg, setup = _configs()["default"]
    for kind, pol, n, first in (("T1", "T", 4096, 0), ("I1", "L", 4096, 64 * 7), ("I2_nested", "T", 1024, 0)):
        dim = _lib.DIS_DIMENSIONS[kind]
        rng = np.random.default_rng(3)
        e = np.sort(rng.uniform(0, 1, (dim, 31)), axis=1)
        edges = np.concatenate([np.zeros((dim, 1)), e, np.ones((dim, 1))], axis=1)
        desc = dev.dis_desc(9.0, 1e-3, pol == "T", (Quark(QuarkType.C, 1.4),), gbw, cfg)
        table = dev.sobol_table(dim, 11)
        sums, hist, count = dev.dis_vegas(desc, kind, cuda(edges), table, n, first_index=first)
        ot = O.sobol_table(dim, 11)
        q = O.sobol_u32(ot, first, n)
        if first == 0:
            q[0] = ot.first_q
        x, jac, bins = O.vegas_map(q, edges, np.zeros(dim), np.ones(dim))
        w = _oracle_values(kind, pol, CHARM, setup, x, amp) * jac
        s = sums.cpu().numpy()
        assert s[0] == pytest.approx(w.sum(), rel=1e-11, abs=1e-12 * np.abs(w).sum())
        assert s[1] == pytest.approx((w * w).sum(), rel=1e-11)
        ref_count = np.stack([np.bincount(bins[:, d], minlength=32) for d in range(dim)])
        np.testing.assert_array_equal(count.cpu().numpy(), ref_count)
        ref_hist = np.stack([np.bincount(bins[:, d], weights=np.abs(w), minlength=32) for d in range(dim)])
        # fixed-point block sums: each |w| rounded to 2^-51 of its block's largest
        np.testing.assert_allclose(hist.cpu().numpy(), ref_hist, rtol=1e-10, atol=1e-12 * np.abs(w).max())


def test_lo_cross_sections_match_reference(g):
    from nlo_bk_optimized_amd.dipole import GBW
    from nlo_bk_optimized_amd.dis import observables as obs
    from nlo_bk_optimized_amd.dis.config import DISConfig, Polarization

    gbw = GBW(Qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0)
    for Q2, xbj, tflag, value, error in g["lo_sigma"]:
        pol = Polarization.T if tflag else Polarization.L
        res = obs.photon_proton_cross_section_LO_d2b(torch.tensor(Q2, dtype=torch.float64),
                                                     torch.tensor(xbj, dtype=torch.float64), pol, gbw, DISConfig())
        assert float(res.value) == pytest.approx(value, rel=1e-12)
        assert float(res.error) == pytest.approx(error, rel=1e-8, abs=1e-15)
        res32 = obs.photon_proton_cross_section_LO_d2b(torch.tensor(Q2, dtype=torch.float32),
                                                       torch.tensor(xbj, dtype=torch.float32), pol, gbw, DISConfig())
        assert float(res32.value) == pytest.approx(value, rel=1e-5)


def test_structure_functions_at_nlo_are_finite_and_consistent():
    from nlo_bk_optimized_amd.dipole import GBW
    from nlo_bk_optimized_amd.dis import F2, FL, FT, DISConfig, DISOrder, sigma_r

    gbw = GBW(Qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0)
    cfg = DISConfig(order=DISOrder.NLO, maxeval=20_000, epsrel=1e-2)
    Q2, xbj = torch.tensor(9.0), torch.tensor(1e-3)
    ft = FT(Q2, xbj, gbw, cfg, seed=3)
    fl = FL(Q2, xbj, gbw, cfg, seed=3 + 1000)
    f2 = F2(Q2, xbj, gbw, cfg, seed=3)
    for r in (ft, fl, f2):
        assert math.isfinite(float(r.value)) and math.isfinite(float(r.error)) and float(r.error) > 0
    assert float(f2.value) == pytest.approx(float(ft.value) + float(fl.value), rel=1e-12)
    sr = sigma_r(Q2, xbj, 318.0, gbw, cfg, seed=3)
    y = 9.0 / (1e-3 * 318.0**2)
    assert float(sr.value) == pytest.approx(float(f2.value) - y * y / (1 + (1 - y) ** 2) * float(fl.value),
                                            rel=1e-10)
