"""Host logic of the Python surface (no GPU): configuration validation as the
reference's BKConfig does it, the r grid, initial conditions, DipoleTable
validation and the legacy table format."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from nlo_bk_optimized_amd import MV, BKConfig, DipoleTable, ICDataFile, load_bk_table, save_bk_table
from nlo_bk_optimized_amd.bk.config import (
    BKOrder,
    RunningCouplingLO,
    RunningCouplingNLO,
    require_gpu_supported,
)
from nlo_bk_optimized_amd.bk.evolution import _r_grid
from oracle import oracle as O


def test_config_defaults_match_reference():
    c = BKConfig()
    assert (c.RPOINTS, c.MINR, c.MAXR, c.MCINTPOINTS, c.DE_SOLVER_STEP, c.KSUB) == (100, 1e-6, 30.0, 100_000, 0.2, 0.65)
    assert c.RC_LO is RunningCouplingLO.BALITSKY_LO and c.Order is BKOrder.NLO_RESUM_DLOG_SLOG


@pytest.mark.parametrize(
    "kwargs",
    [dict(NF=4), dict(MINR=0), dict(RPOINTS=1), dict(DE_SOLVER_STEP=0), dict(C2=0),
     dict(RC_LO=RunningCouplingLO.FIXED_LO), dict(KINEMATICAL_CONSTRAINT=True),
     dict(VEGAS_REUSE_WARMUP_FRACTION=2.0), dict(MC_BINS=300), dict(PRECISION="fp16"),
     dict(RK_METHOD="rk45")],
)
def test_config_validation(kwargs):
    with pytest.raises(ValueError):
        BKConfig(**kwargs)


def test_unsupported_integration_is_rejected_loudly():
    from nlo_bk_optimized_amd.bk.config import IntegrationMethod

    for method in (IntegrationMethod.MISER, IntegrationMethod.MULTIPLE):
        with pytest.raises(NotImplementedError, match="INTMETHOD_NLO"):
            require_gpu_supported(BKConfig(INTMETHOD_NLO=method))
    # every coupling, order and NF choice runs on the device
    require_gpu_supported(BKConfig(RC_NLO=RunningCouplingNLO.PARENT_NLO, NF=5))


def test_auto_rk_method_follows_the_reference_selection():
    from nlo_bk_optimized_amd.bk.config import BKOrder, resolve_rk_method

    f32, f64 = torch.float32, torch.float64
    assert resolve_rk_method(BKConfig(), f32) == "kutta3"        # fixed RK2(3) path
    assert resolve_rk_method(BKConfig(), f64) == "adaptive_rk2"  # float64: adaptive
    assert resolve_rk_method(BKConfig(RPOINTS=255), f32) == "adaptive_rk2"
    assert resolve_rk_method(BKConfig(Order=BKOrder.LO), f32) == "adaptive_rk2"
    assert resolve_rk_method(BKConfig(EULER_METHOD=True), f32) == "euler"
    assert resolve_rk_method(BKConfig(RK_METHOD="rk4", RPOINTS=255), f32) == "rk4"


def test_auto_k1_rule_follows_the_reference_selection():
    """K1_RULE="auto": the fixed rule only where the reference validated it
    (float32, production physics, RPOINTS = 100 and the default K1 limits;
    S/bk/derivatives.py:145-166 with S/bk/evolution.py:160-177), else the
    adaptive rule; explicit choices win; the device lists hold 100 intervals."""

    from nlo_bk_optimized_amd.bk.config import BKOrder, resolve_k1_rule

    f32, f64 = torch.float32, torch.float64
    assert resolve_k1_rule(BKConfig(), f32) == "fixed"
    assert resolve_k1_rule(BKConfig(), f64) == "adaptive"
    assert resolve_k1_rule(BKConfig(RPOINTS=255), f32) == "adaptive"
    assert resolve_k1_rule(BKConfig(INTACCURACY=1e-4), f32) == "adaptive"
    assert resolve_k1_rule(BKConfig(Order=BKOrder.NLO), f32) == "adaptive"
    assert resolve_k1_rule(BKConfig(NF=5), f32) == "adaptive"
    assert resolve_k1_rule(BKConfig(K1_RULE="fixed", RPOINTS=255), f64) == "fixed"
    assert resolve_k1_rule(BKConfig(K1_RULE="adaptive"), f32) == "adaptive"
    with pytest.raises(NotImplementedError, match="100 intervals"):
        resolve_k1_rule(BKConfig(RINTPOINTS=150), f64)
    assert resolve_k1_rule(BKConfig(RINTPOINTS=150, K1_RULE="fixed"), f64) == "fixed"
    with pytest.raises(ValueError, match="K1_RULE"):
        BKConfig(K1_RULE="gauss")


def test_r_grid_matches_oracle():
    c = BKConfig(RPOINTS=255, MINR=1e-6, MAXR=1e2)
    r = _r_grid(c, device="cpu", dtype=torch.float64).numpy()
    assert r.size == 256
    np.testing.assert_allclose(r, O.r_grid(1e-6, 1e2, 255), rtol=2e-15)


def test_mv_matches_oracle():
    r = torch.logspace(-8, 2, 200, dtype=torch.float64)
    for params in ((0.2, 1.0, 1.0), (0.37, 1.13, 7.5)):
        q, g, ec = params
        got = MV(qs0sqr=q, gamma=g, ec=ec).dipole_amplitude(r).numpy()
        np.testing.assert_allclose(got, O.mv(r.numpy(), q, g, ec, 0.241), rtol=2e-12)


def test_mv_product_matches_reference_vectors(golden):
    """The package's MV (the amplitude solve_bk starts from) against the
    reference's own outputs, in the reference's float64 arithmetic."""

    r = torch.tensor(golden["r_grid"], dtype=torch.float64)
    np.testing.assert_allclose(MV(qs0sqr=0.2, gamma=1.0).dipole_amplitude(r).numpy(), golden["mv_N"],
                               rtol=1e-15, atol=0)
    q, g, ec, lam = (float(v) for v in golden["mv2_params"])
    r2 = torch.tensor(golden["mv2_r"], dtype=torch.float64)
    got = MV(qs0sqr=q, gamma=g, ec=ec, lambdaqcd=lam).dipole_amplitude(r2).numpy()
    np.testing.assert_allclose(got, golden["mv2_N"], rtol=1e-15, atol=0)
    # float32 stays float32, as the reference's tensors do
    assert MV().dipole_amplitude(r.float()).dtype == torch.float32


def test_ic_data_file_errors(tmp_path):
    bad = tmp_path / "bad.dat"
    bad.write_text("".join(f"{1e-3 * (k + 1)} {0.01 * k}\n" for k in range(12)) + "0.5\n")
    with pytest.raises(ValueError, match=r"bad.dat:13: expected the two columns"):
        ICDataFile(bad)
    bad.write_text("".join(f"{1e-3 * (k + 1)} {0.01 * k}\n" for k in range(12)) + "0.5 x\n")
    with pytest.raises(ValueError, match=r"bad.dat:13: r and N must be numbers"):
        ICDataFile(bad)
    bad.write_text("# r N extra\n" + "".join(f"{1e-3 * (k + 1)} {0.01 * k} 7\n" for k in range(12)))
    ic = ICDataFile(bad, dtype=torch.float32)
    assert ic.r.dtype == torch.float32 and ic.r.numel() == 12
    assert ic.min_r == pytest.approx(1e-3) and ic.max_r == pytest.approx(1.2e-2)


def test_ic_data_file(tmp_path):
    r = np.logspace(-5, 1, 40)
    N = 1 - np.exp(-r**2)
    path = tmp_path / "ic.dat"
    path.write_text("# r N\n" + "\n".join(f"{a:.17e} {b:.17e}" for a, b in zip(r, N)))
    ic = ICDataFile(path)
    out = ic.dipole_amplitude(torch.tensor([1e-6, r[7], 20.0], dtype=torch.float64))
    assert out[0] == 0.0 and out[2] == 1.0
    assert out[1].item() == pytest.approx(N[7], rel=1e-14)
    with pytest.raises(ValueError):
        (tmp_path / "short.dat").write_text("1 0.1\n2 0.2\n")
        ICDataFile(tmp_path / "short.dat")


def test_dipole_table_validation():
    r = torch.logspace(-3, 1, 5, dtype=torch.float64)
    y = torch.tensor([0.0, 0.5], dtype=torch.float64)
    DipoleTable(r=r, y=y, N=torch.zeros(2, 5, dtype=torch.float64), x0=0.01)
    with pytest.raises(ValueError):
        DipoleTable(r=r, y=y, N=torch.zeros(3, 5, dtype=torch.float64), x0=0.01)
    with pytest.raises(ValueError):
        DipoleTable(r=r.flip(0), y=y, N=torch.zeros(2, 5, dtype=torch.float64), x0=0.01)
    with pytest.raises(ValueError):
        DipoleTable(r=r, y=y, N=torch.full((2, 5), math.nan, dtype=torch.float64), x0=0.01)


def test_legacy_table_round_trip(tmp_path):
    r = 1e-5 * 1.16 ** torch.arange(12, dtype=torch.float64)
    y = torch.tensor([0.0, 0.4, 0.8], dtype=torch.float64)
    N = torch.rand(3, 12, dtype=torch.float64, generator=torch.Generator().manual_seed(0))
    save_bk_table(tmp_path / "t.dat", DipoleTable(r=r, y=y, N=N, x0=0.01))
    t = load_bk_table(tmp_path / "t.dat")
    torch.testing.assert_close(t.r, r, rtol=1e-14, atol=0)
    torch.testing.assert_close(t.N, N, rtol=1e-15, atol=0)
    torch.testing.assert_close(t.y, y)
    assert t.x0 == 0.01


def test_load_reference_nlobk_table():
    # R/xb_001_mc50w.dat, the reference's C++ nlobk NLO table (101 r x 26 Y)
    from pathlib import Path

    path = Path(__file__).resolve().parent / "golden" / "xb_001_mc50w.dat"
    t = load_bk_table(path)
    assert t.N.shape == (26, 101)
    assert t.x0 == 0.01
    assert t.r[0].item() == pytest.approx(1e-5) and t.r[-1].item() == pytest.approx(30.0, rel=1e-12)
    np.testing.assert_allclose(t.y.numpy(), np.arange(26) * 0.4, atol=1e-12)
    assert bool(((t.N >= 0) & (t.N <= 1)).all())
    # evolution increases N at fixed r
    assert bool((t.N[-1] >= t.N[0]).all())
