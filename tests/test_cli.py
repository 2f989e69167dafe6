"""Host logic of the nlo-bk and nlo-dis command lines (S/cli/nlobk.py, S/cli/nlodis.py).

The device work is replaced by recording fakes so that the flag -> config
mapping, the run modes and the output format are checked on the CPU.
"""

from __future__ import annotations

import io
from contextlib import redirect_stdout
from types import SimpleNamespace

import pytest
import torch

from nlo_bk_optimized_amd.bk.config import BKOrder, IntegrationMethod, RunningCouplingLO, RunningCouplingNLO
from nlo_bk_optimized_amd.cli import nlobk, nlodis
from nlo_bk_optimized_amd.dis.config import DISOrder, RunningCouplingIRScheme, RunningCouplingScheme


def test_nlobk_defaults_follow_reference_command():
    cfg = nlobk.config_from_args(nlobk.build_parser().parse_args([]))
    assert cfg.Order is BKOrder.NLO_RESUM_DLOG_SLOG
    assert cfg.RC_LO is RunningCouplingLO.BALITSKY_LO
    assert cfg.RC_NLO is RunningCouplingNLO.SMALLEST_NLO
    assert cfg.INTMETHOD_NLO is IntegrationMethod.VEGAS
    assert (cfg.NF, cfg.MINR, cfg.MAXR, cfg.RPOINTS) == (3, 1e-6, 30.0, 100)
    assert (cfg.MCINTPOINTS, cfg.DE_SOLVER_STEP, cfg.KSUB, cfg.C2) == (100_000, 0.2, 0.65, 1.0)
    assert cfg.FORCE_POSITIVE_N and not cfg.DNDY and not cfg.EULER_METHOD
    assert cfg.PRECISION == "fp32"


def test_nlobk_flags_map_to_config():
    args = nlobk.build_parser().parse_args(
        ["--order", "lo", "--rc-lo", "fixed", "--fixed-alpha-s", "0.3", "--nolimit", "--rpoints", "64",
         "--mcintpoints", "4096", "--step", "0.1", "--precision", "fp64_stable", "--rk-method", "rk4",
         "--mc-bins", "32", "--nf", "5", "--Ksub", "0.5", "--C2", "2.0"])
    cfg = nlobk.config_from_args(args)
    assert cfg.Order is BKOrder.LO
    assert cfg.RC_LO is RunningCouplingLO.FIXED_LO
    # a fixed LO coupling fixes the NLO one as well (S/cli/nlobk.py:83-86)
    assert cfg.RC_NLO is RunningCouplingNLO.FIXED_NLO
    assert cfg.FIXED_AS == 0.3 and not cfg.FORCE_POSITIVE_N
    assert (cfg.RPOINTS, cfg.MCINTPOINTS, cfg.DE_SOLVER_STEP) == (64, 4096, 0.1)
    assert (cfg.PRECISION, cfg.RK_METHOD, cfg.MC_BINS) == ("fp64_stable", "rk4", 32)
    assert (cfg.NF, cfg.KSUB, cfg.C2) == (5, 0.5, 2.0)
    with pytest.raises(SystemExit):
        nlobk.build_parser().parse_args(["--nf", "4"])


class _Recorder:
    def __init__(self):
        self.calls = []

    def make(self, name):
        def fn(Q2, xbj, dipole, config, **opts):
            self.calls.append((name, float(Q2), float(xbj), config, opts, Q2.dtype))
            v = 1.0 if name == "FL" else 2.0
            v += len(config.quarks) * 0.25
            return SimpleNamespace(value=torch.tensor(v), error=torch.tensor(1e-3), seed=opts["seed"],
                                   n_eval=10, converged=True)
        return fn


@pytest.fixture
def fake_dis(monkeypatch):
    rec = _Recorder()
    made = {}

    def fake_dipole(table, interpolation, device=None):
        made["interpolation"] = interpolation
        return SimpleNamespace(table=table)

    monkeypatch.setattr(nlodis, "load_bk_table", lambda path, dtype=None: SimpleNamespace(path=path))
    monkeypatch.setattr(nlodis, "BKDipole", fake_dipole)
    for name in ("F2", "FL", "FT"):
        monkeypatch.setattr(nlodis, name, rec.make(name))
    return rec, made


BASE = ["--datafile", "bk.dat", "--C2", "1.5", "--charm-mass", "1.27", "--proton-area", "8.4",
        "--rc-scheme", "PARENT", "--device", "cpu"]


def test_nlodis_hera_fl_mode(fake_dis, capsys):
    rec, made = fake_dis
    assert nlodis.main(BASE) == 0
    out = capsys.readouterr()
    lines = out.out.strip().splitlines()
    assert lines[0] == "fit,x,Q2,FL"
    assert len(lines) == 1 + len(nlodis.HERA_FL_POINTS) == 21
    assert made["interpolation"] is nlodis.AmplitudeInterpolation.LINEAR_LINEAR
    assert [c[0] for c in rec.calls] == ["FL"] * 20
    name, Q2, xbj, cfg, opts, dtype = rec.calls[0]
    assert Q2 == 1.5 and xbj == pytest.approx(0.000028, rel=1e-7) and dtype is torch.float32
    assert cfg.order is DISOrder.NLO and cfg.rc_scheme is RunningCouplingScheme.PARENT
    assert cfg.rc_ir_scheme is RunningCouplingIRScheme.SMOOTH
    assert cfg.transverse_area == pytest.approx(8.4 * 2.56819)
    assert (cfg.C2_alpha, cfg.maxeval, cfg.epsrel, cfg.cuda_nested_points) == (1.5, 1_000_000, 0.01, 48)
    assert [q.mass for q in cfg.quarks] == [0.01, 1.27]
    assert opts == {"quadrature_points": 24, "batch_size": 65_536, "seed": 0}
    assert lines[1].startswith("nlodis,2.8e-05,1.5,")
    assert out.err.count("# FL") == 20


def test_nlodis_f2fl_grid_mode(fake_dis, capsys):
    rec, _ = fake_dis
    assert nlodis.main(BASE + ["--runmode", "F2FL_GRID", "--dtype", "float64", "--no-header", "--name", "fit1"]) == 0
    lines = capsys.readouterr().out.strip().splitlines()
    xs = []
    x = 1e-5
    while x <= 1e-2:
        xs.append(x)
        x *= 1.5
    assert len(lines) == 3 * len(xs)
    row = lines[0].split(",")
    assert row[0] == "fit1" and float(row[1]) == 1e-5 and float(row[2]) == 4.5 and len(row) == 8
    # one quark per FT/FL call, both quarks and LO order for F2_LO
    kinds = [(c[0], len(c[3].quarks), c[3].order) for c in rec.calls[:5]]
    assert kinds == [("FT", 1, DISOrder.NLO), ("FL", 1, DISOrder.NLO), ("FT", 1, DISOrder.NLO),
                     ("FL", 1, DISOrder.NLO), ("F2", 2, DISOrder.LO)]
    assert rec.calls[0][5] is torch.float64
    # F2 light = FT + FL of the light-quark calls
    assert float(row[3]) == pytest.approx(2.25 + 1.25) and float(row[4]) == pytest.approx(1.25)


def _ic_file(tmp_path):
    path = tmp_path / "ic.dat"
    path.write_text("".join(f"{1e-6 * 10 ** (k / 4):.10e} {min(1.0, 1e-10 * 10 ** (k / 2)):.10e}\n" for k in range(34)))
    return path


def test_nlobk_ic_data_keeps_the_requested_dtype(tmp_path, monkeypatch):
    """--ic-data with default flags: the file is read in float32 (the
    reference passes dtype=dtype to ICDataFile, S/cli/nlobk.py), so
    solve_bk resolves the production Kutta step and the fixed K1 rule."""

    from nlo_bk_optimized_amd.bk.config import resolve_k1_rule, resolve_rk_method

    seen = {}

    def fake_solve(ic, maxy, config, *, device, dtype, seed):
        effective = ic.r.dtype  # solve_bk takes the dtype of an ICDataFile
        seen.update(ic_dtype=effective, dtype=dtype, rk=resolve_rk_method(config, effective),
                    k1=resolve_k1_rule(config, effective))
        return SimpleNamespace(y=torch.zeros(1), r=ic.r)

    monkeypatch.setattr(nlobk, "solve_bk", fake_solve)
    monkeypatch.setattr(nlobk, "save_bk_table", lambda path, table: None)
    with redirect_stdout(io.StringIO()):
        assert nlobk.main(["--ic-data", str(_ic_file(tmp_path)), "--device", "cpu"]) == 0
    assert seen == {"ic_dtype": torch.float32, "dtype": torch.float32, "rk": "kutta3", "k1": "fixed"}
    with redirect_stdout(io.StringIO()):
        assert nlobk.main(["--ic-data", str(_ic_file(tmp_path)), "--device", "cpu", "--dtype", "float64"]) == 0
    assert seen["ic_dtype"] is torch.float64 and seen["rk"] == "adaptive_rk2" and seen["k1"] == "adaptive"


def test_nlobk_dndy_output(monkeypatch):
    def fake_dndy(ic, config, *, device, dtype, seed):
        assert config.DNDY and dtype is torch.float32 and seed == 4
        v = torch.tensor([1e-3, 2e-3], dtype=dtype)
        return SimpleNamespace(r=v, K1=v * 2, K2_Kf=-v, N=v / 10)

    monkeypatch.setattr(nlobk, "compute_dndy", fake_dndy)
    buf = io.StringIO()
    with redirect_stdout(buf):
        assert nlobk.main(["--dndy", "--seed", "4", "--device", "cpu"]) == 0
    lines = buf.getvalue().strip().splitlines()
    assert lines[0] == "# r dN/dy[K1] dN/dy[K2+Kf] N" and len(lines) == 3
    vals = [float(x) for x in lines[2].split()]
    assert vals == pytest.approx([2e-3, 4e-3, -2e-3, 2e-4], rel=1e-6)
