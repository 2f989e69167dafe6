"""The DIS composition on the CPU with the device integrators replaced by
fakes that return known integrals: normalisations, quark charges and
budgets, seeds, the LO(x0) + dipole + qqbar-g sum, F2 = FT + FL, the
VEGAS iteration statistics of ``vegas_device``, and the qqbar-g term's choice
between the fixed-Sobol estimator and VEGAS (S/dis/observables.py:459-532).

Known answers follow the reference's own tests
(T/tests/test_dis_production.py:70-107 for the LO factor
4 alpha_em Nc / (2 pi)^2, :195-231 for the qqbar-g factor
e_q^2 4 Nc alpha_em / (2 pi)^3 2 pi) and the closed forms of
S/numerics/integration.py:500-589 worked by hand.
"""

from __future__ import annotations

import math
from types import SimpleNamespace

import pytest
import torch

from nlo_bk_optimized_amd.dipole.amplitude import GBW
from nlo_bk_optimized_amd.dis import observables as obs
from nlo_bk_optimized_amd.dis.config import NC, AlphaEM, DISConfig, DISOrder, Polarization, Quark, QuarkType
from nlo_bk_optimized_amd.numerics.integration import IntegralResult


def _result(value, error=0.0, n_eval=1, converged=True, seed=None):
    return IntegralResult(torch.tensor(value, dtype=torch.float64), torch.tensor(error, dtype=torch.float64),
                          n_eval, converged, seed)


@pytest.fixture
def fakes(monkeypatch):
    """Replace the device batch (and Sobol tables) by fakes that record every
    integral the composition asks for and answer from tables: the real plan
    builds the items (budgets, iterations, seeds, descriptors)."""

    from nlo_bk_optimized_amd import _lib
    from nlo_bk_optimized_amd import device as dev

    calls = []
    gl_values = {}
    vegas_values = {}
    kinds = {v: k for k, v in _lib.DIS_KINDS.items()}

    def fake_table(dim, seed):
        t = _lib.SobolTable()
        t.dim = dim
        t.shift[0] = seed  # marks the seed for the fake batch
        return t

    def fake_batch(items, tables, *, bins, precision, epsrel, epsabs, device, workspace_budget):
        rows = []
        for it in items:
            kind = kinds[it.contribution]
            desc = SimpleNamespace(Q2=it.desc.Q2, xbj=it.desc.xbj, transverse=it.desc.transverse,
                                   n_quarks=it.desc.n_quarks)
            if it.method == _lib.NLO_DIS_METHOD_GAUSS_LEGENDRE:
                calls.append(("gl", kind, desc, dict(points=it.n, precision=precision)))
                value, error = gl_values.get(kind, (1.0, 0.0))
                rows.append((value, error, it.n**2 + max(2, it.n // 2) ** 2, 1.0))
            else:
                seed = int(tables[it.sobol_table].shift[0])
                calls.append(("vegas", kind, desc, dict(samples=it.n, iterations=it.max_iterations,
                                                         min_iterations=it.min_iterations, seed=seed,
                                                         precision=precision, first=it.first_index)))
                value, error = vegas_values.get(kind, (1.0, 0.0))
                rows.append((value, error, it.n * it.max_iterations, 1.0))
        return torch.tensor(rows, dtype=torch.float64)

    monkeypatch.setattr(dev, "sobol_table", fake_table)
    monkeypatch.setattr(dev, "dis_batch", fake_batch)
    return SimpleNamespace(calls=calls, gl=gl_values, vegas=vegas_values)


LO_FACTOR = 4 * AlphaEM * NC / (2 * math.pi) ** 2
DIP_FACTOR = 4 * NC * AlphaEM / (2 * math.pi) ** 2 * 2 * math.pi
QG_FACTOR = 4 * NC * AlphaEM / (2 * math.pi) ** 3 * 2 * math.pi


def test_lo_normalisation_known_answer(fakes):
    fakes.gl["LO"] = (5.0, 0.25)
    res = obs.photon_proton_cross_section_LO_d2b(torch.tensor(9.0), torch.tensor(1e-3), Polarization.T, GBW(),
                                                 DISConfig(epsrel=0.1))
    torch.testing.assert_close(res.value, torch.tensor(5.0 * LO_FACTOR, dtype=torch.float64))
    torch.testing.assert_close(res.error, torch.tensor(0.25 * LO_FACTOR, dtype=torch.float64))
    (kind, name, desc, opts), = fakes.calls
    assert (kind, name, opts["points"], opts["precision"]) == ("gl", "LO", 24, "fp32")
    assert desc.n_quarks == 2  # LO sums every quark inside the integrand


def test_qg_normalisation_known_answer(fakes):
    charm = Quark(QuarkType.C, 1.4)
    fakes.vegas.update({"I1": (2.0, 0.1), "I2_nested": (-0.5, 0.2), "I3_nested": (0.25, 0.3)})
    res = obs.sigma_qg_d2b(torch.tensor(9.0), torch.tensor(1e-3), Polarization.T, GBW(),
                           DISConfig(quarks=(charm,), maxeval=1000), seed=11)
    factor = charm.charge**2 * QG_FACTOR
    torch.testing.assert_close(res.value, torch.tensor(1.75 * factor, dtype=torch.float64))
    torch.testing.assert_close(res.error, torch.tensor(math.sqrt(0.01 + 0.04 + 0.09) * factor, dtype=torch.float64))
    assert [c[1] for c in fakes.calls] == ["I1", "I2_nested", "I3_nested"]
    assert [c[3]["seed"] for c in fakes.calls] == [11, 12, 13]
    # _vegas_with_budget (S/dis/observables.py:661-686): 8 iterations of maxeval / 8, at least 3
    assert all((c[3]["samples"], c[3]["iterations"], c[3]["min_iterations"]) == (125, 8, 3) for c in fakes.calls)


def test_qg_float64_integrates_the_unnested_forms(fakes):
    obs.sigma_qg_d2b(torch.tensor(9.0, dtype=torch.float64), torch.tensor(1e-3, dtype=torch.float64), Polarization.L,
                     GBW(), DISConfig(quarks=(Quark(QuarkType.C, 1.4),)))
    assert [c[1] for c in fakes.calls] == ["I1", "I2", "I3"]
    assert all(c[3]["precision"] == "fp64" for c in fakes.calls)


def test_dipole_normalisation_budgets_and_seeds(fakes):
    light, charm = Quark(QuarkType.LIGHT, 0.14), Quark(QuarkType.C, 1.27)
    fakes.gl.update({"Omega_L_const": (3.0, 0.0)})
    fakes.vegas.update({"ab": (1.0, 0.3), "cd": (-2.0, 0.4)})
    res = obs.sigma_dip_d2b(torch.tensor(4.0), torch.tensor(1e-2), Polarization.L, GBW(),
                            DISConfig(quarks=(light, charm), maxeval=800), seed=5)
    per_quark = 3.0 + 1.0 - 2.0
    want = (light.charge**2 + charm.charge**2) * per_quark * DIP_FACTOR
    torch.testing.assert_close(res.value, torch.tensor(want, dtype=torch.float64))
    err = math.sqrt((light.charge**4 + charm.charge**4) * (0.09 + 0.16)) * DIP_FACTOR
    torch.testing.assert_close(res.error, torch.tensor(err, dtype=torch.float64))
    vegas = [c for c in fakes.calls if c[0] == "vegas"]
    assert [c[1] for c in vegas] == ["ab", "cd", "ab", "cd"]
    assert [c[3]["seed"] for c in vegas] == [5, 6, 7, 8]
    # light quarks (m < 0.4) get ten times the budget (S/dis/observables.py:421)
    assert [c[3]["samples"] * c[3]["iterations"] for c in vegas] == [8000, 8000, 800, 800]


def test_cross_section_and_structure_functions_compose(fakes):
    fakes.gl.update({"LO": (2.0, 0.0), "T0": (1.0, 0.0)})
    fakes.vegas.update({"T1": (0.5, 0.0), "T2": (0.25, 0.0), "I1": (0.125, 0.0), "I2_nested": (0.0625, 0.0),
                        "I3_nested": (0.03125, 0.0), "ab": (1.0, 0.0), "cd": (1.0, 0.0)})
    fakes.gl["Omega_L_const"] = (1.0, 0.0)
    charm = Quark(QuarkType.C, 1.27)
    cfg = DISConfig(quarks=(charm,), maxeval=800, transverse_area=2.5, order=DISOrder.NLO)
    dip = GBW()
    Q2, x = torch.tensor(10.0), torch.tensor(1e-3)
    sig = obs.photon_proton_cross_section_d2b(Q2, x, Polarization.T, dip, cfg, seed=3)
    e2 = charm.charge**2
    want = 2.0 * LO_FACTOR + e2 * (1.0 + 0.5 + 0.25) * DIP_FACTOR + e2 * (0.125 + 0.0625 + 0.03125) * QG_FACTOR
    torch.testing.assert_close(sig.value, torch.tensor(want, dtype=torch.float64))
    lo_call = next(c for c in fakes.calls if c[1] == "LO")
    assert lo_call[2].xbj == pytest.approx(dip.X0())  # the LO term is taken at x0 at NLO
    assert [c[3]["seed"] for c in fakes.calls if c[1] in ("T1", "T2")] == [3, 4]
    assert [c[3]["seed"] for c in fakes.calls if c[1] in ("I1", "I2_nested", "I3_nested")] == [103, 104, 105]
    fakes.calls.clear()
    ft = obs.FT(Q2, x, dip, cfg, seed=3)
    torch.testing.assert_close(ft.value, torch.tensor(want * 10.0 / (4 * math.pi**2 * AlphaEM) * 2.5,
                                                      dtype=torch.float64))
    fakes.calls.clear()
    f2 = obs.F2(Q2, x, dip, cfg, seed=3)
    fl = obs.FL(Q2, x, dip, cfg, seed=1003)
    torch.testing.assert_close(f2.value, ft.value + fl.value)
    l_seeds = [c[3]["seed"] for c in fakes.calls if c[1] in ("ab", "cd")]
    assert l_seeds[:2] == [1003, 1004]  # F2 seeds its L integrals 1000 later
    lo_only = obs.photon_proton_cross_section_d2b(Q2, x, Polarization.T, dip,
                                                  DISConfig(quarks=(charm,), order=DISOrder.LO))
    torch.testing.assert_close(lo_only.value, torch.tensor(2.0 * LO_FACTOR, dtype=torch.float64))


def test_batch_composition_equals_per_point(fakes):
    """F2_batch / sigma_r_batch over several points: the per-point results,
    and the same integrals (kinds, seeds, budgets) as the per-point calls."""

    light, charm = Quark(QuarkType.LIGHT, 0.14), Quark(QuarkType.C, 1.27)
    cfg = DISConfig(quarks=(light, charm), maxeval=1600, order=DISOrder.NLO)
    fakes.vegas.update({"T1": (0.5, 0.1), "T2": (0.25, 0.2), "I1": (0.125, 0.05), "ab": (1.5, 0.3)})
    pts = [(torch.tensor(q), torch.tensor(x)) for q, x in ((2.0, 1e-3), (20.0, 1e-4), (50.0, 1e-2))]
    batch = obs.F2_batch(pts, GBW(), cfg, seed=7)
    batch_calls = list(fakes.calls)
    fakes.calls.clear()
    single = [obs.F2(q, x, GBW(), cfg, seed=7) for q, x in pts]
    for b, s1 in zip(batch, single):
        assert torch.equal(b.value, s1.value) and torch.equal(b.error, s1.error)
        assert b.n_eval == s1.n_eval and b.converged == s1.converged and b.seed == 7
    key = [(c[0], c[1], c[2].Q2, c[2].xbj, c[2].transverse, tuple(sorted(c[3].items()))) for c in batch_calls]
    assert key == [(c[0], c[1], c[2].Q2, c[2].xbj, c[2].transverse, tuple(sorted(c[3].items())))
                   for c in fakes.calls]
    # sigma_r = FT + (1 - y^2 / Y+) FL with F2's seeds
    sr = obs.sigma_r_batch([pts[0], pts[2]], 318.0, GBW(), cfg, seed=7)
    ft = obs.FT(*pts[0], GBW(), cfg, seed=7)
    fl = obs.FL(*pts[0], GBW(), cfg, seed=1007)
    y = 2.0 / (1e-3 * 318.0**2)
    torch.testing.assert_close(sr[0].value, ft.value + (1 - y * y / (1 + (1 - y) ** 2)) * fl.value)
    # without a seed F2 takes the default 0 for both polarisations, as the reference does
    fakes.calls.clear()
    obs.F2(*pts[0], GBW(), cfg)
    seeds = {c[3]["seed"] for c in fakes.calls if c[0] == "vegas"}
    assert min(seeds) == 0 and max(seeds) < 1000


def _fake_device(monkeypatch, sums_per_iteration):
    from nlo_bk_optimized_amd import device as dev

    seen = {"first": [], "adapt": 0}
    it = iter(sums_per_iteration)

    def fake_vegas(desc, kind, edges, table, n, *, first_index, precision):
        seen["first"].append(first_index)
        s = torch.tensor(next(it), dtype=torch.float64)
        return s, torch.ones(edges.shape[0], edges.shape[1] - 1), torch.ones(edges.shape[0], edges.shape[1] - 1)

    def fake_adapt(edges, hist, count, eps):
        seen["adapt"] += 1
        return edges

    monkeypatch.setattr(dev, "dis_vegas", fake_vegas)
    monkeypatch.setattr(dev, "vegas_adapt", fake_adapt)
    monkeypatch.setattr(dev, "sobol_table", lambda dim, seed: None)
    return seen


def test_vegas_iteration_statistics_known_answers(monkeypatch):
    from nlo_bk_optimized_amd.numerics.integration import vegas_device

    # n = 4: (sum w, sum w^2) = (8, 20) -> estimate 2, variance (20 - 16) / 3 / 4 = 1/3;
    #        (4, 6) -> estimate 1, variance (6 - 4) / 3 / 4 = 1/6.
    # Inverse-variance combination: (2 * 3 + 1 * 6) / (3 + 6) = 4/3, error sqrt(1 / 9) = 1/3.
    seen = _fake_device(monkeypatch, [(8.0, 20.0), (4.0, 6.0)])
    res = vegas_device(None, "T0", samples_per_iteration=4, max_iterations=2, min_iterations=2, bins=4,
                       epsrel=1.0, seed=9, precision="fp64", first_index=128, device="cpu")
    assert float(res.value) == pytest.approx(4 / 3, rel=1e-15)
    assert float(res.error) == pytest.approx(1 / 3, rel=1e-15)
    assert res.converged and res.n_eval == 8 and res.seed == 9
    assert seen["first"] == [128, 192] and seen["adapt"] == 2  # blocks of 64 Sobol indices, adapt every pass


def test_vegas_waits_for_min_iterations_then_stops(monkeypatch):
    from nlo_bk_optimized_amd.numerics.integration import vegas_device

    # identical passes: estimate 2, variance 1/3 each; after k passes error sqrt(1 / (3 k))
    seen = _fake_device(monkeypatch, [(8.0, 20.0)] * 8)
    res = vegas_device(None, "T0", samples_per_iteration=4, max_iterations=8, min_iterations=3, bins=4,
                       epsrel=0.5, seed=0, precision="fp64", device="cpu")
    # error <= 0.5 * 2 already after one pass, but three passes are required
    assert len(seen["first"]) == 3 and res.converged
    assert float(res.value) == pytest.approx(2.0, rel=1e-15)
    assert float(res.error) == pytest.approx(math.sqrt(1 / 9), rel=1e-15)
    # never converging: all eight passes, converged False, error sqrt(1 / 24)
    seen = _fake_device(monkeypatch, [(8.0, 20.0)] * 8)
    res = vegas_device(None, "T0", samples_per_iteration=4, max_iterations=8, min_iterations=3, bins=4,
                       epsrel=1e-6, seed=0, precision="fp64", device="cpu")
    assert len(seen["first"]) == 8 and not res.converged and res.n_eval == 32
    assert float(res.error) == pytest.approx(math.sqrt(1 / 24), rel=1e-15)


def test_vegas_zero_variance_is_floored_at_tiny(monkeypatch):
    from nlo_bk_optimized_amd.numerics.integration import vegas_device

    # a constant integrand: sum w^2 = (sum w)^2 / n -> variance floored at the dtype's tiny
    _fake_device(monkeypatch, [(12.0, 36.0)] * 3)
    res = vegas_device(None, "T0", samples_per_iteration=4, max_iterations=3, min_iterations=3, bins=4,
                       epsrel=1e-3, seed=0, precision="fp32", device="cpu")
    assert float(res.value) == 3.0 and res.converged
    assert float(res.error) == pytest.approx(math.sqrt(torch.finfo(torch.float32).tiny / 3), rel=1e-12)


def _fixed(total_value, total_error, converged, seed):
    from nlo_bk_optimized_amd.dis.fixed_sobol import FIXED_SOBOL_EVALUATIONS, FixedSobolResult

    t = _result(total_value, total_error, FIXED_SOBOL_EVALUATIONS, converged, seed)
    return FixedSobolResult((t, t, t), t)


def test_fixed_sobol_support_is_limited_to_the_validated_workload():
    """T/tests/test_dis_production.py:110-116 and S/dis/observables.py:640-658."""

    from nlo_bk_optimized_amd.dis.fixed_sobol import fixed_sobol_supported

    f32, f64 = torch.tensor(9.0), torch.tensor(9.0, dtype=torch.float64)
    assert fixed_sobol_supported(f32, GBW(), DISConfig())
    assert not fixed_sobol_supported(f32, GBW(), DISConfig(maxeval=100_000))
    assert not fixed_sobol_supported(f64, GBW(), DISConfig())
    assert not fixed_sobol_supported(f32, GBW(), DISConfig(cuda_nested_points=32))
    assert not fixed_sobol_supported(f32, GBW(), DISConfig(cuda_nested=False))
    assert not fixed_sobol_supported(f32, GBW(), DISConfig(cuda_fusion=False))
    assert not fixed_sobol_supported(f32, object(), DISConfig())


def test_qg_dispatches_to_fixed_sobol(fakes, monkeypatch):
    """T/tests/test_dis_production.py:195-231: the fixed-Sobol totals, scaled by
    e_q^2 4 Nc alpha_em / (2 pi)^3 2 pi, with no VEGAS integral."""

    from nlo_bk_optimized_amd.dis import fixed_sobol

    seen = []

    def fake(entries, dipole, config, **kw):
        seen.extend(entries)
        return [_fixed(2.0, 0.1, True, e[4]) for e in entries]

    monkeypatch.setattr(fixed_sobol, "gbw_qg_fixed_sobol_batch", fake)
    light, charm = Quark(QuarkType.LIGHT, 0.14), Quark(QuarkType.C, 1.27)
    res = obs.sigma_qg_d2b(torch.tensor(9.0), torch.tensor(1e-3), Polarization.T, GBW(),
                           DISConfig(quarks=(light, charm)), seed=11)
    want = (light.charge**2 + charm.charge**2) * 2.0 * QG_FACTOR
    torch.testing.assert_close(res.value, torch.tensor(want, dtype=torch.float64))
    err = math.sqrt(light.charge**4 + charm.charge**4) * 0.1 * QG_FACTOR
    torch.testing.assert_close(res.error, torch.tensor(err, dtype=torch.float64))
    assert [e[4] for e in seen] == [11, 12]  # seed + quark index
    assert [e[3] for e in seen] == [light, charm]
    assert not fakes.calls


def test_qg_falls_back_to_vegas_when_a_total_misses_epsrel(fakes, monkeypatch):
    """S/dis/observables.py:476-491: one unconverged quark sends every quark of
    the term to VEGAS, seeded seed + 0, 1, ..."""

    from nlo_bk_optimized_amd.dis import fixed_sobol

    def fake(entries, dipole, config, **kw):
        return [_fixed(2.0, 0.1, k == 0, e[4]) for k, e in enumerate(entries)]

    monkeypatch.setattr(fixed_sobol, "gbw_qg_fixed_sobol_batch", fake)
    fakes.vegas.update({"I1": (1.0, 0.0), "I2_nested": (0.5, 0.0), "I3_nested": (0.25, 0.0)})
    light, charm = Quark(QuarkType.LIGHT, 0.14), Quark(QuarkType.C, 1.27)
    res = obs.sigma_qg_d2b(torch.tensor(9.0), torch.tensor(1e-3), Polarization.L, GBW(),
                           DISConfig(quarks=(light, charm)), seed=3)
    want = (light.charge**2 + charm.charge**2) * 1.75 * QG_FACTOR
    torch.testing.assert_close(res.value, torch.tensor(want, dtype=torch.float64))
    assert [c[1] for c in fakes.calls] == ["I1", "I2_nested", "I3_nested"] * 2
    assert [c[3]["seed"] for c in fakes.calls] == [3, 4, 5, 6, 7, 8]


def test_fixed_sobol_arguments_are_checked_on_the_host():
    """S/dis/fixed_sobol.py:79-85, 286-294: sizes and scalars are refused before
    any device work."""

    from nlo_bk_optimized_amd.dis.fixed_sobol import gbw_qg_fixed_sobol_batch, make_sobol_points

    charm = Quark(QuarkType.C, 1.4)
    q, x = torch.tensor(9.0), torch.tensor(1e-3)
    entry = [(q, x, Polarization.T, charm, 0)]
    with pytest.raises(TypeError):
        gbw_qg_fixed_sobol_batch(entry, object(), DISConfig())
    with pytest.raises(ValueError, match="powers of two"):
        gbw_qg_fixed_sobol_batch(entry, GBW(), DISConfig(), samples_per_replicate=1000)
    with pytest.raises(ValueError, match="replicates must exceed one"):
        gbw_qg_fixed_sobol_batch(entry, GBW(), DISConfig(), replicates=1)
    with pytest.raises(ValueError, match="matching"):
        gbw_qg_fixed_sobol_batch([(q, x.double(), Polarization.T, charm, 0)], GBW(), DISConfig())
    with pytest.raises(ValueError, match="matching"):
        gbw_qg_fixed_sobol_batch([(9.0, 1e-3, Polarization.T, charm, 0)], GBW(), DISConfig())
    with pytest.raises(ValueError, match="powers of two"):
        make_sobol_points(5, training_samples=100, training_rounds=1, samples_per_replicate=64, replicates=2,
                          device="cpu", dtype=torch.float32, seed=0)
