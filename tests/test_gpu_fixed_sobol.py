"""The GBW fixed-Sobol qqbar-g estimator (nlo_dis_fixed_sobol) and the
paired-endpoint map (nlo_dis_endpoint_map) on the GPU against the oracle
(oracle/fixed_sobol_oracle.py).

* Endpoint map: float32 and float64, with and without the angle fold, through a
  non-uniform grid: every output against the oracle's
  map_importance_grid + paired_endpoint_inputs (x inverse density) at a few
  float32 / float64 ulps of the value (the device's cos / sin / atan2 are not
  numpy's; angle fractions compared on the circle).
* The whole estimator in FP64 at reduced sizes (two training sets of 2048
  points, three replicates of 2048, 8 bins), both polarisations, at the
  benchmark point (charm) and a light-quark point (m = 0.14, Q^2 = 50,
  x = 1e-2, maxr = 30): the learned
  grid, I1, I2, I3 and the total (value and error) against the oracle at 1e-9
  relative; in FP32 (the reference's path: float32 map and endpoint split,
  nested I2 / I3 in FP32, I1 in FP64) at the FP32 integrand tolerance.
* Edge sizes in FP64 against the oracle: sets smaller than a block (64 and 32
  points, 2 bins), two points per set, and the most replicates (32) with the
  most bins (255) on 256 points (1e-8 on the edges, 1e-7 on the estimates:
  bins holding one point or none amplify the summation order); the endpoint
  map of no point and of one point.
* An estimator batched with others equals it alone, bit for bit; the frozen
  grid of a run, passed back, reproduces the replicate estimates with the
  replicate evaluations only.
* At the reference's benchmark point and full sizes (851 968 evaluations per
  polarisation) the six I1 / I2 / I3 values against the reference's 20-worker
  C++ values with their errors (T/benchmarks/dis_fixed_sobol_all_qg.py:69-76).
* sigma_qg_d2b takes the fixed-Sobol path where the reference does
  (S/dis/observables.py:470-491) and VEGAS below the evaluation budget.
* The generic forms over Python callables: learn_triple_importance_grid on
  the device against the reference's own grids (golden vectors, 1e-11),
  paired_endpoint_inputs against the oracle, integrate_triple_fixed_sobol's
  known answer (T/tests/test_dis_production.py:153-192).
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import dis_oracle as D
from oracle import fixed_sobol_oracle as F

pytestmark = pytest.mark.gpu

Q2, XBJ, MASS, MAXR = 9.0, 1e-3, 1.4, 5.0


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nlo_bk_optimized_amd import device

    return device


def _gbw():
    from nlo_bk_optimized_amd.dipole import GBW

    return GBW(Qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0)


def _config(**kw):
    from nlo_bk_optimized_amd.dis.config import DISConfig, Quark, QuarkType

    return DISConfig(quarks=(Quark(QuarkType.C, MASS),), maxr=MAXR, nf_alphas=1, cuda_nested_points=48, **kw)


def _grid(bins: int, seed: int = 4) -> np.ndarray:
    rng = np.random.default_rng(seed)
    inner = np.sort(rng.uniform(0.02, 0.98, (5, bins - 1)), axis=1)
    return np.concatenate([np.zeros((5, 1)), inner, np.ones((5, 1))], axis=1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("fold", [True, False])
def test_endpoint_map_matches_oracle(dev, dtype, fold):
    print(f"endpoint map dtype={np.dtype(dtype).name} fold={fold}", flush=True)
    points = F.sobol_points(21, 4096, dtype)
    edges = _grid(16).astype(dtype)
    first, w1, second, w2 = F.endpoint_map(points, edges, MAXR, fold_angle=fold)
    got = dev.dis_endpoint_map(torch.from_numpy(points).cuda(), torch.from_numpy(edges).cuda(), MAXR,
                               fold_angle=fold)
    got = [g.cpu().numpy() for g in got]
    assert all(g.dtype == np.dtype(dtype) for g in got)
    ulp = float(np.finfo(dtype).eps)
    # mapped coordinates 0..2 and the first sector: map arithmetic plus cos / sin
    np.testing.assert_allclose(got[0][:, :3], first[:, :3], rtol=2 * ulp, atol=0)
    np.testing.assert_allclose(got[2][:, :3], second[:, :3], rtol=2 * ulp, atol=0)
    np.testing.assert_allclose(got[0][:, 3:], first[:, 3:], rtol=64 * ulp, atol=8 * ulp)
    np.testing.assert_allclose(got[1], w1, rtol=64 * ulp, atol=0)
    # second sector: radial coordinate and weight, angle on the circle
    np.testing.assert_allclose(got[2][:, 3], second[:, 3], rtol=256 * ulp, atol=16 * ulp)
    np.testing.assert_allclose(got[3], w2, rtol=256 * ulp, atol=16 * ulp)
    da = np.abs(got[2][:, 4].astype(np.float64) - second[:, 4])
    assert float(np.max(np.minimum(da, 1 - da))) <= 256 * ulp


def _device_run(dev, precision: str, transverse: bool, *, seed=7, training=2048, rounds=2, reps_n=2048, reps=3,
                bins=8, return_edges=True, entries=None, point=(Q2, XBJ, MASS, MAXR)):
    from nlo_bk_optimized_amd.dis import Polarization
    from nlo_bk_optimized_amd.dis.config import DISConfig, Quark, QuarkType
    from nlo_bk_optimized_amd.dis.fixed_sobol import gbw_qg_fixed_sobol_batch

    dtype = torch.float32 if precision == "fp32" else torch.float64
    q2, xbj, mass, maxr = point
    quark = Quark(QuarkType.C if mass > 1 else QuarkType.LIGHT, mass)
    cfg = DISConfig(quarks=(quark,), maxr=maxr, nf_alphas=1 if mass > 1 else 3, cuda_nested_points=48)
    if entries is None:
        entries = [(torch.tensor(q2, dtype=dtype), torch.tensor(xbj, dtype=dtype),
                    Polarization.T if transverse else Polarization.L, quark, seed)]
    return gbw_qg_fixed_sobol_batch(entries, _gbw(), cfg, training_samples=training, training_rounds=rounds,
                                    samples_per_replicate=reps_n, replicates=reps, bins=bins,
                                    return_edges=return_edges)


def _oracle_run(dtype, transverse: bool, *, seed=7, training=2048, rounds=2, reps_n=2048, reps=3, bins=8,
                point=(Q2, XBJ, MASS, MAXR)):
    q2, xbj, mass, maxr = point
    charge = 2 / 3 if mass > 1 else math.sqrt(2 / 3)
    return F.gbw_qg_fixed_sobol(q2, xbj, transverse, D.QuarkSpec(mass, charge),
                                D.GBWParams(qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0),
                                D.DisSetup(maxr=maxr, nf_alphas=1 if mass > 1 else 3), seed=seed, dtype=dtype,
                                training_samples=training, training_rounds=rounds, samples_per_replicate=reps_n,
                                replicates=reps, bins=bins, return_edges=True)


# the benchmark point and a light-quark point at larger Q^2 and x with the default maxr
POINTS = [(Q2, XBJ, MASS, MAXR), (50.0, 1e-2, 0.14, 30.0)]


def _rows(result):
    return list(result.contributions) + [result.total]


@pytest.mark.parametrize("point", POINTS)
@pytest.mark.parametrize("transverse", [False, True])
def test_fixed_sobol_fp64_matches_oracle(dev, transverse, point):
    print(f"fixed-Sobol FP64 transverse={transverse} point={point}", flush=True)
    (res,), edges = _device_run(dev, "fp64", transverse, point=point)
    contributions, total, oedges = _oracle_run(np.float64, transverse, point=point)
    np.testing.assert_allclose(edges[0].cpu().numpy(), oedges, rtol=0, atol=1e-11)
    for got, want in zip(_rows(res), list(contributions) + [total]):
        print(f"  {float(got.value):+.12e} +- {float(got.error):.3e}  oracle {want.value:+.12e} +- {want.error:.3e}")
        assert float(got.value) == pytest.approx(want.value, rel=1e-9, abs=1e-14)
        assert float(got.error) == pytest.approx(want.error, rel=1e-6, abs=1e-14)
        assert got.n_eval == want.n_eval == 2 * 2 * 2048 + 2 * 3 * 2048
        assert got.converged == want.converged


@pytest.mark.parametrize("point", POINTS)
@pytest.mark.parametrize("transverse", [False, True])
def test_fixed_sobol_fp32_matches_oracle(dev, transverse, point):
    """FP32 path: the learned grid within 1e-4 (FP32 integrand values move the
    histogram), the estimates within 2e-4 relative of the oracle's
    (FP64-integrand) values or a tenth of their standard error."""

    print(f"fixed-Sobol FP32 transverse={transverse} point={point}", flush=True)
    (res,), edges = _device_run(dev, "fp32", transverse, point=point)
    contributions, total, oedges = _oracle_run(np.float32, transverse, point=point)
    np.testing.assert_allclose(edges[0].cpu().numpy(), oedges, rtol=0, atol=1e-4)
    for got, want in zip(_rows(res), list(contributions) + [total]):
        print(f"  {float(got.value):+.10e} +- {float(got.error):.3e}  oracle {want.value:+.10e} +- {want.error:.3e}")
        tol = max(2e-4 * abs(want.value), 0.1 * want.error)
        assert abs(float(got.value) - want.value) <= tol
        assert float(got.error) == pytest.approx(want.error, rel=0.05, abs=1e-7)


def test_batched_estimators_equal_estimators_alone(dev):
    from nlo_bk_optimized_amd.dis import Polarization

    cfg = _config()
    q = (torch.tensor(Q2, dtype=torch.float32), torch.tensor(XBJ, dtype=torch.float32))
    q2 = (torch.tensor(30.0, dtype=torch.float32), torch.tensor(1e-2, dtype=torch.float32))
    entries = [(*q, Polarization.L, cfg.quarks[0], 3), (*q2, Polarization.T, cfg.quarks[0], 3),
               (*q, Polarization.T, cfg.quarks[0], 9)]
    print("fixed-Sobol batch of three vs alone", flush=True)
    batch, bedges = _device_run(dev, "fp32", True, entries=entries)
    for k, e in enumerate(entries):
        (alone,), aedges = _device_run(dev, "fp32", True, entries=[e])
        assert torch.equal(bedges[k], aedges[0])
        for b, a in zip(_rows(batch[k]), _rows(alone)):
            assert torch.equal(b.value, a.value) and torch.equal(b.error, a.error) and b.n_eval == a.n_eval


def test_frozen_grid_reproduces_the_replicates(dev):
    from nlo_bk_optimized_amd.dis import Polarization
    from nlo_bk_optimized_amd.dis.fixed_sobol import gbw_qg_fixed_sobol_batch

    cfg = _config()
    e = [(torch.tensor(Q2, dtype=torch.float32), torch.tensor(XBJ, dtype=torch.float32), Polarization.L,
          cfg.quarks[0], 11)]
    print("fixed-Sobol frozen grid", flush=True)
    (learned,), edges = gbw_qg_fixed_sobol_batch(e, _gbw(), cfg, training_samples=4096, samples_per_replicate=2048,
                                                 return_edges=True)
    (frozen,) = gbw_qg_fixed_sobol_batch(e, _gbw(), cfg, training_samples=4096, samples_per_replicate=2048,
                                         grid=edges)
    for a, b in zip(_rows(learned), _rows(frozen)):
        assert torch.equal(a.value, b.value) and torch.equal(a.error, b.error)
        assert a.n_eval == 2 * 3 * 4096 + 2 * 7 * 2048 and b.n_eval == 2 * 7 * 2048


def test_qqbarg_fixed_sobol_matches_reference_cpp_values(dev):
    """T/benchmarks/dis_fixed_sobol_all_qg.py:69-76: the reference's 20-worker
    C++ values; ours by the fixed-Sobol estimator at full sizes (FP32 path,
    seed 3100 as the reference's benchmark) within 4 combined standard errors."""

    from nlo_bk_optimized_amd.dis import Polarization
    from nlo_bk_optimized_amd.dis.fixed_sobol import FIXED_SOBOL_EVALUATIONS, gbw_qg_fixed_sobol_batch

    ref = {
        "L": ((0.10206739088678346, 0.0002535000977528945), (-0.024439455902300525, 0.00004764427185133986),
              (0.009414936380564324, 0.000026852425497947778)),
        "T": ((0.40263591173620866, 0.0017956984452001683), (-0.08719375357937224, 0.00026850014906050766),
              (0.04848797141670969, 0.00013338145018266037)),
    }
    cfg = _config()
    q = (torch.tensor(Q2, dtype=torch.float32), torch.tensor(XBJ, dtype=torch.float32))
    entries = [(*q, Polarization.L, cfg.quarks[0], 3100), (*q, Polarization.T, cfg.quarks[0], 3100)]
    print("fixed-Sobol six qqbar-g terms at the reference benchmark point", flush=True)
    results = gbw_qg_fixed_sobol_batch(entries, _gbw(), cfg)
    for pol, res in zip(("L", "T"), results):
        assert res.total.n_eval == FIXED_SOBOL_EVALUATIONS
        for name, got, (value, error) in zip(("I1", "I2", "I3"), res.contributions, ref[pol]):
            ours, err = float(got.value), float(got.error)
            pull = (ours - value) / math.hypot(err, error)
            print(f"  {pol}_{name}: {ours:.8f} +- {err:.2e} vs {value:.8f} +- {error:.2e} ({pull:+.2f} sigma)")
            assert abs(pull) <= 4.0


def test_sigma_qg_takes_the_fixed_sobol_path(dev):
    """The fixed-Sobol totals times e_q^2 4 Nc alpha_em / (2 pi)^3 2 pi when
    every quark's total converges; VEGAS for the term when one does not, and
    below the fixed-Sobol evaluation budget (S/dis/observables.py:470-532)."""

    from nlo_bk_optimized_amd.dis import Polarization
    from nlo_bk_optimized_amd.dis.config import NC, AlphaEM
    from nlo_bk_optimized_amd.dis.fixed_sobol import FIXED_SOBOL_EVALUATIONS, gbw_qg_fixed_sobol
    from nlo_bk_optimized_amd.dis.observables import sigma_qg_d2b

    q2, x = torch.tensor(Q2, dtype=torch.float32), torch.tensor(XBJ, dtype=torch.float32)
    print("sigma_qg fixed-Sobol dispatch", flush=True)
    loose = _config(epsrel=1e-2)
    sigma = sigma_qg_d2b(q2, x, Polarization.T, _gbw(), loose, seed=11)
    fixed = gbw_qg_fixed_sobol(q2, x, Polarization.T, loose.quarks[0], _gbw(), loose, seed=11)
    assert fixed.total.converged
    factor = loose.quarks[0].charge ** 2 * 4 * NC * AlphaEM / (2 * math.pi) ** 3 * 2 * math.pi
    assert float(sigma.value) == pytest.approx(float(fixed.total.value) * factor, rel=1e-14)
    assert float(sigma.error) == pytest.approx(float(fixed.total.error) * factor, rel=1e-14)
    assert sigma.n_eval == FIXED_SOBOL_EVALUATIONS
    # a total that misses epsrel sends the term to VEGAS
    tight = _config(epsrel=1e-3)
    missed = gbw_qg_fixed_sobol(q2, x, Polarization.T, tight.quarks[0], _gbw(), tight, seed=11)
    assert not missed.total.converged
    fallback = sigma_qg_d2b(q2, x, Polarization.T, _gbw(), tight, seed=11)
    assert fallback.n_eval != FIXED_SOBOL_EVALUATIONS
    assert abs(float(fallback.value) - float(sigma.value)) <= 5 * math.hypot(float(fallback.error),
                                                                             float(sigma.error))
    # below the fixed-Sobol budget the term is integrated by VEGAS
    small = _config(maxeval=100_000, epsrel=1e-2)
    vegas = sigma_qg_d2b(q2, x, Polarization.T, _gbw(), small, seed=11)
    assert vegas.n_eval != FIXED_SOBOL_EVALUATIONS
    assert abs(float(vegas.value) - float(sigma.value)) <= 5 * math.hypot(float(vegas.error), float(sigma.error))


@pytest.mark.parametrize("sizes", [
    dict(training=64, rounds=1, reps_n=32, reps=2, bins=2),     # partial blocks, two bins
    dict(training=256, rounds=3, reps_n=128, reps=32, bins=255),  # the most replicates and bins
    dict(training=2, rounds=2, reps_n=2, reps=3, bins=4),       # two points per set
])
def test_fixed_sobol_edge_sizes_match_oracle(dev, sizes):
    print(f"fixed-Sobol FP64 edge sizes {sizes}", flush=True)
    (res,), edges = _device_run(dev, "fp64", True, **sizes)
    contributions, total, oedges = _oracle_run(np.float64, True, **sizes)
    # with 255 bins and 256 points most bins hold one point or none, and the
    # refinement divides by their small smoothed importances: the summation
    # order of the histograms (per block, then over blocks, against the
    # oracle's point order) shows up at 1e-9 in the edges and in the estimates
    np.testing.assert_allclose(edges[0].cpu().numpy(), oedges, rtol=0, atol=1e-8)
    for got, want in zip(_rows(res), list(contributions) + [total]):
        assert float(got.value) == pytest.approx(want.value, rel=1e-7, abs=1e-14)
        assert float(got.error) == pytest.approx(want.error, rel=1e-5, abs=1e-14)
        assert got.n_eval == want.n_eval


def test_endpoint_map_empty_and_single_point(dev):
    edges = torch.linspace(0, 1, 9, dtype=torch.float32).expand(5, -1).contiguous().cuda()
    empty = dev.dis_endpoint_map(torch.empty((0, 5), dtype=torch.float32, device="cuda"), edges, MAXR)
    assert all(t.shape[0] == 0 for t in empty)
    one = torch.tensor([[0.3, 0.6, 0.5, 0.25, 0.2]], dtype=torch.float32)
    got = [t.cpu().numpy() for t in dev.dis_endpoint_map(one.cuda(), edges, MAXR, fold_angle=True)]
    want = F.endpoint_map(one.numpy(), edges.cpu().numpy(), MAXR, fold_angle=True)
    for g, w in zip(got, want):
        np.testing.assert_allclose(g, w, rtol=1e-5, atol=1e-6)


def _synthetic_torch(x):
    """tests/golden/make_fixed_sobol_golden.py synthetic() as the reference's
    function(x) returns it (regular, sensitive in float64, I2, I3)."""

    phi = 2 * math.pi * x[:, 4]
    regular = x[:, 0] * (1 + x[:, 3]) * torch.cos(phi) + 0.1
    sensitive = 0.01 * x[:, 1] * x[:, 2]
    i2 = torch.exp(-3 * x[:, 1]) * x[:, 2] - 0.05
    i3 = (x[:, 3] - 0.4) * x[:, 0] ** 2 * (1 + torch.cos(phi) ** 2)
    return regular, sensitive.double(), i2, i3


def test_generic_grid_learning_matches_reference_vectors(dev):
    """learn_triple_importance_grid over a Python callable, on the device,
    against the reference's own function's grids
    (tests/golden/fixed_sobol_reference_vectors.npz)."""

    from pathlib import Path

    from nlo_bk_optimized_amd.dis.fixed_sobol import learn_triple_importance_grid, make_sobol_points

    golden = np.load(Path(__file__).resolve().parent / "golden" / "fixed_sobol_reference_vectors.npz")
    for k in range(3):
        seed, sets, count, bins, fold, i1_weight, maxr = golden[f"case{k}_params"]
        print(f"generic grid learning case {k}", flush=True)
        pts = make_sobol_points(5, training_samples=int(count), training_rounds=int(sets), samples_per_replicate=2,
                                replicates=2, device="cuda", dtype=torch.float64, seed=int(seed))
        edges = learn_triple_importance_grid(_synthetic_torch, pts.training, maxr=float(maxr), bins=int(bins),
                                             fold_angle=bool(fold), i1_weight=float(i1_weight))
        np.testing.assert_allclose(edges.cpu().numpy(), golden[f"case{k}_edges"], rtol=0, atol=1e-11)


def test_generic_paired_endpoint_inputs_match_oracle(dev):
    from nlo_bk_optimized_amd.dis.fixed_sobol import paired_endpoint_inputs

    src = F.sobol_points(31, 2048, np.float64)
    for fold in (False, True):
        got = [t.cpu().numpy() for t in paired_endpoint_inputs(torch.from_numpy(src).cuda(), MAXR, fold_angle=fold)]
        want = F.paired_endpoint_inputs(src, MAXR, fold_angle=fold)
        for g, w in zip(got, want):
            np.testing.assert_allclose(g, w, rtol=1e-13, atol=1e-14)


def test_generic_integrate_known_answer(dev, monkeypatch):
    """T/tests/test_dis_production.py:153-192 on device tensors."""

    from nlo_bk_optimized_amd.dis import fixed_sobol

    monkeypatch.setattr(fixed_sobol, "dis_sobol_endpoint_map", lambda points, edges, maxr, fold_angle: (
        points, torch.ones(points.shape[0], device=points.device), points,
        torch.zeros(points.shape[0], device=points.device)))

    def i1_sums(x, weights):
        return x[:, 0] * weights, torch.zeros(1, dtype=torch.float64, device=x.device)

    def i2_i3(x):
        return x[:, 1], -0.5 * x[:, 0]

    replicates = tuple(torch.tensor(v, device="cuda") for v in
                       ([[0.1, 0.8], [0.3, 0.4]], [[0.5, 0.2], [0.7, 0.6]], [[0.2, 0.9], [0.4, 0.1]]))
    result = fixed_sobol.integrate_triple_fixed_sobol(i1_sums, i2_i3, torch.empty(0), replicates, maxr=5.0,
                                                      epsrel=1.0, seed=13, training_evaluations=10)
    totals = torch.tensor([0.7, 0.7, 0.65], dtype=torch.float64)
    torch.testing.assert_close(result.total.value.cpu(), totals.mean())
    torch.testing.assert_close(result.total.error.cpu(), totals.std(unbiased=True) / math.sqrt(3))
    assert result.total.n_eval == 22 and result.total.seed == 13


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("transverse", [False, True])
def test_nested_pair_equals_the_single_terms(dev, precision, transverse):
    """nlo_dis_qg_i23 (both nested terms from one pass over the inner rule)
    against the nested I2 and I3 kinds of nlo_dis_integrand at the same points
    (the same operations: FP64 to 1e-13, FP32 to 1e-6 relative) and, in FP64,
    against the oracle."""

    from nlo_bk_optimized_amd.dis import Polarization
    from nlo_bk_optimized_amd.dis.observables import integrand_qg_nested_massive, integrand_qg_nested_pair

    cfg = _config()
    pol = Polarization.T if transverse else Polarization.L
    dtype = torch.float32 if precision == "fp32" else torch.float64
    q2, x = torch.tensor(Q2, dtype=dtype), torch.tensor(XBJ, dtype=dtype)
    pts = torch.from_numpy(F.sobol_points(41, 4096, np.float64))
    print(f"nested pair {precision} transverse={transverse}", flush=True)
    i2, i3 = integrand_qg_nested_pair(pts, q2, x, pol, cfg.quarks[0], _gbw(), cfg)
    s2 = integrand_qg_nested_massive(pts, q2, x, pol, cfg.quarks[0], "I2", _gbw(), cfg)
    s3 = integrand_qg_nested_massive(pts, q2, x, pol, cfg.quarks[0], "I3", _gbw(), cfg)
    rtol = 1e-13 if precision == "fp64" else 1e-6
    for a, b in ((i2, s2), (i3, s3)):
        scale = float(b.abs().max())
        np.testing.assert_allclose(a.cpu().numpy(), b.cpu().numpy(), rtol=rtol, atol=rtol * 1e-3 * scale)
    if precision == "fp64":
        amp = lambda r, Y: D.gbw(r, Y, D.GBWParams(qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0))  # noqa: E731
        setup = D.DisSetup(maxr=MAXR, nf_alphas=1)
        for c, got in (("I2", i2), ("I3", i3)):
            want = D.qg_integrand(pts.numpy(), Q2, XBJ, transverse, D.QuarkSpec(MASS, 2 / 3), c, amp, setup,
                                  nested_points=48)
            scale = float(np.max(np.abs(want)))
            np.testing.assert_allclose(got.cpu().numpy(), want, rtol=1e-9, atol=1e-12 * scale)
