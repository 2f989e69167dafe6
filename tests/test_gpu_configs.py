"""The workload configurations and every fixed-step rapidity method on the
HIP path, against the oracle's own RK driver (tests/bk_helpers.oracle_evolve,
numpy stages over the C oracle's right-hand side).

* The 64-node plain-FP64 workload (R = 64 on [1e-6, 1e2], MV Qs0^2 = 0.2,
  gamma = 1, 2^14 Sobol points per parent, midpoint RK2, dY = 0.1 to Y = 1)
  through the drop-in ``solve_bk``.
* Euler, midpoint RK2, Kutta's third-order step (the reference's production
  step, S/bk/evolution.py:190-236) and classical RK4 on the device evolver.
* The batched driver, each run against ``oracle_evolve`` for its own grid.
* The batched workload's shape (64 runs, R ~ U{128..512}, 2^18 points per
  parent) for two steps: batch equal to single runs, finite, in [0, 1].
* ``compute_dndy`` (S/bk/evolution.py:61-82) against the oracle's
  right-hand side.

Tolerances: 1e-10 relative in FP64 with the cancellation-free formula; the
reference K2/Kf formula additionally within the rounding scale it carries
where its terms cancel (unit roundoff times the integrated sum of |terms| the
oracle reports, per right-hand side).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.bk_helpers import bk_params, oracle_evolve, oracle_rhs, random_edges, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nlo_bk_optimized_amd import device

    return device


def cuda(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def reference_formula_scale(r, N0, edges, n, p):
    """Per node, the rounding scale of one right-hand side on the reference
    K2/Kf formula: unit roundoff times the integrated sum of |terms| / n that
    the oracle reports (its terms cancel by up to 1e12 at r ~ 1e-3 and more
    below, so two evaluation orders agree only to this scale there)."""

    _, sS = O.evolution_splines(r, N0)
    lo, width = O.nlo_bounds(r)
    act = np.nonzero((np.arange(r.size) >= 1) & (N0 <= 0.99999))[0]
    mag = np.zeros(r.size)
    mag[act] = O.bk_nlo_sums(r[act], sS, edges[act], lo, width, O.sobol_table(4, 0), 0, n, p,
                             want_magnitude=True)[3]
    return 2.0**-52 * mag / n


@pytest.mark.parametrize("precision", ["fp64", "fp64_stable"])
def test_64_node_fp64_rk2_to_y1_matches_oracle(dev, precision):
    """R = 64, 2^14 Sobol points (seed 0), midpoint RK2, dY = 0.1, Y: 0 -> 1,
    through solve_bk.  "fp64" is the plain FP64 path (the reference K2/Kf
    formula on the S spline): below r ~ 1e-3 its terms cancel by many orders
    of magnitude, so GPU and oracle (different operation order) agree there
    only to the formula's own rounding scale (unit roundoff times the
    integrated sum of |terms| per right-hand side), accumulated over the
    steps.  "fp64_stable" (the FP32 path's cancellation-free formula in FP64)
    agrees to 1e-10 relative everywhere."""

    from nlo_bk_optimized_amd import MV, BKConfig, solve_bk
    from nlo_bk_optimized_amd.bk import evolution as evo

    n, steps, h = 1 << 14, 10, 0.1
    cfg = BKConfig(RPOINTS=63, MINR=1e-6, MAXR=1e2, MCINTPOINTS=n, DE_SOLVER_STEP=h, RK_METHOD="rk2",
                   K1_RULE="fixed", PRECISION=precision, LambdaQCD=0.241, NF=3)
    ic = MV(qs0sqr=0.2, gamma=1.0, ec=1.0)
    print(f"64-node {precision} RK2 to Y = 1 through solve_bk", flush=True)
    table = solve_bk(ic, 1.0, cfg, device="cuda", dtype=torch.float64, seed=0)
    assert table.N.shape == (11, 64)
    np.testing.assert_allclose(table.y.cpu().numpy(), np.arange(11) * 0.1, rtol=0, atol=1e-15)
    setup = evo.prepare(ic, cfg, device=torch.device("cuda"), seed=0)  # deterministic: solve_bk's grids
    r, N0, edges = (x.cpu().numpy() for x in (setup.r, setup.N0, setup.edges))
    stable = precision == "fp64_stable"
    ref = oracle_evolve(r, N0, edges, O.sobol_table(4, 0), n, bk_params(stable), steps=steps, h=h, method="rk2")
    got = table.N.cpu().numpy()
    err = np.abs(got - ref)
    sN, sS = O.evolution_splines(r, N0)
    act = np.nonzero((np.arange(r.size) >= 1) & (N0 <= 0.99999))[0]
    # K1: 1e-14 of its integrated |integrand| per right-hand side (its own cancellation at small r)
    k1mag = np.zeros(r.size)
    k1mag[act] = O.k1_fixed(r[act], sN, p=bk_params(stable), magnitude=True)
    tol = 1e-10 * np.abs(ref) + steps * h * 2.0 * 1e-14 * k1mag + 1e-300
    if not stable:
        # the scale grows with N over the evolution: twice the initial one, per step
        tol = tol + steps * h * 2.0 * reference_formula_scale(r, N0, edges, n, bk_params(False))
    bad = err > tol
    print(f"max |dN| {err.max():.3e}; max rel {np.max(err / np.maximum(np.abs(ref), 1e-300)):.3e}; "
          f"max err / tol {np.max(err / tol):.3e}")
    assert not bad.any(), (np.argwhere(bad)[:5], err[bad][:5], ref[bad][:5])


@pytest.mark.parametrize("method", ["euler", "rk2", "kutta3", "rk4"])
@pytest.mark.parametrize("precision", ["fp64_stable", "fp64"])
def test_evolver_methods_match_oracle(dev, method, precision):
    from nlo_bk_optimized_amd.bk.config import BKConfig

    R, n, bins, steps, h = 12, 2048, 50, 3, 0.2
    r, N0 = workload(R=R, minr=1e-3)
    edges = random_edges(R, bins, seed=7)
    config = BKConfig(MINR=1e-3)
    table = dev.sobol_table(4, 0)
    print(f"evolver {method} {precision}", flush=True)
    ev = dev.Evolver(cuda(r), cuda(edges), table, config, n_samples=n, precision=precision, rk_method=method,
                     step=h)
    ev.state.copy_(cuda(N0))
    got = [N0]
    for _ in range(steps):
        ev.step()
        got.append(ev.state.cpu().numpy().copy())
    assert int(ev.flag.item()) == 0
    ev.close()
    p = O.params(minr=1e-3, stable=precision == "fp64_stable")
    ref = oracle_evolve(r, N0, edges, O.sobol_table(4, 0), n, p, steps=steps, h=h, method=method)
    tol = 1e-10 * np.abs(ref) + 1e-18
    if precision == "fp64":
        tol = tol + steps * h * 2.0 * reference_formula_scale(r, N0, edges, n, p)
    err = np.abs(np.array(got) - ref)
    print(f"max rel {np.max(err / np.maximum(np.abs(ref), 1e-300)):.3e}, max err / tol {np.max(err / tol):.3e}")
    assert np.all(err <= tol), (np.argwhere(err > tol)[:5], err[err > tol][:5])
    # the methods differ: a wrong tableau cannot hide behind another one
    if method != "euler":
        other = oracle_evolve(r, N0, edges, O.sobol_table(4, 0), n, p, steps=1, h=h, method="euler")
        assert np.abs(other[1] - ref[1]).max() > 1e-6


@pytest.mark.parametrize("method", ["rk2", "kutta3"])
def test_batch_runs_match_oracle(dev, method):
    from nlo_bk_optimized_amd.bk.config import BKConfig

    n, steps, h = 2048, 2, 0.1
    runs = []
    for k, (R, q) in enumerate(((9, 0.2), (14, 0.45), (11, 0.1))):
        r, N = workload(R=R, qs0sqr=q)
        runs.append((r, N, random_edges(R, 50, seed=31 + k)))
    cfg = BKConfig(MINR=1e-6)
    table = dev.sobol_table(4, 0)
    print(f"batch {method} vs oracle per run", flush=True)
    bev = dev.BatchEvolver(cuda(np.concatenate([x[0] for x in runs])), cuda(np.concatenate([x[2] for x in runs])),
                           [x[0].size for x in runs], table, cfg, n_samples=n, precision="fp64_stable",
                           rk_method=method, step=h, spt=8, grid_blocks=64)
    bev.state.copy_(cuda(np.concatenate([x[1] for x in runs])))
    for _ in range(steps):
        bev.step()
    got = bev.state.cpu().numpy()
    assert int(bev.flag.item()) == 0
    bev.close()
    off = 0
    for r, N, e in runs:
        ref = oracle_evolve(r, N, e, O.sobol_table(4, 0), n, bk_params(True), steps=steps, h=h, method=method)
        np.testing.assert_allclose(got[off:off + r.size], ref[-1], rtol=1e-10, atol=1e-18)
        off += r.size


def test_batched_workload_shape_two_steps(dev):
    """64 runs with Qs0^2 ~ U[0.1, 0.5], gamma ~ U[0.9, 1.2], e_c ~ U[1, 20],
    R ~ U{128..512} on [1e-6, 1e2], 2^18 Sobol points per parent, FP32, RK2,
    dY = 0.1, two steps: size-independent properties (equal to the runs
    evolved one at a time with the same launch split, finite, in [0, 1])."""

    from nlo_bk_optimized_amd.bk.config import BKConfig

    rng = np.random.default_rng(2024)
    runs = 64
    qs = rng.uniform(0.1, 0.5, runs)
    gam = rng.uniform(0.9, 1.2, runs)
    ec = rng.uniform(1.0, 20.0, runs)
    sizes = rng.integers(128, 513, runs)
    n, h = 1 << 18, 0.1
    cfg = BKConfig(MINR=1e-6)
    table = dev.sobol_table(4, 0)
    rs, N0s, es = [], [], []
    for k in range(runs):
        r = O.r_grid(1e-6, 1e2, int(sizes[k]) - 1)
        rs.append(r)
        N0s.append(O.mv(r, qs[k], gam[k], ec[k], 0.241))
        es.append(random_edges(r.size, 50, seed=100 + k))
    print(f"batched workload shape: {runs} runs, sum R = {int(sizes.sum())}, 2^18 points", flush=True)
    bev = dev.BatchEvolver(cuda(np.concatenate(rs)), cuda(np.concatenate(es)), sizes.tolist(), table, cfg,
                           n_samples=n, precision="fp32", rk_method="rk2", step=h, spt=16, grid_blocks=0)
    bev.state.copy_(cuda(np.concatenate(N0s)))
    for _ in range(2):
        bev.step()
    got = bev.state.cpu().numpy()
    assert int(bev.flag.item()) == 0
    evals = int(bev.active_evaluations.item())
    bev.close()
    assert np.all(np.isfinite(got)) and np.all((got >= 0) & (got <= 1))
    off, exact, ref_evals = 0, 0, 0
    for k in range(runs):
        R = rs[k].size
        ev = dev.Evolver(cuda(rs[k]), cuda(es[k]), table, cfg, n_samples=n, precision="fp32", rk_method="rk2",
                         step=h)
        ev.state.copy_(cuda(N0s[k]))
        for _ in range(2):
            ev.step()
        single = ev.state.cpu().numpy()
        ref_evals += int(ev.active_evaluations.item())
        ev.close()
        part = got[off:off + R]
        # R > 128: the single-run plan also takes 16 samples per lane -> the same sums bit for bit
        if R > 128:
            np.testing.assert_array_equal(part, single)
            exact += 1
        else:
            np.testing.assert_allclose(part, single, rtol=0, atol=1e-12)
        off += R
    assert evals == ref_evals
    print(f"{exact} of {runs} runs bit-identical to single runs")


def test_compute_dndy_matches_oracle_rhs(dev):
    """compute_dndy's columns (active parents only) against the oracle."""

    from nlo_bk_optimized_amd import MV, BKConfig, compute_dndy
    from nlo_bk_optimized_amd.bk import evolution as evo

    cfg = BKConfig(RPOINTS=23, MINR=1e-6, MAXR=1e2, MCINTPOINTS=4096, K1_RULE="fixed", PRECISION="fp64_stable",
                   VEGAS_ADAPT_ITERATIONS=2)
    ic = MV(qs0sqr=0.3, gamma=1.1, ec=2.0)
    res = compute_dndy(ic, cfg, device="cuda", dtype=torch.float64, seed=5)
    setup = evo.prepare(ic, cfg, device=torch.device("cuda"), seed=5)
    r, N0, edges = (x.cpu().numpy() for x in (setup.r, setup.N0, setup.edges))
    k1, k2kf = oracle_rhs(r, N0, edges, O.sobol_table(4, 5), 4096, bk_params(True))
    active = (np.arange(r.size) >= 1) & (N0 <= 0.99999)
    np.testing.assert_array_equal(res.r.cpu().numpy(), r[active])
    np.testing.assert_array_equal(res.N.cpu().numpy(), N0[active])
    mag = O.k1_fixed(r[active], O.evolution_splines(r, N0)[0], p=bk_params(True), magnitude=True)
    assert np.all(np.abs(res.K1.cpu().numpy() - k1[active]) <= 1e-10 * np.abs(k1[active]) + 1e-13 * mag)
    np.testing.assert_allclose(res.K2_Kf.cpu().numpy(), k2kf[active], rtol=1e-10, atol=1e-300)
    # float32 output as the reference's production dtype
    res32 = compute_dndy(ic, cfg, device="cuda", dtype=torch.float32, seed=5)
    assert res32.K1.dtype == torch.float32 and res32.r.shape == res.r.shape
