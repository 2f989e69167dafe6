"""The batched driver (many evolutions of uneven grid sizes per launch)
against the single-run driver on the same runs, through the C ABI.

Each run's right-hand side is the single-run one; only the split of its
samples into per-block partial sums may differ (samples per lane), so the
batched evolution equals the individual ones to rounding (1e-12 relative in
FP64, 1e-6 absolute in FP32).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests.bk_helpers import random_edges, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nlo_bk_optimized_amd import device

    return device


def cuda(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


RUNS = ((17, 0.2), (24, 0.35), (33, 0.12))  # (R, Qs0^2)


def _runs():
    out = []
    for k, (R, q) in enumerate(RUNS):
        r, N = workload(R=R, qs0sqr=q)
        out.append((r, N, random_edges(R, 50, seed=11 + k)))
    return out


@pytest.mark.parametrize("precision,spt,grid", [("fp64_stable", 16, 64), ("fp32", 8, 0), ("fp64", 0, 0)])
def test_batch_matches_single_runs(dev, precision, spt, grid):
    from nlo_bk_optimized_amd.bk.config import BKConfig

    n, steps, h = 4096, 2, 0.1
    cfg = BKConfig(MINR=1e-6)
    table = dev.sobol_table(4, 0)
    runs = _runs()
    r = cuda(np.concatenate([x[0] for x in runs]))
    edges = cuda(np.concatenate([x[2] for x in runs]))
    N0 = np.concatenate([x[1] for x in runs])
    bev = dev.BatchEvolver(r, edges, [x[0].size for x in runs], table, cfg, n_samples=n, precision=precision,
                           rk_method="rk2", step=h, spt=spt, grid_blocks=grid)
    bev.state.copy_(cuda(N0))
    for _ in range(steps):
        bev.step()
    got = bev.state.cpu().numpy()
    tuned = bev.tuning()
    evals = int(bev.active_evaluations.item())
    bev.close()
    print(precision, "batch launch shape (spt, grid, ms):", tuned)
    assert tuned[0] in (4, 8, 16) and tuned[1] >= 1
    if spt:
        assert tuned[0] == spt
    ref, ref_evals, off = [], 0, 0
    for rr, NN, ee in runs:
        ev = dev.Evolver(cuda(rr), cuda(ee), table, cfg, n_samples=n, precision=precision, rk_method="rk2",
                         step=h)
        ev.state.copy_(cuda(NN))
        for _ in range(steps):
            ev.step()
        ref.append(ev.state.cpu().numpy())
        ref_evals += int(ev.active_evaluations.item())
        ev.close()
    ref = np.concatenate(ref)
    assert evals == ref_evals
    if precision == "fp32":
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)
    else:
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-300)


def test_solve_bk_batch_matches_solve_bk(dev):
    from nlo_bk_optimized_amd import MV, BKConfig, solve_bk, solve_bk_batch

    base = dict(MINR=1e-6, MAXR=1e2, MCINTPOINTS=1 << 12, DE_SOLVER_STEP=0.1, VEGAS_ADAPT_ITERATIONS=1,
                RK_METHOD="rk2", PRECISION="fp64_stable")
    ics = [MV(qs0sqr=0.2, gamma=1.0, ec=1.0), MV(qs0sqr=0.4, gamma=1.1, ec=5.0)]
    cfgs = [BKConfig(RPOINTS=20, **base), BKConfig(RPOINTS=31, **base)]
    stats: dict = {}
    tables = solve_bk_batch(ics, 0.2, cfgs, device="cuda", dtype=torch.float64, seed=3, stats=stats)
    print("solve_bk_batch stats:", stats)
    for ic, c, t in zip(ics, cfgs, tables):
        ref = solve_bk(ic, 0.2, c, device="cuda", dtype=torch.float64, seed=3)
        np.testing.assert_array_equal(t.r.cpu().numpy(), ref.r.cpu().numpy())
        np.testing.assert_array_equal(t.y.cpu().numpy(), ref.y.cpu().numpy())
        np.testing.assert_allclose(t.N.cpu().numpy(), ref.N.cpu().numpy(), rtol=1e-12, atol=1e-300)
    with pytest.raises(ValueError, match="share"):
        solve_bk_batch(ics, 0.2, [cfgs[0], BKConfig(RPOINTS=31, **{**base, "MCINTPOINTS": 1 << 13})])
