"""Edge cases of the BK path through the C ABI against the oracle: the
smallest grid (R = 3: one interior parent), the largest (R = 2048, the LDS
limit), sample counts and Sobol offsets that are not multiples of a wave or
a block, a single sample, saturated parents in the middle of the grid, and a
batch of one run.  Tolerances as in test_gpu_parity.py."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.bk_helpers import bk_params, oracle_evolve, random_edges, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nlo_bk_optimized_amd import device

    return device


def cuda(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _config(**kw):
    from nlo_bk_optimized_amd.bk.config import BKConfig

    return BKConfig(MINR=1e-6, **kw)


def _k1_check(dev, r, N, idx):
    """Fixed and adaptive K1 on every node (GPU) against the oracle at idx."""

    rt, Nt = cuda(r), cuda(N)
    t64, _ = dev.spline_build(rt, Nt)
    cfg = _config()
    k1 = dev.bk_k1(rt, Nt, t64, cfg).cpu().numpy()
    stats = {}
    k1a = dev.bk_k1(rt, Nt, t64, cfg, rule="adaptive", stats=stats).cpu().numpy()
    sN, _ = O.evolution_splines(r, N)
    active = (np.arange(r.size) >= 1) & (N <= 0.99999)
    assert k1[0] == 0.0 and np.all(k1[~active] == 0.0) and np.all(k1a[~active] == 0.0)
    idx = np.array([i for i in idx if active[i]])
    ref = O.k1_fixed(r[idx], sN, p=bk_params())
    mag = O.k1_fixed(r[idx], sN, p=bk_params(), magnitude=True)
    assert np.all(np.abs(k1[idx] - ref) <= 1e-10 * np.abs(ref) + 1e-13 * mag), (k1[idx], ref)
    refa, erra, _ = O.k1_adaptive(r[idx], sN, p=bk_params())
    err = stats["error"].cpu().numpy()[idx]
    assert np.all(np.abs(k1a[idx] - refa) <= 1e-10 * np.abs(refa) + err + erra + 1e-13 * mag)


def test_smallest_grid(dev):
    """R = 3: node 0 is never a parent, so one or two parents."""

    print("R=3 grid", flush=True)
    r, N = workload(R=3)
    _k1_check(dev, r, N, [1, 2])
    n, bins = 1000, 50
    edges = random_edges(3, bins, seed=2)
    rt, Nt = cuda(r), cuda(N)
    t64, t32 = dev.spline_build(rt, Nt)
    table = dev.sobol_table(4, 0)
    sums, _, _ = dev.bk_nlo(rt, Nt, t64, t32, cuda(edges), table, n, _config(), precision="fp64_stable")
    sN, _ = O.evolution_splines(r, N)
    lo, width = O.nlo_bounds(r)
    active = np.nonzero((np.arange(3) >= 1) & (N <= 0.99999))[0]
    ref = O.bk_nlo_sums(r[active], sN, edges[active], lo, width, O.sobol_table(4, 0), 0, n, bk_params(True))[0]
    s = sums.cpu().numpy()
    np.testing.assert_allclose(s[active, 0], ref[:, 0], rtol=1e-10, atol=0)
    assert np.all(s[0] == 0.0)
    # one RK2 step of the evolver
    ev = dev.Evolver(rt, cuda(edges), table, _config(), n_samples=n, precision="fp64_stable", rk_method="rk2",
                     step=0.1)
    try:
        ev.state.copy_(Nt)
        ev.step()
        got = ev.state.cpu().numpy()
    finally:
        ev.close()
    want = oracle_evolve(r, N, edges, O.sobol_table(4, 0), n, bk_params(True), steps=1, h=0.1, method="rk2")[-1]
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-18)


@pytest.mark.parametrize("first,n", [(0, 1), (0, 63), (64 * 5, 1000), (0, 4096 + 13), (64 * 37, 300)])
@pytest.mark.parametrize("precision", ["fp64_stable", "fp32"])
def test_ragged_sample_counts(dev, precision, first, n):
    """Sample ranges that do not fill a wave, a block or a chunk; a Sobol
    offset first_index (a multiple of 64)."""

    print(f"ragged n={n} first={first} {precision}", flush=True)
    R, bins = 12, 50
    r, N = workload(R=R)
    edges = random_edges(R, bins, seed=3)
    rt, Nt = cuda(r), cuda(N)
    t64, t32 = dev.spline_build(rt, Nt)
    table = dev.sobol_table(4, 1)
    sums, _, _ = dev.bk_nlo(rt, Nt, t64, t32, cuda(edges), table, n, _config(), first_index=first,
                            precision=precision)
    sN, _ = O.evolution_splines(r, N)
    lo, width = O.nlo_bounds(r)
    active = np.nonzero((np.arange(R) >= 1) & (N <= 0.99999))[0]
    ref = O.bk_nlo_sums(r[active], sN, edges[active], lo, width, O.sobol_table(4, 1), first, n,
                        bk_params(True))[0]
    s = sums.cpu().numpy()[active]
    if precision == "fp64_stable":
        np.testing.assert_allclose(s[:, 0], ref[:, 0], rtol=1e-10, atol=1e-300)
        np.testing.assert_allclose(s[:, 1], ref[:, 1], rtol=1e-10, atol=1e-300)
    else:
        # 2e-5 of the sum holds for sums over many samples; a handful of samples
        # keep single-sample FP32 errors (up to ~1e-4 where the terms cancel)
        scale = np.maximum(np.abs(ref[:, 0]), 1e-3 * np.sqrt(ref[:, 1] * n))
        tol = 2e-5 * scale if n >= 256 else 5e-4 * np.sqrt(ref[:, 1])
        err = np.abs(s[:, 0] - ref[:, 0])
        print("fp32 error / tolerance:", np.max(err / (tol + 1e-300)), flush=True)
        assert np.all(err <= tol + 1e-300)


def test_largest_grid(dev):
    """R = 2048 (the LDS limit of the K1 and K2 + Kf kernels): K1 on both rules
    and the FP64 K2 + Kf sums at a spread of parents."""

    print("R=2048 grid", flush=True)
    R = 2048
    r, N = workload(R=R, minr=1e-7, maxr=1e2)
    idx = list(range(1, R, 293))
    _k1_check(dev, r, N, idx)
    n, bins = 2048, 50
    edges = random_edges(R, bins, seed=4)
    rt, Nt = cuda(r), cuda(N)
    t64, t32 = dev.spline_build(rt, Nt)
    table = dev.sobol_table(4, 0)
    sums, _, _ = dev.bk_nlo(rt, Nt, t64, t32, cuda(edges), table, n, _config(), precision="fp64_stable")
    sN, _ = O.evolution_splines(r, N)
    lo, width = O.nlo_bounds(r)
    active = np.array([i for i in idx if N[i] <= 0.99999])
    ref = O.bk_nlo_sums(r[active], sN, edges[active], lo, width, O.sobol_table(4, 0), 0, n, bk_params(True))[0]
    np.testing.assert_allclose(sums.cpu().numpy()[active, 0], ref[:, 0], rtol=1e-10, atol=0)


def test_saturated_parents_in_the_middle(dev):
    """A bump of N above 0.99999 inside the grid: those parents get K1 = 0 and
    K2 + Kf = 0 (the active set is index >= 1 and N <= 0.99999, per node)."""

    print("saturated bump", flush=True)
    R, bins, n = 16, 50, 1024
    r, N = workload(R=R)
    N = N.copy()
    N[5:8] = 0.999995
    edges = random_edges(R, bins, seed=6)
    rt, Nt = cuda(r), cuda(N)
    t64, t32 = dev.spline_build(rt, Nt)
    k1 = dev.bk_k1(rt, Nt, t64, _config()).cpu().numpy()
    sums, _, _ = dev.bk_nlo(rt, Nt, t64, t32, cuda(edges), dev.sobol_table(4, 0), n, _config(),
                            precision="fp32")
    assert np.all(k1[5:8] == 0.0) and np.all(sums.cpu().numpy()[5:8] == 0.0)
    active = (np.arange(R) >= 1) & (N <= 0.99999)
    assert np.all(k1[active] != 0.0)


def test_batch_of_one_run(dev):
    from nlo_bk_optimized_amd import MV, solve_bk, solve_bk_batch
    from nlo_bk_optimized_amd.bk.config import BKConfig

    print("batch of one", flush=True)
    ic = MV(qs0sqr=0.3, x0=0.01, ec=2.0, lambdaqcd=0.241, gamma=1.1)
    cfg = BKConfig(RPOINTS=20, MINR=1e-6, MAXR=1e2, MCINTPOINTS=1 << 11, DE_SOLVER_STEP=0.1, RK_METHOD="rk2",
                   K1_RULE="fixed")
    (t,) = solve_bk_batch([ic], 0.3, cfg, seed=0)
    single = solve_bk(ic, 0.3, cfg, seed=0)
    # equal up to the grouping of the K2 + Kf sums (the batch tunes its launch shape)
    torch.testing.assert_close(t.N, single.N, rtol=1e-6, atol=1e-12)
