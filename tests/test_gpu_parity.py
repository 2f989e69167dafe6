"""HIP path (through the C ABI) against the oracle and torch's SobolEngine.

Tolerances: integer work (Sobol states, VEGAS bins, counts) and the VEGAS
grid refinement are bit-exact; FP64 sums within 1e-10 relative (the
reference K2/Kf formula: 1e-10 where it is well conditioned, else within
1e-18 of its integrated rounding scale); the FP32 integrand within 2e-5
relative of the FP64 evaluation of the same (cancellation-free) formula.
"""

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


# ------------------------------------------------------------------ Sobol


@pytest.mark.parametrize("dim,seed", [(4, 0), (5, 7), (1, 3)])
@pytest.mark.parametrize("first,n", [(0, 1), (0, 4096 + 13), (37, 1000), (4096, 2048)])
def test_sobol_bit_exact_with_sobolengine(dev, dim, seed, first, n):
    print(f"sobol dim={dim} seed={seed} first={first} n={n}", flush=True)
    table = dev.sobol_table(dim, seed)
    for kind, dtype in (("f32", torch.float32), ("f64", torch.float64)):
        eng = torch.quasirandom.SobolEngine(dim, scramble=True, seed=seed)
        ref = eng.draw(first + n, dtype=dtype)[first:]
        got = dev.sobol_draw(table, n, first_index=first, kind=kind).cpu()
        assert torch.equal(got, ref), kind
    q = dev.sobol_draw(table, n, first_index=first, kind="u32").cpu().numpy().view(np.uint32)
    ot = O.sobol_table(dim, seed)
    np.testing.assert_array_equal(q, O.sobol_u32(ot, first, n))


# ------------------------------------------------------------------ VEGAS


@pytest.mark.parametrize("dim", [4, 5])  # 4: the vector load / store path
def test_vegas_map_bins_bit_exact(dev, dim):
    bins, n = 50, 20000
    table = dev.sobol_table(dim, 11)
    u = dev.sobol_draw(table, n, kind="f64")
    edges = random_edges(1, bins, seed=3)[0, :4]
    edges = np.concatenate([edges, edges[:1]], axis=0)[:dim]
    lo = np.array([-3.0, 0.5, 0.0, 1.0, -1.0])[:dim]
    width = np.array([6.0, 2.0, 6.283185307179586, 0.5, 4.0])[:dim]
    x, jac, b = dev.vegas_map(u, cuda(edges), cuda(lo), cuda(width))
    q = np.round(u.cpu().numpy() * 2.0**30).astype(np.uint32)
    ox, ojac, ob = O.vegas_map(q, edges, lo, width)
    np.testing.assert_array_equal(b.cpu().numpy().astype(np.int32), ob)
    np.testing.assert_array_equal(x.cpu().numpy(), ox)
    np.testing.assert_array_equal(jac.cpu().numpy(), ojac)


def test_vegas_map_uniform_grid_is_identity(dev):
    # S/dis/fixed_sobol.py map_importance_grid on uniform edges: x = u, jac = 1
    # (the identity pinned by T/tests/test_dis_production.py:139-143).
    n, dim, bins = 4096, 5, 32
    table = dev.sobol_table(dim, 2)
    u = dev.sobol_draw(table, n, kind="f64")
    edges = torch.linspace(0, 1, bins + 1, dtype=torch.float64, device="cuda").expand(dim, -1).contiguous()
    x, jac, _ = dev.vegas_map(u, edges, torch.zeros(dim, dtype=torch.float64, device="cuda"),
                              torch.ones(dim, dtype=torch.float64, device="cuda"))
    torch.testing.assert_close(x, u, rtol=0, atol=1e-15)
    torch.testing.assert_close(jac, torch.ones_like(jac), rtol=1e-13, atol=0)


@pytest.mark.parametrize("n", [1, 4095, 4096 * 5 + 17])
def test_vegas_histogram_matches_bincount(dev, n):
    dim, bins = 4, 50
    rng = np.random.default_rng(n)
    b = rng.integers(0, bins, (n, dim)).astype(np.uint8)
    w = rng.exponential(1.0, n) * np.exp(rng.uniform(-20, 5, n))
    hist, count = dev.vegas_histogram(cuda(b, torch.uint8), cuda(w), bins)
    for d in range(dim):
        np.testing.assert_array_equal(count[d].cpu().numpy(), np.bincount(b[:, d], minlength=bins))
        ref = np.bincount(b[:, d], weights=w, minlength=bins)
        np.testing.assert_allclose(hist[d].cpu().numpy(), ref, rtol=1e-12, atol=1e-14 * w.sum())


def test_vegas_adapt_bit_exact_with_reference(dev, golden):
    for bins in (50, 32):
        edges = cuda(golden[f"adapt{bins}_edges"])
        new = dev.vegas_adapt(edges, cuda(golden[f"adapt{bins}_hist"]), cuda(golden[f"adapt{bins}_count"]),
                              float(np.finfo(np.float64).eps))
        np.testing.assert_array_equal(new.cpu().numpy(), golden[f"adapt{bins}_new"])


# ------------------------------------------------------------------ BK


def test_spline_tables_match_oracle(dev):
    r, N = workload(R=64)
    t64, t32 = dev.spline_build(cuda(r), cuda(N))
    sN, sS = O.evolution_splines(r, N)
    R = r.size
    t = t64.cpu().numpy()
    assert t[0] == r[0] and t[1] == r[-1]
    np.testing.assert_allclose(t[2 : 2 + R], sN.x, rtol=1e-15)
    for row, spline in ((0, sN), (1, sS)):
        rec = t[2 + R + row * 4 * (R - 1) : 2 + R + (row + 1) * 4 * (R - 1)].reshape(R - 1, 4).T
        ref = spline.coef.reshape(4, R - 1)
        scale = np.abs(ref).max(axis=1, keepdims=True)
        np.testing.assert_allclose(rec, ref, rtol=0, atol=1e-12 * scale.max())
    # FP32 table: local coordinate t = (ln r - ln r_0)/h - i, base-2 exponent
    t3 = t32.cpu().numpy().astype(np.float64)
    np.testing.assert_array_equal(t3[:2 + R], t[:2 + R].astype(np.float32))
    hav = (sN.x[-1] - sN.x[0]) / (R - 1)
    scale = np.log2(np.e) * hav ** np.arange(4)
    for row in (0, 1):
        o = 2 + R + row * 4 * (R - 1)
        rec = t3[o : o + 4 * (R - 1)].reshape(R - 1, 4)
        ref = t[o : o + 4 * (R - 1)].reshape(R - 1, 4) * scale
        np.testing.assert_allclose(rec, ref, rtol=1e-7, atol=1e-30)
        # N at the nodes: 2^(a') reproduces the node values
    a0 = t3[2 + R : 2 + R + 4 * (R - 1) : 4]
    # log2 N ~ -40 at the smallest r: float32 rounding of the exponent gives 2e-6
    np.testing.assert_allclose(2.0 ** a0[1:], np.clip(N[1:-1], 0, 1), rtol=3e-6)


def test_k1_fp64_matches_oracle(dev):
    from nlo_bk_optimized_amd.bk.config import BKConfig

    r, N = workload(R=24)
    config = BKConfig(MINR=1e-6)
    rt, Nt = cuda(r), cuda(N)
    t64, _ = dev.spline_build(rt, Nt)
    k1 = dev.bk_k1(rt, Nt, t64, config).cpu().numpy()
    sN, _ = O.evolution_splines(r, N)
    active = (np.arange(r.size) >= 1) & (N <= 0.99999)
    ref = np.zeros(r.size)
    mag = np.zeros(r.size)
    ref[active] = O.k1_fixed(r[active], sN, p=bk_params())
    mag[active] = O.k1_fixed(r[active], sN, p=bk_params(), magnitude=True)
    assert k1[0] == 0.0 and np.all(k1[~active] == 0.0)
    # 1e-10 relative, or 1e-13 of the integrated |integrand| where K1 cancels
    assert np.all(np.abs(k1 - ref) <= 1e-10 * np.abs(ref) + 1e-13 * mag), (k1, ref)


def _nlo_case(dev, precision, histogram=False, R=12, n=4096, seed=0, minr=1e-6):
    from nlo_bk_optimized_amd.bk.config import BKConfig

    r, N = workload(R=R, minr=minr)
    config = BKConfig(MINR=1e-6)
    bins = 50
    edges = random_edges(R, bins, seed=seed)
    rt, Nt = cuda(r), cuda(N)
    t64, t32 = dev.spline_build(rt, Nt)
    table = dev.sobol_table(4, seed)
    sums, hist, count = dev.bk_nlo(rt, Nt, t64, t32, cuda(edges), table, n, config,
                                   precision=precision, histogram=histogram)
    sN, sS = O.evolution_splines(r, N)
    lo, width = O.nlo_bounds(r)
    ot = O.sobol_table(4, seed)
    active = np.nonzero((np.arange(R) >= 1) & (N <= 0.99999))[0]
    out = {}
    # "direct": reference integrand on the S spline; "cos": the same with the
    # torch code's X^2 = r^2 + z^2 - 2 r z cos(theta) (their spread measures
    # the reference formula's rounding sensitivity); "stable": N spline.
    for key, spline, prm in (("direct", sS, bk_params(False)),
                             ("stable", sN, bk_params(True))):
        out[key] = O.bk_nlo_sums(r[active], spline, edges[active], lo, width, ot, 0, n, prm,
                                 want_hist=histogram, want_magnitude=True)
    return active, sums.cpu().numpy(), (None if hist is None else hist.cpu().numpy()), \
        (None if count is None else count.cpu().numpy()), out


def test_nlo_fp64_reference_formula_matches_oracle(dev):
    # r >= 1e-2: the reference K2/Kf formula is well conditioned -> 1e-10
    active, sums, _, _, o = _nlo_case(dev, "fp64", R=16, minr=1e-2)
    ref, _, _, mag = o["direct"]
    # 1e-10 relative, plus one unit of rounding of the integrated terms
    rel = np.abs(sums[active, 0] - ref[:, 0]) / np.abs(ref[:, 0])
    assert np.all(np.abs(sums[active, 0] - ref[:, 0]) <= 1e-10 * np.abs(ref[:, 0]) + 2.2e-16 * mag)
    assert np.median(rel) <= 1e-12, rel
    np.testing.assert_allclose(sums[active, 1], ref[:, 1], rtol=1e-10, atol=0)


def test_nlo_fp64_reference_formula_within_its_rounding_scale(dev):
    # down to r = 1e-6 the formula's terms cancel by up to 1e19 (sum of
    # |jac| x the rounding scale of every term, computed by the oracle);
    # GPU and oracle must then agree to 1e-18 of that scale (1e-10 relative
    # where the scale allows)
    active, sums, _, _, o = _nlo_case(dev, "fp64", R=24)
    ref, _, _, mag = o["direct"]
    got = sums[active, 0]
    assert np.all(np.abs(got - ref[:, 0]) <= 1e-10 * np.abs(ref[:, 0]) + 1e-18 * mag), (got, ref[:, 0], mag)


def test_nlo_fp64_stable_matches_oracle(dev):
    active, sums, _, _, o = _nlo_case(dev, "fp64_stable", R=24)
    ref = o["stable"][0]
    np.testing.assert_allclose(sums[active, 0], ref[:, 0], rtol=1e-10, atol=0)
    np.testing.assert_allclose(sums[active, 1], ref[:, 1], rtol=1e-10, atol=0)
    # component sums (K2, Kf) of cancelling terms: relative to sqrt(n sum w^2)
    scale = np.sqrt(ref[:, 1] * 4096)
    for k in (2, 3):
        assert np.all(np.abs(sums[active, k] - ref[:, k]) <= 1e-10 * np.maximum(np.abs(ref[:, k]), scale))


def test_nlo_fp32_within_tolerance_of_fp64_oracle(dev):
    active, sums, _, _, o = _nlo_case(dev, "fp32", R=24)
    ref = o["stable"][0]
    scale = np.maximum(np.abs(ref[:, 0]), 1e-3 * np.sqrt(ref[:, 1] * 4096))
    err = np.abs(sums[active, 0] - ref[:, 0]) / scale
    print("fp32 vs fp64 oracle, relative sum error per parent:", err)
    assert np.all(err <= 2e-5), err


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp64_stable"])
def test_nlo_sum_only_equals_first_column(dev, precision):
    """NLO_BK_SUM_ONLY (the evolver's instance) forms sum w exactly as the full
    instance does (same order): bit for bit, the other columns 0."""

    from nlo_bk_optimized_amd.bk.config import BKConfig

    R, n = 24, 4096 + 640
    r, N = workload(R=R)
    edges = cuda(random_edges(R, 50, seed=5))
    rt, Nt = cuda(r), cuda(N)
    t64, t32 = dev.spline_build(rt, Nt)
    table = dev.sobol_table(4, 2)
    print(f"nlo sum only {precision}", flush=True)
    full, _, _ = dev.bk_nlo(rt, Nt, t64, t32, edges, table, n, BKConfig(MINR=1e-6), precision=precision)
    only, _, _ = dev.bk_nlo(rt, Nt, t64, t32, edges, table, n, BKConfig(MINR=1e-6), precision=precision,
                            sum_only=True)
    assert torch.equal(only[:, 0], full[:, 0])
    assert torch.count_nonzero(only[:, 1:]) == 0
    assert torch.count_nonzero(full[:, 0]) > 0


def test_nlo_histogram_matches_oracle(dev):
    active, sums, hist, count, o = _nlo_case(dev, "fp64_stable", histogram=True, R=24)
    _, oh, oc, _ = o["stable"]
    np.testing.assert_array_equal(count[active], oc)
    tot = oh.sum(axis=2, keepdims=True)
    # fixed-point block sums: each term rounded to 2^-51 of its block's largest |w|
    assert np.all(np.abs(hist[active] - oh) <= 1e-10 * tot + 1e-300)
    assert np.all(hist[0] == 0) and np.all(count[0] == 0)


def test_evolver_matches_oracle_rk2(dev):
    from nlo_bk_optimized_amd.bk.config import BKConfig

    R, n, bins, steps, h = 10, 2048, 50, 2, 0.1
    r, N0 = workload(R=R)
    edges = random_edges(R, bins, seed=5)
    config = BKConfig(MINR=1e-6)
    table = dev.sobol_table(4, 0)
    ev = dev.Evolver(cuda(r), cuda(edges), table, config, n_samples=n, precision="fp64_stable",
                     rk_method="rk2", step=h)
    ev.state.copy_(cuda(N0))
    got = [N0]
    for _ in range(steps):
        ev.step()
        got.append(ev.state.cpu().numpy().copy())
    ev.close()
    ref = oracle_evolve(r, N0, edges, O.sobol_table(4, 0), n, bk_params(True), steps=steps, h=h,
                        method="rk2")
    np.testing.assert_allclose(np.array(got), ref, rtol=1e-10, atol=1e-18)


def test_solve_bk_fp32_vs_fp64(dev):
    from nlo_bk_optimized_amd import MV, BKConfig, solve_bk

    base = dict(RPOINTS=31, MINR=1e-6, MAXR=1e2, MCINTPOINTS=4096, DE_SOLVER_STEP=0.1,
                RK_METHOD="rk2", VEGAS_ADAPT_ITERATIONS=2)
    ic = MV(qs0sqr=0.2, gamma=1.0)
    t32 = solve_bk(ic, 0.3, BKConfig(PRECISION="fp32", **base), dtype=torch.float64, seed=0)
    t64 = solve_bk(ic, 0.3, BKConfig(PRECISION="fp64_stable", **base), dtype=torch.float64, seed=0)
    assert t32.N.shape == (4, 32)
    torch.testing.assert_close(t32.y.cpu(), torch.tensor([0.0, 0.1, 0.2, 0.30000000000000004], dtype=torch.float64))
    dmax = (t32.N - t64.N).abs().max().item()
    print(f"max |dN| fp32 vs fp64_stable: {dmax:.3e}")
    assert dmax <= 1e-5
    assert bool(torch.all((t64.N >= 0) & (t64.N <= 1)))


def test_largest_workload_shape_one_rk4_step(dev):
    """The largest configuration's shape (R = 1024 on [1e-7, 1e2], 2^24 Sobol
    points per parent per right-hand side, classical RK4, dY = 0.05) for one
    step: FP32 against the FP64 evaluation of the same formula at the full
    size, where the oracle would take hours (size-independent checks: the
    step stays in [0, 1] and small, and both precisions agree)."""

    from nlo_bk_optimized_amd import MV
    from nlo_bk_optimized_amd.bk.config import BKConfig
    from nlo_bk_optimized_amd.bk import evolution as evo

    base = dict(RPOINTS=1023, MINR=1e-7, MAXR=1e2, MCINTPOINTS=1 << 24, DE_SOLVER_STEP=0.05, RK_METHOD="rk4",
                VEGAS_ADAPT_ITERATIONS=1, VEGAS_ADAPT_POINTS=1 << 16)
    ic = MV(qs0sqr=0.2, gamma=1.0)
    out = {}
    for precision in ("fp32", "fp64_stable"):
        config = BKConfig(PRECISION=precision, **base)
        setup = evo.prepare(ic, config, device=torch.device("cuda"), seed=0)
        ev = dev.Evolver(setup.r, setup.edges, setup.table, config, n_samples=1 << 24, precision=precision,
                         rk_method="rk4", step=0.05)
        ev.state.copy_(setup.N0)
        ev.step()
        out[precision] = ev.state.cpu().numpy().copy()
        assert int(ev.flag.item()) == 0
        ev.close()
        N0 = setup.N0.cpu().numpy()
    a, b = out["fp32"], out["fp64_stable"]
    assert np.all((a >= 0) & (a <= 1)) and np.all((b >= 0) & (b <= 1))
    assert np.abs(b - N0).max() < 0.05  # one small step
    d = np.abs(a - b).max()
    k = int(np.argmax(np.abs(a - b)))
    print(f"R=1024, 2^24 points, one RK4 step: max |dN| fp32 vs fp64 = {d:.3e} at r = {setup.r[k].item():.3e}, "
          f"N = {b[k]:.6f}, step dN = {b[k] - N0[k]:.3e}")
    assert d <= 1e-5
