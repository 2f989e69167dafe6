"""Benchmark: full NLO BK evolution on one MI355X (the 256-node workload).

Workload: 256 log-spaced r in [1e-6, 1e2] GeV^-1; MV initial condition
(Qs0^2 = 0.2 GeV^2, gamma = 1, Lambda = 0.241 GeV, e_c = 1), Nf = 3, Balitsky
LO / smallest-distance NLO coupling, double- and single-log resummation;
K2 + Kf by VEGAS over 2^20 scrambled-Sobol points (seed 0) per parent per
right-hand side through 50-bin grids adapted 5 times on the initial
condition and then frozen; K1 on the 8 x 24 Gauss-Legendre rule per log-r
panel; midpoint RK2 with dY = 0.1.

A "step" is one RK2 rapidity step (two right-hand sides) over all parents,
replayed from its hipGraph; the default 100 steps are the whole evolution
Y: 0 -> 10.  ``value`` is integrand samples per second of the FP32 path
(FP32 integrand, FP64 accumulation), counting only active parents
(N <= 0.99999), with all inputs resident in HBM.  The same evolution in
plain FP64 (the reference's K2/Kf formula) and in FP64 with the FP32 path's
formula is run for the speedup and max |dN| fields.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

METRIC = "NLO-BK evolution wall time + integrand samples/s, FP32 vs FP64 at fixed error"

# Algorithmic flops per K2 + Kf sample of the FP32 path (VEGAS map, the
# cancellation-free integrand and the FP64 accumulation): the operation count
# of the formula in oracle/flop_count.py (add/sub/mul/div, reciprocal, sqrt,
# rsqrt, exp, exp2, log, log1p, expm1, sin, coth = 1, fma = 2, compares,
# selects, min/max and integer work = 0), which tests/test_flop_count.py
# checks against the C oracle by value and against this constant.
FLOPS_PER_SAMPLE = 216
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per launch of the FP32 K2 + Kf kernel on this workload, from the
# FETCH_SIZE / WRITE_SIZE counter runs calibrated on a device copy.
COUNTERS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "nlo_kernel_fp32_256r_counters.md")


def measured_traffic(rpoints: int, samples_log2: int) -> float | None:
    """traffic_bytes_per_launch from the counter summary (default workload only)."""

    if rpoints != 255 or samples_log2 != 20 or not os.path.exists(COUNTERS):
        return None
    with open(COUNTERS) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def parse() -> argparse.Namespace:
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--samples-log2", type=int, default=20)
    p.add_argument("--rpoints", type=int, default=255)
    p.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison runs")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-config1", action="store_true",
                   help="skip the 64-node plain-FP64 RK2 evolution to Y = 1 on the GPU and on the CPU oracle")
    p.add_argument("--kernel-launches", type=int, default=5)
    p.add_argument("--cpu-samples-log2", type=int, default=20,
                   help="Sobol points per parent in the CPU baseline sample")
    p.add_argument("--no-dis", action="store_true", help="skip the DIS VEGAS-pass measurements")
    p.add_argument("--no-dropin", action="store_true",
                   help="skip the end-to-end solve_bk run with the default K1_RULE (adaptive K1 at this size)")
    p.add_argument("--no-hbm", action="store_true", help="skip the Sobol / VEGAS map / histogram bandwidth runs")
    p.add_argument("--large-maxy", type=float, default=0.0,
                   help="run the 1024-node, 2^24-point RK4 evolution to this Y in FP32 and FP64 (0 skips it)")
    p.add_argument("--large-precision", default="both", choices=("both", "fp32", "fp64_stable"),
                   help="which path(s) the large run evolves")
    p.add_argument("--large-save", default=None, help="npz file for N at every unit of Y of the large run")
    p.add_argument("--large-resume", default=None, help="npz file (from --large-save) to continue from")
    p.add_argument("--batch-runs", type=int, default=64,
                   help="evolutions in the batched uneven-grid run (0 skips it)")
    p.add_argument("--dis-points", type=int, default=32,
                   help="(Q^2, x) points of the NLO DIS batch (512 = the full 32 x 16 grid; 0 skips it)")
    p.add_argument("--dis-precisions", default="fp32,fp64",
                   help="integrand precisions of the DIS batch (fp32 always runs)")
    p.add_argument("--dis-frozen-points", type=int, default=32,
                   help="points of the DIS batch whose terms are also compared FP32 vs FP64 through one frozen grid")
    p.add_argument("--kernel-probe", action="store_true",
                   help="only run the K2+Kf launch and a 256 MiB device copy (for counter runs)")
    p.add_argument("--probe-precision", default="fp32", choices=("fp32", "fp64", "fp64_stable"),
                   help="integrand precision of the --kernel-probe launches")
    p.add_argument("--probe-dis", action="store_true",
                   help="with --kernel-probe: also one NLO DIS batch (8 points, GBW dipole, light + charm, "
                        "2^22 samples per integral) in the probe precision")
    p.add_argument("--probe-fixed-sobol", action="store_true",
                   help="with --kernel-probe: also the six-term fixed-Sobol workload (GBW, Q^2 = 9, x = 1e-3, "
                        "charm, L and T), learned once and evaluated through the frozen grids "
                        "--kernel-launches times")
    return p.parse_args()


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(
            f"bench.py runs on one GPU only (got --gpus {args.gpus}): the BK right-hand side "
            "would shard over parents with one all-gather of N per RK stage, which is not "
            "implemented; run with --gpus 1",
            file=sys.stderr,
        )
        return 2

    import torch

    from nlo_bk_optimized_amd import MV, BKConfig
    from nlo_bk_optimized_amd import device as dev
    from nlo_bk_optimized_amd.bk import evolution as evo

    if not torch.cuda.is_available():
        print("no HIP device available", file=sys.stderr)
        return 1

    n = 1 << args.samples_log2
    base = dict(
        RPOINTS=args.rpoints, MINR=1e-6, MAXR=1e2, MCINTPOINTS=n, MC_BINS=50,
        VEGAS_ADAPT_ITERATIONS=5, DE_SOLVER_STEP=0.1, RK_METHOD="rk2",
    )
    ic = MV(qs0sqr=0.2, x0=0.01, ec=1.0, lambdaqcd=0.241, gamma=1.0)
    device = torch.device("cuda")

    def log(msg: str) -> None:
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    def evolve(precision: str, steps: int, warmup: int):
        # set-up: grid adaptation (5 VEGAS passes with histograms on the
        # initial condition), then the evolver (buffers and the captured step
        # graph); both are part of a full evolution's wall time
        config = BKConfig(PRECISION=precision, **base)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        setup = evo.prepare(ic, config, device=device, seed=0)
        torch.cuda.synchronize()
        t_adapt = time.perf_counter() - t0
        t0 = time.perf_counter()
        ev = dev.Evolver(setup.r, setup.edges, setup.table, config, n_samples=n, precision=precision,
                         rk_method="rk2", step=config.DE_SOLVER_STEP)
        ev.state.copy_(setup.N0)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        for _ in range(warmup):
            ev.step()
        ev.state.copy_(setup.N0)
        torch.cuda.synchronize()
        count0 = int(ev.active_evaluations.item())
        t0 = time.perf_counter()
        for _ in range(steps):
            ev.step()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        evaluations = int(ev.active_evaluations.item()) - count0
        flag = int(ev.flag.item())
        N = ev.state.clone()
        ev.close()
        return dict(setup=setup, seconds=elapsed, adapt_seconds=t_adapt, setup_seconds=t_setup,
                    evaluations=evaluations, N=N, nonfinite=flag, config=config)

    if args.kernel_probe:
        # Counter runs (rocprofv3 --pmc) need only the dominant kernel and a copy
        # of known size whose FETCH_SIZE / WRITE_SIZE calibrate the counters.
        config = BKConfig(PRECISION=args.probe_precision, **base)
        setup = evo.prepare(ic, config, device=device, seed=0)
        t64, t32 = dev.spline_build(setup.r, setup.N0)
        for _ in range(args.kernel_launches):
            dev.bk_nlo(setup.r, setup.N0, t64, t32, setup.edges, setup.table, n, config,
                       precision=args.probe_precision, sum_only=True)
        for _ in range(args.kernel_launches):  # K1 on the fixed rule, as the evolution runs it
            dev.bk_k1(setup.r, setup.N0, t64, config, rule="fixed")
        if args.probe_dis:
            from nlo_bk_optimized_amd.dipole import GBW
            from nlo_bk_optimized_amd.dis import structure_functions_batch
            from nlo_bk_optimized_amd.dis.config import DISConfig, DISOrder, Quark, QuarkType

            dt = torch.float32 if args.probe_precision == "fp32" else torch.float64
            dcfg = DISConfig(order=DISOrder.NLO, quarks=(Quark(QuarkType.LIGHT, 0.14), Quark(QuarkType.C, 1.27)),
                             maxeval=1 << 22, epsrel=1e-12)
            pts = [(torch.tensor(q2, dtype=dt), torch.tensor(xb, dtype=dt))
                   for q2 in (2.0, 30.0) for xb in (1e-5, 1e-4, 1e-3, 1e-2)]
            structure_functions_batch(pts, GBW(Qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0), dcfg, seed=0)
        if args.probe_fixed_sobol:
            from nlo_bk_optimized_amd.dipole import GBW
            from nlo_bk_optimized_amd.dis import Polarization
            from nlo_bk_optimized_amd.dis.config import DISConfig, Quark, QuarkType
            from nlo_bk_optimized_amd.dis.fixed_sobol import gbw_qg_fixed_sobol_batch

            gbw = GBW(Qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0)
            fcfg = DISConfig(quarks=(Quark(QuarkType.C, 1.4),), maxr=5.0, nf_alphas=1, cuda_nested_points=48)
            q2t, xt = torch.tensor(9.0, dtype=torch.float32), torch.tensor(1e-3, dtype=torch.float32)
            entries = [(q2t, xt, pol, fcfg.quarks[0], 3100) for pol in (Polarization.L, Polarization.T)]
            _, grids = gbw_qg_fixed_sobol_batch(entries, gbw, fcfg, return_edges=True)
            for _ in range(args.kernel_launches):
                gbw_qg_fixed_sobol_batch(entries, gbw, fcfg, grid=grids)
        src = torch.ones(1 << 25, dtype=torch.float64, device=device)  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"probe": f"nlo_kernel<{args.probe_precision}, sum only>", "launches": args.kernel_launches,
                          "copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8}), flush=True)
        return 0

    # One-time initialisation (library load, kernel attributes, the first
    # allocations of the caching allocator): a first set-up, reported apart
    # from the set-up every evolution pays.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evo.prepare(ic, BKConfig(PRECISION="fp32", **base), device=device, seed=0)
    torch.cuda.synchronize()
    first_setup_s = time.perf_counter() - t0

    log(f"fp32 evolution: {args.steps} steps, 2^{args.samples_log2} samples/parent/RHS")
    fp32 = evolve("fp32", args.steps, args.warmup)
    samples = fp32["evaluations"] * n
    value = samples / fp32["seconds"]
    log(f"fp32: {fp32['seconds']:.3f} s, {value:.4e} samples/s, adaptation {fp32['adapt_seconds']:.3f} s, "
        f"evolver set-up {fp32['setup_seconds']:.3f} s (first set-up in the process {first_setup_s:.3f} s)")

    # Dominant kernel: the K2 + Kf integrand launch as the evolver runs it
    # (sum w only) with its per-parent reduce, timed with HIP events on the
    # stream it is launched on (torch's current stream), on the state the
    # timed steps ended at (the graph replays of the timed region cannot be
    # bracketed one kernel at a time) and, for reference, at the initial
    # condition.  The GPU clock drops in the pause before these launches and
    # ramps back over ~20 ms of load (the kernel trace shows the first
    # launches after a pause up to 25 % slower than the back-to-back launches
    # of the timed steps), so 20 untimed launches precede each timed block.
    setup = fp32["setup"]
    config = fp32["config"]
    stream = torch.cuda.current_stream()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    idx = torch.arange(setup.r.numel(), device=device)

    def time_launches(N):
        t64, t32 = dev.spline_build(setup.r, N)
        active = int(((idx >= 1) & (N <= 0.99999)).sum())
        for _ in range(20):
            dev.bk_nlo(setup.r, N, t64, t32, setup.edges, setup.table, n, config, precision="fp32", sum_only=True)
        launches = max(args.kernel_launches, 10)
        start.record(stream)
        for _ in range(launches):
            dev.bk_nlo(setup.r, N, t64, t32, setup.edges, setup.table, n, config, precision="fp32", sum_only=True)
        stop.record(stream)
        stop.synchronize()
        ms = start.elapsed_time(stop) / launches
        return ms, active, FLOPS_PER_SAMPLE * active * n / (ms * 1e-3) / 1e12

    launch_ms, active_end, achieved = time_launches(fp32["N"])
    launch_ms0, active0, achieved0 = time_launches(setup.N0)
    log(f"nlo fp32 launch at the end of the timed steps: {launch_ms:.3f} ms for {active_end} parents x {n} "
        f"samples -> {achieved:.2f} TFLOP/s; at the initial condition {launch_ms0:.3f} ms for {active0} "
        f"parents -> {achieved0:.2f} TFLOP/s")
    t64, t32 = dev.spline_build(setup.r, setup.N0)

    # K1 per right-hand side on the initial condition: the fixed 8 x 24 rule
    # the evolution above uses, and the nested adaptive Gauss-Kronrod rule the
    # reference takes at this grid size (RPOINTS != 100).
    k1_ms, k1_stats = {}, {}
    for rule in ("fixed", "adaptive"):
        dev.bk_k1(setup.r, setup.N0, t64, config, rule=rule)
        start.record(stream)
        for _ in range(args.kernel_launches):
            dev.bk_k1(setup.r, setup.N0, t64, config, rule=rule, stats=k1_stats if rule == "adaptive" else None)
        stop.record(stream)
        stop.synchronize()
        k1_ms[rule] = start.elapsed_time(stop) / args.kernel_launches
    fixed_k1 = dev.bk_k1(setup.r, setup.N0, t64, config, rule="fixed")
    adaptive_k1 = dev.bk_k1(setup.r, setup.N0, t64, config, rule="adaptive")
    act = (torch.arange(setup.r.numel(), device=device) >= 1) & (setup.N0 <= 0.99999)
    k1_rel = float(((adaptive_k1 - fixed_k1).abs() / fixed_k1.abs().clamp_min(1e-300))[act].max())
    k1_evals = k1_stats["evaluations"][act].double()
    log(f"K1 per RHS: fixed {k1_ms['fixed']:.3f} ms ({active0} x {(setup.r.numel() - 1) * 8 * 24} nodes), "
        f"adaptive {k1_ms['adaptive']:.3f} ms (mean {float(k1_evals.mean()):.0f} evaluations per parent), "
        f"max rel difference {k1_rel:.2e}")

    dropin = {}
    if not args.no_dropin:
        # The drop-in call itself: solve_bk with the same configuration and
        # K1_RULE="auto", which at this grid size takes the nested adaptive K1
        # as the reference does (RPOINTS != 100), end to end (grid adaptation,
        # evolver set-up, every step, the table) to the same rapidity.
        from nlo_bk_optimized_amd import solve_bk
        from nlo_bk_optimized_amd.bk.config import resolve_k1_rule

        cfg_d = BKConfig(PRECISION="fp32", **base)
        solve_bk(ic, 0.2, cfg_d, device=device, seed=0)  # first-call costs
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solve_bk(ic, args.steps * cfg_d.DE_SOLVER_STEP, cfg_d, device=device, seed=0)
        torch.cuda.synchronize()
        dropin = {"solve_bk_dropin_fp32_s": time.perf_counter() - t0,
                  "solve_bk_dropin_k1_rule": resolve_k1_rule(cfg_d, torch.float32),
                  "solve_bk_dropin_maxy": args.steps * cfg_d.DE_SOLVER_STEP}
        log(f"solve_bk drop-in (K1 rule {dropin['solve_bk_dropin_k1_rule']}): "
            f"{dropin['solve_bk_dropin_fp32_s']:.3f} s to Y = {dropin['solve_bk_dropin_maxy']:g}")

    extra: dict = {
        **dropin,
        "k1_fixed_ms_per_rhs": k1_ms["fixed"],
        "k1_adaptive_ms_per_rhs": k1_ms["adaptive"],
        "k1_adaptive_mean_evaluations_per_parent": float(k1_evals.mean()),
        "k1_fixed_vs_adaptive_max_rel": k1_rel,
    }
    if not args.no_fp64:
        log("fp64 (reference K2/Kf formula) evolution")
        fp64 = evolve("fp64", args.steps, 0)
        log(f"fp64: {fp64['seconds']:.3f} s")
        log("fp64 (FP32 path's formula) evolution")
        fp64s = evolve("fp64_stable", args.steps, 0)
        log(f"fp64_stable: {fp64s['seconds']:.3f} s")
        extra.update({
            "fp64_samples_per_s": fp64["evaluations"] * n / fp64["seconds"],
            "fp64_evolution_s": fp64["seconds"],
            "fp64_full_evolution_s": fp64["adapt_seconds"] + fp64["setup_seconds"] + fp64["seconds"],
            "speedup_fp32_vs_fp64": fp64["seconds"] / fp32["seconds"],
            "max_abs_dN_fp32_vs_fp64_same_formula": float((fp32["N"] - fp64s["N"]).abs().max()),
            "max_abs_dN_fp32_vs_fp64_reference_formula": float((fp32["N"] - fp64["N"]).abs().max()),
            "max_abs_dN_fp64_formulas": float((fp64s["N"] - fp64["N"]).abs().max()),
        })

    if not args.no_dis:
        # DIS: one VEGAS pass (in-kernel Sobol, map, integrand, FP64 sums and
        # histogram) of the qqbar-g terms at the reference's DIS benchmark point
        # (charm, GBW Qs0^2 = 1, Q^2 = 9, x = 1e-3, maxr = 5, 32-bin grids).
        from nlo_bk_optimized_amd.dipole import GBW
        from nlo_bk_optimized_amd.dis.config import DISConfig, Quark, QuarkType

        gbw = GBW(Qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0)
        dcfg = DISConfig(quarks=(Quark(QuarkType.C, 1.4),), maxr=5.0, nf_alphas=1)
        desc = dev.dis_desc(9.0, 1e-3, True, dcfg.quarks, gbw, dcfg)
        nd = 1 << 22
        for kind, dim in (("I1", 5), ("I2_nested", 5), ("I3_nested", 5)):
            edges_d = torch.linspace(0, 1, 33, dtype=torch.float64, device=device).expand(dim, -1).contiguous()
            table_d = dev.sobol_table(dim, 0)
            for prec in ("fp32", "fp64"):
                dev.dis_vegas(desc, kind, edges_d, table_d, nd, precision=prec)
                st = torch.cuda.current_stream()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(st)
                reps = 3
                for _ in range(reps):
                    dev.dis_vegas(desc, kind, edges_d, table_d, nd, precision=prec)
                e1.record(st)
                e1.synchronize()
                ms = e0.elapsed_time(e1) / reps
                extra[f"dis_{kind}_T_{prec}_samples_per_s"] = nd / (ms * 1e-3)
                log(f"dis {kind} T {prec}: {ms:.3f} ms per 2^22-sample VEGAS pass")

        # The reference's six-term fixed-Sobol workload at the same point
        # (T/benchmarks/dis_fixed_sobol_all_qg.py): I1, I2 and I3 for L and T,
        # seed 3100, 3 x 65 536 training and 7 x 32 768 replicate points per
        # polarisation, FP32.  First use (grid learning and replicates) and the
        # frozen-grid evaluation (the replicates through the learned grids), each
        # the median of 7 runs as the reference times them, against its
        # 20-worker C++ total of 3.625 s; the six values against the C++ ones.
        from nlo_bk_optimized_amd.dis import Polarization
        from nlo_bk_optimized_amd.dis.fixed_sobol import gbw_qg_fixed_sobol_batch

        fcfg = DISConfig(quarks=(Quark(QuarkType.C, 1.4),), maxr=5.0, nf_alphas=1, cuda_nested_points=48)
        q2t, xt = torch.tensor(9.0, dtype=torch.float32), torch.tensor(1e-3, dtype=torch.float32)
        entries = [(q2t, xt, pol, fcfg.quarks[0], 3100) for pol in (Polarization.L, Polarization.T)]
        fixed, grids = gbw_qg_fixed_sobol_batch(entries, gbw, fcfg, return_edges=True)
        times = {"first_use": [], "frozen_grid": []}
        for _ in range(7):
            for key, grid in (("first_use", None), ("frozen_grid", grids)):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                gbw_qg_fixed_sobol_batch(entries, gbw, fcfg, grid=grid)
                torch.cuda.synchronize()
                times[key].append(time.perf_counter() - t0)
        cpp = {"L": ((0.10206739088678346, 0.0002535000977528945), (-0.024439455902300525, 4.764427185133986e-05),
                     (0.009414936380564324, 2.6852425497947778e-05)),
               "T": ((0.40263591173620866, 0.0017956984452001683), (-0.08719375357937224, 0.00026850014906050766),
                     (0.04848797141670969, 0.00013338145018266037))}
        pulls = [(float(c.value) - v) / math.hypot(float(c.error), e)
                 for pol, res in zip(("L", "T"), fixed) for c, (v, e) in zip(res.contributions, cpp[pol])]
        med = {k: sorted(v)[len(v) // 2] for k, v in times.items()}
        extra.update({
            "fixed_sobol_six_terms_first_use_s": med["first_use"],
            "fixed_sobol_six_terms_frozen_grid_s": med["frozen_grid"],
            "fixed_sobol_cpp20_over_frozen_grid": 3.625 / med["frozen_grid"],
            "fixed_sobol_cpp20_over_first_use": 3.625 / med["first_use"],
            "fixed_sobol_max_abs_pull_vs_cpp": max(abs(x) for x in pulls),
        })
        log(f"fixed-Sobol six terms: first use {med['first_use'] * 1e3:.2f} ms, frozen-grid evaluation "
            f"{med['frozen_grid'] * 1e3:.2f} ms ({3.625 / med['frozen_grid']:.0f}x the 20-worker C++ 3.625 s); "
            f"pulls vs C++ " + " ".join(f"{x:+.2f}" for x in pulls))

    if args.dis_points > 0:
        # NLO DIS batch: (Q^2, x) on 32 log Q^2 in [1, 100] x 16 log x in
        # [1e-5, 1e-2] (every k-th point of that grid unless --dis-points 512),
        # light (m = 0.14) + charm (m = 1.27), T and L, LO + dipole + qqbar-g,
        # 2^22 VEGAS samples per integral (all eight iterations run: epsrel is
        # below reach), N(r, Y) from the 256-node evolution taken to Y = 12 so
        # that every qqbar-g rapidity ln(W^2 z2) lies inside the table.  The
        # same batch in FP32 and FP64 gives the relative F2 / FL / sigma_r
        # differences (sigma_r at sqrt(s) = 318 GeV where y <= 1).
        import numpy as np

        from nlo_bk_optimized_amd import solve_bk
        from nlo_bk_optimized_amd.dipole import BKDipole
        from nlo_bk_optimized_amd.dis import structure_functions_batch
        from nlo_bk_optimized_amd.dis.config import DISConfig, DISOrder, Quark, QuarkType

        log("DIS batch: BK table to Y = 12 (untimed)")
        table = solve_bk(ic, 12.0, BKConfig(PRECISION="fp32", **base), device=device, seed=0)
        dipole = BKDipole(table, device=device)
        dcfg = DISConfig(order=DISOrder.NLO, quarks=(Quark(QuarkType.LIGHT, 0.14), Quark(QuarkType.C, 1.27)),
                         maxeval=1 << 22, epsrel=1e-12)
        qs_all, xs_all = np.geomspace(1.0, 100.0, 32), np.geomspace(1e-5, 1e-2, 16)
        step = max(1, int(round(math.sqrt(512 / max(1, args.dis_points)))))
        points = [(q, x) for q in qs_all[::step] for x in xs_all[::step]]
        res = {}
        for dtype, name in ((torch.float32, "fp32"), (torch.float64, "fp64")):
            if name not in args.dis_precisions:
                res[name] = res["fp32"]
                continue
            # every integral of every point (FT with seed 0, FL with seed
            # 1000, as F2 seeds them) in one device batch
            pts = [(torch.tensor(q2, dtype=dtype), torch.tensor(xb, dtype=dtype)) for q2, xb in points]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pairs = structure_functions_batch(pts, dipole, dcfg, seed=0)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            vals = [(ft.value.item(), fl.value.item(), ft.error.item(), fl.error.item()) for ft, fl in pairs]
            evals = sum(ft.n_eval + fl.n_eval for ft, fl in pairs)
            res[name] = (np.array(vals), wall, evals)
            log(f"DIS batch {name}: {len(points)} points in {wall:.2f} s ({evals:.3e} integrand evaluations)")
        v32, v64 = res["fp32"][0], res["fp64"][0]
        f2_32, f2_64 = v32[:, 0] + v32[:, 1], v64[:, 0] + v64[:, 1]
        # Integrand precision alone: one VEGAS pass of every VEGAS term (dipole
        # and qqbar-g, T and L, both quarks) through the same frozen (uniform)
        # grid at the same 2^22 Sobol points on the FP32 and the FP64 path.  (The adaptive runs above refine
        # their grids with the dtype's epsilon in the smoothing, as the
        # reference does, so they differ by about one Monte Carlo error.)
        from nlo_bk_optimized_amd import _lib

        frozen, worst = [], None
        for k, (q2, xb) in enumerate(points[:args.dis_frozen_points]):
            if k % 64 == 0 and k > 0:
                log(f"DIS frozen-grid comparison: {k} of {len(points)} points")
            for transverse in (True, False):
                kinds = (("T1", "T2") if transverse else ("ab", "cd")) + ("I1", "I2_nested", "I3_nested")
                for quark in dcfg.quarks:
                    desc = dev.dis_desc(q2, xb, transverse, (quark,), dipole, dcfg)
                    for kind in kinds:
                        dim = _lib.DIS_DIMENSIONS[kind]
                        ed = torch.linspace(0, 1, 33, dtype=torch.float64, device=device).expand(dim, -1).contiguous()
                        tb = dev.sobol_table(dim, 0)
                        s32, _, _ = dev.dis_vegas(desc, kind, ed, tb, 1 << 22, precision="fp32")
                        s64, _, _ = dev.dis_vegas(desc, kind, ed, tb, 1 << 22, precision="fp64")
                        a32, a64 = s32.cpu().numpy(), s64.cpu().numpy()
                        nn = float(1 << 22)
                        stderr = math.sqrt(max(a64[1] - a64[0] ** 2 / nn, 0.0) / (nn - 1))
                        row = (abs(a32[0] - a64[0]) / max(abs(a64[0]), 1e-300),
                               abs(a32[0] - a64[0]) / max(stderr, 1e-300))
                        frozen.append(row)
                        if worst is None or row[1] > worst[0][1]:
                            worst = (row, f"Q2={q2:.3g} x={xb:.3g} {'T' if transverse else 'L'} "
                                          f"m={quark.mass} {kind}")
        frozen = np.array(frozen).reshape(-1, 2)
        if not len(frozen):  # no frozen-grid comparison requested
            frozen = np.full((1, 2), np.nan)
            worst = (None, None)
        q2a = np.array([p[0] for p in points])
        xa = np.array([p[1] for p in points])
        yv = q2a / (xa * 318.0**2)
        ok = yv <= 1
        c = 1 - yv**2 / (1 + (1 - yv) ** 2)
        sr32, sr64 = v32[:, 0] + c * v32[:, 1], v64[:, 0] + c * v64[:, 1]
        extra.update({
            "dis_batch_points": len(points),
            "dis_batch_fp32_s": res["fp32"][1],
            "dis_batch_fp64_s": res["fp64"][1],
            "dis_batch_speedup_fp32_vs_fp64": res["fp64"][1] / res["fp32"][1],
            "dis_batch_fp32_evaluations": res["fp32"][2],
            "dis_batch_fp32_samples_per_s": res["fp32"][2] / res["fp32"][1],
            "dis_batch_fp64_samples_per_s": res["fp64"][2] / res["fp64"][1],
            "dis_frozen_grid_points": min(len(points), args.dis_frozen_points),
            "dis_max_rel_dF2_fp32_vs_fp64": float(np.max(np.abs(f2_32 - f2_64) / np.abs(f2_64))),
            "dis_max_rel_dFL_fp32_vs_fp64": float(np.max(np.abs(v32[:, 1] - v64[:, 1]) / np.abs(v64[:, 1]))),
            "dis_max_rel_dsigma_r_fp32_vs_fp64": (float(np.max(np.abs(sr32 - sr64)[ok] / np.abs(sr64[ok])))
                                                  if ok.any() else None),
            "dis_sigma_r_points": int(ok.sum()),
            "dis_median_rel_stderr_F2": float(np.median(np.hypot(v64[:, 2], v64[:, 3]) / np.abs(f2_64))),
            "dis_max_dF2_in_stderr": float(np.max(np.abs(f2_32 - f2_64) / np.hypot(v64[:, 2], v64[:, 3]))),
            "dis_frozen_grid_max_rel_diff_fp32_vs_fp64": float(frozen[:, 0].max()),
            "dis_frozen_grid_median_rel_diff_fp32_vs_fp64": float(np.median(frozen[:, 0])),
            "dis_frozen_grid_max_diff_in_stderr": float(frozen[:, 1].max()),
            "dis_frozen_grid_worst_case": worst[1],
        })
        log(f"DIS batch: max rel dF2 {extra['dis_max_rel_dF2_fp32_vs_fp64']:.2e} "
            f"({extra['dis_max_dF2_in_stderr']:.2f} MC errors), max rel dsigma_r "
            f"{extra['dis_max_rel_dsigma_r_fp32_vs_fp64']}; frozen grid, same points: max rel "
            f"{frozen[:, 0].max():.2e}, median {np.median(frozen[:, 0]):.2e}, max {frozen[:, 1].max():.3f} errors "
            f"({worst[1]})")

    if args.large_maxy > 0:
        # The large run: 1024 r in [1e-7, 1e2], 2^24 Sobol points per parent per
        # RHS, classical RK4 with dY = 0.05, FP32 integrand with FP64 sums
        # against the FP64 evaluation of the same formula (the GPU oracle of
        # the FP32 path), max |dN| at every unit of Y up to --large-maxy.  A
        # long run can be split: --large-precision runs one path only,
        # --large-save writes N at every unit of Y, --large-resume continues
        # from the last state such a file holds (the grids are adapted on the
        # initial condition again, deterministically).
        import numpy as np

        nl = 1 << 24
        large = dict(RPOINTS=1023, MINR=1e-7, MAXR=1e2, MCINTPOINTS=nl, MC_BINS=50, VEGAS_ADAPT_ITERATIONS=5,
                     DE_SOLVER_STEP=0.05, RK_METHOD="rk4")
        per_unit = int(round(1 / 0.05))
        precs = ("fp32", "fp64_stable") if args.large_precision == "both" else (args.large_precision,)
        y_start, rows_in = 0, {}
        if args.large_resume:
            with np.load(args.large_resume) as f:
                rows_in = {k: f[k] for k in f.files}
            y_start = int(rows_in["y"][-1])
        steps_l = int(round((args.large_maxy - y_start) / 0.05))
        states, times = {}, {}
        for prec in precs:
            cfg = BKConfig(PRECISION=prec, **large)
            setup_l = evo.prepare(ic, cfg, device=device, seed=0)
            ev = dev.Evolver(setup_l.r, setup_l.edges, setup_l.table, cfg, n_samples=nl, precision=prec,
                             rk_method="rk4", step=0.05)
            if y_start > 0:
                ev.state.copy_(torch.as_tensor(rows_in[prec][-1], dtype=torch.float64, device=device))
            else:
                ev.state.copy_(setup_l.N0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rows = []
            for k in range(1, steps_l + 1):
                ev.step()
                if k % per_unit == 0:
                    rows.append(ev.state.clone())
                    torch.cuda.synchronize()
                    log(f"large {prec}: Y = {y_start + k * 0.05:.0f} after {time.perf_counter() - t0:.1f} s")
            torch.cuda.synchronize()
            times[prec] = time.perf_counter() - t0
            states[prec] = rows
            ev.close()
        extra.update({"large_y_from": y_start, "large_maxy": args.large_maxy})
        for prec in precs:
            extra[f"large_{prec}_s"] = times[prec]
        if len(precs) == 2:
            dn = [float((a - b).abs().max()) for a, b in zip(states["fp32"], states["fp64_stable"])]
            extra.update({
                "large_speedup_fp32_vs_fp64_stable": times["fp64_stable"] / times["fp32"],
                "large_max_abs_dN_by_unit_Y": dn,
            })
            log(f"large: fp32 {times['fp32']:.1f} s, fp64 {times['fp64_stable']:.1f} s, max |dN| per unit Y {dn}")
        if args.large_save:
            out = {prec: np.stack([r.cpu().numpy() for r in states[prec]]) for prec in precs}
            out["y"] = np.arange(y_start + 1, y_start + 1 + len(states[precs[0]]))
            np.savez(args.large_save, **out)

    if args.batch_runs > 0:
        # Batched uneven evolutions: runs with Qs0^2 ~ U[0.1, 0.5], gamma ~
        # U[0.9, 1.2], e_c ~ U[1, 20], R ~ U{128..512} (numpy default_rng seed
        # 20261202), 2^18 Sobol points per parent, midpoint RK2 to Y = 8 with
        # dY = 0.1, FP32 integrand; all runs in one set of launches per stage
        # (launch shape tuned during the warmup steps), against the same runs
        # one after another with the single-run driver.
        import numpy as np

        rng = np.random.default_rng(20261202)
        nb = 1 << 18
        qs = rng.uniform(0.1, 0.5, args.batch_runs)
        gam = rng.uniform(0.9, 1.2, args.batch_runs)
        ecs = rng.uniform(1.0, 20.0, args.batch_runs)
        Rs = rng.integers(128, 513, args.batch_runs)
        bsteps = 80
        setups, cfgs = [], []
        t0 = time.perf_counter()
        for q, g, ec, R in zip(qs, gam, ecs, Rs):
            cfg = BKConfig(**{**base, "RPOINTS": int(R) - 1, "MCINTPOINTS": nb, "PRECISION": "fp32"})
            ic_b = MV(qs0sqr=float(q), x0=0.01, ec=float(ec), lambdaqcd=0.241, gamma=float(g))
            setups.append(evo.prepare(ic_b, cfg, device=device, seed=0))
            cfgs.append(cfg)
        torch.cuda.synchronize()
        t_prep = time.perf_counter() - t0
        log(f"batch: {args.batch_runs} runs, sum R = {int(Rs.sum())}, grids adapted in {t_prep:.1f} s")
        rb = torch.cat([s.r for s in setups])
        eb = torch.cat([s.edges for s in setups]).contiguous()
        N0b = torch.cat([s.N0 for s in setups])
        bev = dev.BatchEvolver(rb, eb, [int(v) for v in Rs], setups[0].table, cfgs[0], n_samples=nb,
                               precision="fp32", rk_method="rk2", step=0.1)
        bev.state.copy_(N0b)
        for _ in range(2):  # the first step tunes the K2 + Kf launch
            bev.step()
        bev.state.copy_(N0b)
        torch.cuda.synchronize()
        c0 = int(bev.active_evaluations.item())
        t0 = time.perf_counter()
        for _ in range(bsteps):
            bev.step()
        torch.cuda.synchronize()
        t_batch = time.perf_counter() - t0
        b_evals = int(bev.active_evaluations.item()) - c0
        Nb = bev.state.clone()
        tuning = bev.tuning()
        bflag = int(bev.flag.item())
        bev.close()
        t_seq, s_evals, off, dmax = 0.0, 0, 0, 0.0
        for s, cfg in zip(setups, cfgs):
            ev1 = dev.Evolver(s.r, s.edges, s.table, cfg, n_samples=nb, precision="fp32", rk_method="rk2",
                              step=0.1)
            ev1.state.copy_(s.N0)
            ev1.step()
            ev1.state.copy_(s.N0)
            torch.cuda.synchronize()
            c1 = int(ev1.active_evaluations.item())
            t0 = time.perf_counter()
            for _ in range(bsteps):
                ev1.step()
            torch.cuda.synchronize()
            t_seq += time.perf_counter() - t0
            s_evals += int(ev1.active_evaluations.item()) - c1
            R = s.r.numel()
            dmax = max(dmax, float((ev1.state - Nb[off:off + R]).abs().max()))
            off += R
            ev1.close()
        extra.update({
            "batch_runs": args.batch_runs,
            "batch_sum_r": int(Rs.sum()),
            "batch_wall_s": t_batch,
            "batch_samples_per_s": b_evals * nb / t_batch,
            "batch_sequential_wall_s": t_seq,
            "batch_speedup_vs_sequential": t_seq / t_batch,
            "batch_tuned_spt": tuning[0],
            "batch_tuned_grid_blocks": tuning[1],
            "batch_max_abs_dN_vs_sequential": dmax,
            "batch_nonfinite": bflag,
            "batch_same_evaluations": b_evals == s_evals,
        })
        log(f"batch: {t_batch:.2f} s ({b_evals * nb / t_batch:.3e} samples/s), one at a time {t_seq:.2f} s, "
            f"tuned spt {tuning[0]} grid {tuning[1]}, max |dN| vs single runs {dmax:.1e}")
        del setups, rb, eb, N0b, Nb

    if not args.no_hbm:
        # The HBM-bound sampling kernels on their own, 2^24 points x 4
        # dimensions: bytes moved per launch / launch time against 8 TB/s.
        # Sobol f32: writes 16 B per point; VEGAS map: reads u (32 B), writes
        # x (32 B), jac (8 B) and bins (4 B); histogram: reads bins (4 B) and
        # |w| (8 B).
        nh = 1 << 24
        table_h = dev.sobol_table(4, 0)
        edges_h = torch.linspace(0, 1, 51, dtype=torch.float64, device=device).expand(4, -1).contiguous()
        lo_h = torch.zeros(4, dtype=torch.float64, device=device)
        wd_h = torch.ones(4, dtype=torch.float64, device=device)
        u_h = dev.sobol_draw(table_h, nh, kind="f64", device=device)
        x_h, jac_h, b_h = dev.vegas_map(u_h, edges_h, lo_h, wd_h)

        def timed(fn, reps=5):
            fn()
            st = torch.cuda.current_stream()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(reps):
                fn()
            e1.record(st)
            e1.synchronize()
            return e0.elapsed_time(e1) / reps * 1e-3

        for name, fn, nbytes in (
            ("sobol_f32", lambda: dev.sobol_draw(table_h, nh, kind="f32", device=device), 16 * nh),
            ("sobol_f64", lambda: dev.sobol_draw(table_h, nh, kind="f64", device=device), 32 * nh),
            ("vegas_map", lambda: dev.vegas_map(u_h, edges_h, lo_h, wd_h), 76 * nh),
            ("vegas_histogram", lambda: dev.vegas_histogram(b_h, jac_h, 50), 12 * nh),
        ):
            sec = timed(fn)
            extra[f"hbm_{name}_GBps"] = nbytes / sec / 1e9
            extra[f"hbm_{name}_frac"] = nbytes / sec / 8e12
            log(f"{name}: {sec * 1e3:.3f} ms for 2^24 x 4, {nbytes / sec / 1e9:.0f} GB/s "
                f"({100 * nbytes / sec / 8e12:.1f} % of 8 TB/s)")
        del u_h, x_h, jac_h, b_h

    def host_cpu() -> dict:
        model = "unknown"
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
        return {"nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count(), "cpu_model": model}

    if not args.no_config1:
        # The 64-node plain-FP64 workload: R = 64 on [1e-6, 1e2], MV Qs0^2 =
        # 0.2, gamma = 1, 2^14 Sobol points (seed 0) per parent, midpoint RK2
        # with dY = 0.1 to Y = 1, K1 by the reference's rule at this grid size
        # (the nested adaptive Gauss-Kronrod rule).  On the GPU through
        # solve_bk, end to end (grid adaptation, evolver, 10 steps; one
        # untimed call first); on the CPU the oracle's RK driver (FP64, every
        # OpenMP thread) on the same grids and Sobol points.
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        import numpy as np

        from nlo_bk_optimized_amd import solve_bk
        from oracle import oracle as O

        c1 = BKConfig(RPOINTS=63, MINR=1e-6, MAXR=1e2, MCINTPOINTS=1 << 14, MC_BINS=50, VEGAS_ADAPT_ITERATIONS=5,
                      DE_SOLVER_STEP=0.1, RK_METHOD="rk2", PRECISION="fp64", K1_RULE="adaptive")
        solve_bk(ic, 1.0, c1, device=device, dtype=torch.float64, seed=0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        table1 = solve_bk(ic, 1.0, c1, device=device, dtype=torch.float64, seed=0)
        torch.cuda.synchronize()
        gpu1_s = time.perf_counter() - t0
        s1 = evo.prepare(ic, c1, device=device, seed=0)
        r1, N01, e1 = (x.cpu().numpy() for x in (s1.r, s1.N0, s1.edges))
        log("64-node FP64 RK2 to Y = 1: CPU oracle run")
        t0 = time.perf_counter()
        ref1 = O.evolve(r1, N01, e1, O.sobol_table(4, 0), 1 << 14, O.params(minr=1e-6), steps=10, h=0.1,
                        method="rk2", k1_rule="adaptive")
        cpu1_s = time.perf_counter() - t0
        d1 = np.abs(table1.N.cpu().numpy() - ref1)
        extra.update({
            "config1_gpu_fp64_s": gpu1_s,
            "config1_cpu_oracle_s": cpu1_s,
            "config1_cpu_threads": O.num_threads(),
            "config1_speedup_gpu_vs_cpu": cpu1_s / gpu1_s,
            "config1_max_abs_dN_gpu_vs_cpu": float(d1.max()),
            **{f"config1_{k}": v for k, v in host_cpu().items()},
        })
        log(f"64-node FP64 RK2 to Y = 1: GPU {gpu1_s:.3f} s (solve_bk end to end), CPU oracle {cpu1_s:.2f} s on "
            f"{O.num_threads()} threads, max |dN| {d1.max():.2e}")

    cpu = None
    if not args.no_cpu_baseline:
        # The FP64 oracle (C restatement of the reference CPU algorithm) on a
        # bounded sample of the same right-hand side: every active parent x
        # 2^cpu_samples_log2 Sobol points of K2 + Kf (reference formula) plus
        # their K1, at the initial condition with the adapted grids, on every
        # host core OpenMP uses (about 10 s at the default size).
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        import numpy as np

        from oracle import oracle as O

        r = setup.r.cpu().numpy()
        N0 = setup.N0.cpu().numpy()
        edges = setup.edges.cpu().numpy()
        sN, sS = O.evolution_splines(r, N0)
        lo, width = O.nlo_bounds(r)
        pick = np.nonzero((np.arange(r.size) >= 1) & (N0 <= 0.99999))[0]
        m = 1 << args.cpu_samples_log2
        table = O.sobol_table(4, 0)
        t0 = time.perf_counter()
        O.bk_nlo_sums(r[pick], sS, edges[pick], lo, width, table, 0, m, O.params(minr=1e-6))
        O.k1_fixed(r[pick], sN, p=O.params(minr=1e-6))
        cpu_s = time.perf_counter() - t0
        cpu = {
            "value": pick.size * m / cpu_s,
            "unit": "samples/s",
            "cores": O.num_threads(),
            "kind": "port",
            "sample": f"{pick.size} active parents x 2^{args.cpu_samples_log2} Sobol points of K2+Kf (FP64 "
                      f"reference formula) plus their K1 (8x24 GL per panel) at Y=0, {cpu_s:.2f} s wall",
            **host_cpu(),
        }
        log(f"cpu baseline: {cpu['value']:.4e} samples/s on {cpu['cores']} threads")

    ms_per_step = fp32["seconds"] / args.steps * 1e3
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"NLO BK evolution: {args.rpoints + 1} r in [1e-6,1e2], MV Qs0^2=0.2 gamma=1, "
                        f"2^{args.samples_log2} Sobol points/parent/RHS, VEGAS 50 bins adapted x5 then "
                        f"frozen, RK2 dY=0.1, {args.steps} steps",
            "r_points": args.rpoints + 1,
            "samples_per_parent": n,
            "rk": "midpoint RK2",
            "k1": "fixed 8x24 Gauss-Legendre per log-r panel",
            "dY": 0.1,
            "accumulate": "f64",
        },
        "evolution_s": fp32["seconds"],
        "adaptation_s": fp32["adapt_seconds"],
        "evolver_setup_s": fp32["setup_seconds"],
        "full_evolution_s": fp32["adapt_seconds"] + fp32["setup_seconds"] + fp32["seconds"],
        "first_setup_in_process_s": first_setup_s,
        "integrand_samples": samples,
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": measured_traffic(args.rpoints, args.samples_log2),
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of nlo_kernel<FP32, sum only> on this "
                              "workload, calibrated on a device copy (profiles/nlo_kernel_fp32_256r_counters.md)",
            "kernel": "nlo_kernel<FP32, sum only> + per-parent reduce (the evolver's instance), on the state "
                      "the timed steps ended at",
            "launch_ms": launch_ms,
            "active_parents": active_end,
            "flops_per_sample": FLOPS_PER_SAMPLE,
            "launch_ms_initial_condition": launch_ms0,
            "achieved_initial_condition": achieved0,
            "frac_initial_condition": achieved0 / FP32_PEAK_TFLOPS,
        },
        "cpu_baseline": cpu,
        "nonfinite": fp32["nonfinite"],
    }
    result.update(extra)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
