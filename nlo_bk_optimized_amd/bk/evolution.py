"""BK radial grid, VEGAS grid adaptation and rapidity stepping on the GPU.

Entry points mirror S/bk/evolution.py:33-82 of the reference:
``solve_bk(initial_condition, maxy, config, *, device, dtype, seed) -> DipoleTable``
and ``compute_dndy(...) -> DNDYResult``.  The right-hand side
(S/bk/evolution.py:354-488) is evaluated by libnlo_gfx950: K1 on the fixed
production Gauss-Legendre rule, K2 + Kf as VEGAS sums over scrambled Sobol
points (seed ``seed``) mapped through per-parent importance grids that are
adapted on the initial condition and then frozen.  Each RK step is one
hipGraph replay.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from nlo_bk_optimized_amd import device as dev
from nlo_bk_optimized_amd.bk.config import BKConfig, require_gpu_supported, resolve_k1_rule, resolve_rk_method
from nlo_bk_optimized_amd.bk.initial_conditions import MV, ICDataFile
from nlo_bk_optimized_amd.dipole.table import DipoleTable


@dataclass(frozen=True, slots=True)
class DNDYResult:
    r: torch.Tensor
    K1: torch.Tensor
    K2_Kf: torch.Tensor
    N: torch.Tensor


def _r_grid(config: BKConfig, *, device: torch.device | str, dtype: torch.dtype) -> torch.Tensor:
    """r_i = MINR * step^i, step = (MAXR / MINR)^(1 / RPOINTS) (S/bk/evolution.py:584-592)."""

    rind = torch.arange(config.RPOINTS + 1, dtype=dtype, device=device)
    step = (config.MAXR / config.MINR) ** (1 / config.RPOINTS)
    return config.MINR * torch.pow(rind.new_tensor(step), rind)


def _resolve_device(device: torch.device | str | None) -> torch.device:
    d = torch.device("cuda" if device is None else device)
    if d.type != "cuda":
        raise ValueError(
            f"device={d}: the MI355X implementation runs on the GPU only (HIP device 'cuda')"
        )
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is available")
    return d


@dataclass
class BKSetup:
    """Grid, initial amplitude, Sobol table and adapted VEGAS grids."""

    r: torch.Tensor  # float64 [R]
    N0: torch.Tensor  # float64 [R]
    table: object
    edges: torch.Tensor  # float64 [R, 4, bins + 1]
    n_samples: int


def uniform_edges(R: int, bins: int, device: torch.device) -> torch.Tensor:
    e = torch.linspace(0, 1, bins + 1, dtype=torch.float64, device=device)
    return e.expand(R, 4, bins + 1).contiguous()


def adapt_grids(
    r: torch.Tensor,
    N: torch.Tensor,
    table,
    config: BKConfig,
    *,
    iterations: int,
    n_points: int,
    precision: str,
) -> torch.Tensor:
    """VEGAS refinement of every parent's grid on amplitude N.

    Iteration k histograms |w| over Sobol points [(k + 1) n, (k + 2) n) and
    refines the edges as S/numerics/integration.py:804-866 does; the
    production estimate later uses points [0, n).
    """

    R = r.numel()
    edges = uniform_edges(R, config.MC_BINS, r.device)
    if iterations == 0 or not config.Order.has_nlo_kernels:
        return edges
    t64, t32 = dev.spline_build(r, N, config.FORCE_POSITIVE_N)
    eps = dev.dtype_epsilon(torch.float64)
    for k in range(iterations):
        _, hist, count = dev.bk_nlo(
            r, N, t64, t32, edges, table, n_points, config,
            first_index=(k + 1) * n_points, precision=precision, histogram=True,
        )
        edges = dev.vegas_adapt(edges, hist, count, eps)
    return edges


def prepare(
    initial_condition: MV | ICDataFile,
    config: BKConfig,
    *,
    device: torch.device,
    seed: int | None,
) -> BKSetup:
    require_gpu_supported(config)
    r = _r_grid(config, device="cpu", dtype=torch.float64).to(device)
    N0 = initial_condition.dipole_amplitude(
        r if not isinstance(initial_condition, ICDataFile) else r.to(initial_condition.r.device, initial_condition.r.dtype)
    ).to(device=device, dtype=torch.float64)
    table = dev.sobol_table(4, 0 if seed is None else int(seed))
    n = int(config.MCINTPOINTS)
    n_adapt = int(config.VEGAS_ADAPT_POINTS or n)
    # adaptation blocks must start at multiples of 64 Sobol indices
    n_adapt_aligned = (n_adapt + 63) // 64 * 64
    if (config.VEGAS_ADAPT_ITERATIONS + 1) * max(n_adapt_aligned, n) > 2**30:
        raise ValueError("MCINTPOINTS x adaptation iterations exceed the 2^30 Sobol points")
    edges = adapt_grids(
        r, N0, table, config,
        iterations=config.VEGAS_ADAPT_ITERATIONS,
        n_points=n_adapt_aligned,
        precision=config.PRECISION,
    )
    return BKSetup(r=r, N0=N0, table=table, edges=edges, n_samples=n)


def rhs(setup: BKSetup, N: torch.Tensor, config: BKConfig,
        history: tuple[torch.Tensor, torch.Tensor] | None = None,
        k1_rule: str = "fixed") -> tuple[torch.Tensor, torch.Tensor]:
    """K1 and K2 + Kf at amplitude N (S/bk/evolution.py:354-488); with the
    kinematical constraint ``history`` = (rapidities, amplitudes), N its last row.
    ``k1_rule``: "fixed" or "adaptive" (resolve_k1_rule gives the reference's choice)."""

    t64, t32 = dev.spline_build(setup.r, N, config.FORCE_POSITIVE_N)
    k1 = dev.bk_k1(setup.r, N, t64, config, history=history if config.KINEMATICAL_CONSTRAINT else None,
                   rule=k1_rule)
    if config.Order.has_nlo_kernels:
        sums, _, _ = dev.bk_nlo(
            setup.r, N, t64, t32, setup.edges, setup.table, setup.n_samples, config,
            precision=config.PRECISION,
        )
        n = setup.n_samples
    else:
        sums = torch.zeros((N.numel(), 4), dtype=torch.float64, device=N.device)
        n = 1
    _, k2kf = dev.rhs_finalize(N, k1, sums, n)
    return k1, k2kf


def solve_bk(
    initial_condition: MV | ICDataFile,
    maxy: float,
    config: BKConfig = BKConfig(),
    *,
    device: torch.device | str = "cuda",
    dtype: torch.dtype = torch.float32,
    seed: int | None = None,
) -> DipoleTable:
    """Evolve through full output steps and return the final in-memory table."""

    if isinstance(initial_condition, ICDataFile):
        dtype = initial_condition.r.dtype
    d = _resolve_device(device)
    setup = prepare(initial_condition, config, device=d, seed=seed)
    rk = resolve_rk_method(config, dtype)
    k1_rule = resolve_k1_rule(config, dtype)
    tolerance = 1e-12 * max(1.0, abs(maxy), abs(config.DE_SOLVER_STEP))
    # output rapidities of fixed steps, as S/bk/evolution.py:108-159 produces them
    steps, y = 0, 0.0
    while True:
        y = y + config.DE_SOLVER_STEP
        steps += 1
        if y >= maxy - tolerance:
            break
    R = setup.r.numel()
    history_y = torch.zeros(steps + 1, dtype=torch.float64, device=d)
    history = torch.empty((steps + 1, R), dtype=torch.float64, device=d)
    history[0].copy_(setup.N0)
    kinematical = bool(config.KINEMATICAL_CONSTRAINT)
    evolver = dev.Evolver(
        setup.r, setup.edges, setup.table, config,
        n_samples=setup.n_samples, precision=config.PRECISION, rk_method=rk,
        step=config.DE_SOLVER_STEP, nlo=config.Order.has_nlo_kernels,
        eps=float(torch.finfo(dtype).eps), history=(history_y, history) if kinematical else None,
        k1_rule=k1_rule,
    )
    ys = [0.0]
    try:
        evolver.state.copy_(setup.N0)
        rows = [history]
        t = 0
        while True:
            evolver.step()
            if rk == "adaptive_rk2":
                y = evolver.adaptive_state()[0]  # the controller's own sum of sub-steps
            else:
                y = ys[-1] + config.DE_SOLVER_STEP
            t += 1
            if t <= steps:
                if not kinematical:  # the kinematical evolver appends its rows itself
                    history[t].copy_(evolver.state)
            else:  # an adaptive run whose rounded rapidity needs one more output step
                rows.append(evolver.state.clone().unsqueeze(0))
            ys.append(y)
            if y >= maxy - tolerance:
                break
        if int(evolver.flag.item()) != 0:
            raise FloatingPointError("non-finite BK amplitude during the rapidity evolution")
    finally:
        evolver.close()
    N_out = torch.cat([history[: min(t, steps) + 1]] + rows[1:], dim=0)
    r_out = setup.r.to(dtype)
    return DipoleTable(
        r=r_out,
        y=torch.tensor(ys, dtype=dtype, device=d),
        N=N_out.to(dtype),
        x0=initial_condition.x0,
    )


_SHARED_FIELDS = ("NC", "NF", "LambdaQCD", "MCINTPOINTS", "DE_SOLVER_STEP", "FIXED_AS", "RC_LO", "RC_NLO",
                  "RESUM_RC", "INTMETHOD_NLO", "FORCE_POSITIVE_N", "SYMMETRIZE_Z_Z2_INTEGRATION", "Order", "KSUB",
                  "KINEMATICAL_CONSTRAINT", "EULER_METHOD", "C2", "MC_BINS", "PRECISION", "RK_METHOD",
                  "K1_RADIAL_ORDER", "K1_ANGULAR_ORDER", "K1_RULE", "RINTPOINTS", "THETAINTPOINTS",
                  "INTACCURACY")


def solve_bk_batch(
    initial_conditions,
    maxy: float,
    configs: BKConfig | list[BKConfig] = BKConfig(),
    *,
    device: torch.device | str = "cuda",
    dtype: torch.dtype = torch.float32,
    seed: int | None = None,
    spt: int = 0,
    grid_blocks: int = 0,
    stats: dict | None = None,
) -> list[DipoleTable]:
    """Evolve many initial conditions together: ``solve_bk`` per run
    (S/bk/evolution.py:33-58), advanced in one set of launches per RK stage.

    The runs may differ in their r grid (RPOINTS, MINR, MAXR) and initial
    condition; every other BKConfig field must agree.  Stepping is fixed-step
    ("auto" means Kutta's third-order step); the kinematical constraint is not
    batched.  ``spt`` / ``grid_blocks`` fix the K2 + Kf launch shape (0: tuned on
    the first step); ``stats`` (a dict) receives the tuning and the evaluation count.
    "auto" stepping and K1 rules resolve as ``solve_bk`` resolves them for each
    run (the dtype of an ``ICDataFile`` run is the file's); the runs must
    resolve alike, and a resolution to the adaptive RK2 raises
    NotImplementedError (the batch steps with a fixed step size).
    """

    ics = list(initial_conditions)
    cfgs = list(configs) if isinstance(configs, (list, tuple)) else [configs] * len(ics)
    if not ics or len(cfgs) != len(ics):
        raise ValueError("one config (or one shared config) per initial condition is required")
    c0 = cfgs[0]
    for c in cfgs[1:]:
        diff = [f for f in _SHARED_FIELDS if getattr(c, f) != getattr(c0, f)]
        if diff:
            raise ValueError(f"batched runs must share {diff}")
    if c0.KINEMATICAL_CONSTRAINT:
        raise NotImplementedError("the kinematical constraint is not batched: use solve_bk per run")
    dtypes = [ic.r.dtype if isinstance(ic, ICDataFile) else dtype for ic in ics]
    methods = {resolve_rk_method(c, t) for c, t in zip(cfgs, dtypes)}
    if len(methods) != 1:
        raise ValueError(f"the runs resolve to different stepping methods {sorted(methods)}; set RK_METHOD")
    rk = methods.pop()
    if rk == "adaptive_rk2":
        raise NotImplementedError("the batched driver takes fixed-step methods only (RK_METHOD resolves to "
                                  "adaptive_rk2 here, as solve_bk would): set RK_METHOD to euler, rk2, kutta3 "
                                  "or rk4")
    rules = {resolve_k1_rule(c, t) for c, t in zip(cfgs, dtypes)}
    if len(rules) != 1:
        raise ValueError("the runs resolve to different K1 rules (K1_RULE='auto' with RPOINTS = 100 in "
                         "some runs only); set K1_RULE to 'fixed' or 'adaptive'")
    k1_rule = rules.pop()
    d = _resolve_device(device)
    setups = [prepare(ic, c, device=d, seed=seed) for ic, c in zip(ics, cfgs)]
    sizes = [s.r.numel() for s in setups]
    r = torch.cat([s.r for s in setups])
    edges = torch.cat([s.edges for s in setups]).contiguous()
    N0 = torch.cat([s.N0 for s in setups])
    tolerance = 1e-12 * max(1.0, abs(maxy), abs(c0.DE_SOLVER_STEP))
    steps, y, ys = 0, 0.0, [0.0]
    while True:
        y = y + c0.DE_SOLVER_STEP
        steps += 1
        ys.append(y)
        if y >= maxy - tolerance:
            break
    ev = dev.BatchEvolver(r, edges, sizes, setups[0].table, c0, n_samples=setups[0].n_samples,
                          precision=c0.PRECISION, rk_method=rk, step=c0.DE_SOLVER_STEP,
                          nlo=c0.Order.has_nlo_kernels, spt=spt, grid_blocks=grid_blocks, k1_rule=k1_rule)
    try:
        history = torch.empty((steps + 1, r.numel()), dtype=torch.float64, device=d)
        history[0].copy_(N0)
        ev.state.copy_(N0)
        for t in range(steps):
            ev.step()
            history[t + 1].copy_(ev.state)
        if int(ev.flag.item()) != 0:
            raise FloatingPointError("non-finite BK amplitude during the rapidity evolution")
        if stats is not None:
            stats["spt"], stats["grid_blocks"], stats["ms_per_nlo_launch"] = ev.tuning()
            stats["active_evaluations"] = int(ev.active_evaluations.item())
            stats["n_samples"] = setups[0].n_samples
    finally:
        ev.close()
    out, off = [], 0
    for s, ic, R, t in zip(setups, ics, sizes, dtypes):
        out.append(DipoleTable(r=s.r.to(t), y=torch.tensor(ys, dtype=t, device=d),
                               N=history[:, off:off + R].to(t), x0=ic.x0))
        off += R
    return out


def compute_dndy(
    initial_condition: MV | ICDataFile,
    config: BKConfig = BKConfig(),
    *,
    device: torch.device | str = "cuda",
    dtype: torch.dtype = torch.float32,
    seed: int | None = None,
) -> DNDYResult:
    """Diagnostic columns at the initial rapidity (S/bk/evolution.py:61-82)."""

    if isinstance(initial_condition, ICDataFile):
        dtype = initial_condition.r.dtype
    d = _resolve_device(device)
    setup = prepare(initial_condition, config, device=d, seed=seed)
    history = (torch.zeros(1, dtype=torch.float64, device=d), setup.N0.unsqueeze(0).contiguous())
    k1, k2kf = rhs(setup, setup.N0, config, history=history, k1_rule=resolve_k1_rule(config, dtype))
    keep = torch.arange(setup.r.numel(), device=d) > 0
    keep = keep & (setup.N0 <= 0.99999)
    return DNDYResult(
        r=setup.r[keep].to(dtype),
        K1=k1[keep].to(dtype),
        K2_Kf=k2kf[keep].to(dtype),
        N=setup.N0[keep].to(dtype),
    )


__all__ = ["BKSetup", "DNDYResult", "adapt_grids", "compute_dndy", "prepare", "rhs", "solve_bk", "solve_bk_batch"]
