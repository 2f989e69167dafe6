"""``nlo-bk``: evolve a dipole amplitude with the BK equation on the GPU
(S/cli/nlobk.py:22-117).

Same flags as the reference command, plus ``--precision`` (fp32 / fp64 /
fp64_stable), ``--rk-method`` and ``--mc-bins``.  Physics choices the device
path does not implement raise NotImplementedError (no silent fallback).

    python -m nlo_bk_optimized_amd.cli.nlobk --maxy 10 --output bk.dat
"""

from __future__ import annotations

import argparse

import torch

from nlo_bk_optimized_amd.bk.config import (
    BKConfig,
    BKOrder,
    IntegrationMethod,
    ResummationCoupling,
    RunningCouplingLO,
    RunningCouplingNLO,
)
from nlo_bk_optimized_amd.bk.evolution import compute_dndy, solve_bk
from nlo_bk_optimized_amd.bk.initial_conditions import MV, ICDataFile
from nlo_bk_optimized_amd.dipole.table import save_bk_table

ORDERS = {
    "lo": BKOrder.LO,
    "lo_resum_dlog": BKOrder.LO_RESUM_DLOG,
    "lo_resum_dlog_slog": BKOrder.LO_RESUM_DLOG_SLOG,
    "nlo": BKOrder.NLO,
    "nlo_resum_dlog": BKOrder.NLO_RESUM_DLOG,
    "nlo_resum_dlog_slog": BKOrder.NLO_RESUM_DLOG_SLOG,
}
RC_LO = {
    "fixed": RunningCouplingLO.FIXED_LO,
    "parent": RunningCouplingLO.PARENT_LO,
    "smallest": RunningCouplingLO.SMALLEST_LO,
    "balitsky": RunningCouplingLO.BALITSKY_LO,
    "fac": RunningCouplingLO.FAC_LO,
    "beuf": RunningCouplingLO.BEUF_LO,
}
RC_NLO = {"parent": RunningCouplingNLO.PARENT_NLO, "smallest": RunningCouplingNLO.SMALLEST_NLO}
RESUM_RC = {
    "parent": ResummationCoupling.RESUM_RC_PARENT,
    "smallest": ResummationCoupling.RESUM_RC_SMALLEST,
    "fixed": ResummationCoupling.RESUM_RC_FIXED,
}
INTEGRATION = {"vegas": IntegrationMethod.VEGAS, "miser": IntegrationMethod.MISER,
               "multiple": IntegrationMethod.MULTIPLE}


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="Evolve a dipole amplitude with the BK equation (MI355X)")
    p.add_argument("--ic-data", help="two-column r N initial condition")
    p.add_argument("--x0", type=float, default=0.01)
    p.add_argument("--qs0sqr", type=float, default=0.10)
    p.add_argument("--gamma", type=float, default=1.0)
    p.add_argument("--ec", type=float, default=1.0)
    p.add_argument("--lambdaqcd", type=float, default=0.241)
    p.add_argument("--maxy", type=float, default=10.0)
    p.add_argument("--output", default="output.dat")
    p.add_argument("--order", choices=tuple(ORDERS), default="nlo_resum_dlog_slog")
    p.add_argument("--rc-lo", choices=tuple(RC_LO), default="balitsky")
    p.add_argument("--rc-nlo", choices=tuple(RC_NLO), default="smallest")
    p.add_argument("--resum-rc", choices=tuple(RESUM_RC), default="smallest")
    p.add_argument("--integration", choices=tuple(INTEGRATION), default="vegas")
    p.add_argument("--nf", type=int, choices=(0, 3, 5), default=3)
    p.add_argument("--C2", type=float, default=1.0)
    p.add_argument("--fixed-alpha-s", type=float, default=0.2)
    p.add_argument("--Ksub", type=float, default=0.65)
    p.add_argument("--minr", type=float, default=1e-6)
    p.add_argument("--maxr", type=float, default=30.0)
    p.add_argument("--rpoints", type=int, default=100)
    p.add_argument("--mcintpoints", type=int, default=100_000)
    p.add_argument("--mc-bins", type=int, default=50)
    p.add_argument("--step", type=float, default=0.2)
    p.add_argument("--euler", action="store_true")
    p.add_argument("--rk-method", choices=("auto", "euler", "rk2", "kutta3", "rk4", "adaptive_rk2"),
                   default="auto")
    p.add_argument("--precision", choices=("fp32", "fp64", "fp64_stable"), default="fp32")
    p.add_argument("--kinematical-constraint", action="store_true")
    p.add_argument("--nolimit", action="store_true")
    p.add_argument("--dndy", action="store_true")
    p.add_argument("--seed", type=int)
    p.add_argument("--device", default="cuda")
    p.add_argument("--dtype", choices=("float32", "float64"), default="float32")
    return p


def config_from_args(args: argparse.Namespace) -> BKConfig:
    fixed = args.rc_lo == "fixed"
    return BKConfig(
        NF=args.nf,
        MAXR=args.maxr,
        MINR=args.minr,
        RPOINTS=args.rpoints,
        MCINTPOINTS=args.mcintpoints,
        DE_SOLVER_STEP=args.step,
        FIXED_AS=args.fixed_alpha_s,
        RC_LO=RC_LO[args.rc_lo],
        RC_NLO=RunningCouplingNLO.FIXED_NLO if fixed else RC_NLO[args.rc_nlo],
        RESUM_RC=RESUM_RC[args.resum_rc],
        INTMETHOD_NLO=INTEGRATION[args.integration],
        FORCE_POSITIVE_N=not args.nolimit,
        DNDY=args.dndy,
        Order=ORDERS[args.order],
        KSUB=args.Ksub,
        KINEMATICAL_CONSTRAINT=args.kinematical_constraint,
        EULER_METHOD=args.euler,
        C2=args.C2,
        MC_BINS=args.mc_bins,
        PRECISION=args.precision,
        RK_METHOD=args.rk_method,
    )


def main(argv: list[str] | None = None) -> int:
    parser = build_parser()
    args = parser.parse_args(argv)
    device = torch.device(args.device)
    dtype = torch.float32 if args.dtype == "float32" else torch.float64
    ic = (ICDataFile(args.ic_data, x0=args.x0, dtype=dtype) if args.ic_data
          else MV(qs0sqr=args.qs0sqr, x0=args.x0, ec=args.ec, lambdaqcd=args.lambdaqcd, gamma=args.gamma))
    config = config_from_args(args)
    if args.dndy:
        result = compute_dndy(ic, config, device=device, dtype=dtype, seed=args.seed)
        print("# r dN/dy[K1] dN/dy[K2+Kf] N")
        rows = torch.stack((result.r, result.K1, result.K2_Kf, result.N), dim=1)
        for row in rows.detach().cpu().tolist():
            print(" ".join(f"{value:.15e}" for value in row))
        return 0
    table = solve_bk(ic, args.maxy, config, device=device, dtype=dtype, seed=args.seed)
    save_bk_table(args.output, table)
    print(f"saved {table.y.numel()} rapidity slices and {table.r.numel()} r points to {args.output}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
