"""``nlo-dis``: DIS structure-function grids on the GPU (S/cli/nlodis.py:48-164).

Same flags and run modes as the reference command: HERA_FL (FL at the 20
HERA kinematic points) and F2FL_GRID (light / charm F2 and FL and the LO F2
for Q^2 in {4.5, 45, 100} and x from 1e-5 up by factors of 1.5).  The BK
table dipole uses LINEAR_LINEAR interpolation, as the reference command does.

    python -m nlo_bk_optimized_amd.cli.nlodis --datafile bk.dat --C2 1 --charm-mass 1.27 \\
        --proton-area 8.4 --rc-scheme SMALLEST
"""

from __future__ import annotations

import argparse
import sys
from dataclasses import replace

import torch

from nlo_bk_optimized_amd.dipole.amplitude import AmplitudeInterpolation, BKDipole
from nlo_bk_optimized_amd.dipole.table import load_bk_table
from nlo_bk_optimized_amd.dis.config import (
    DISConfig,
    DISOrder,
    Quark,
    QuarkType,
    RunningCouplingIRScheme,
    RunningCouplingScheme,
)
from nlo_bk_optimized_amd.dis.observables import F2, FL, FT
from nlo_bk_optimized_amd.numerics.integration import IntegralResult

# (Q^2, x) of the HERA FL measurement (S/cli/nlodis.py:23-44)
HERA_FL_POINTS = (
    (1.5, 0.000028), (2.0, 0.000043), (2.5, 0.000059), (3.5, 0.000088), (5.0, 0.000129),
    (6.5, 0.000169), (8.5, 0.000224), (12.0, 0.000319), (15.0, 0.000402), (20.0, 0.000540),
    (25.0, 0.000687), (35.0, 0.000958), (45.0, 0.001210), (60.0, 0.001570), (90.0, 0.002430),
    (120.0, 0.003030), (150.0, 0.004020), (200.0, 0.005410), (250.0, 0.007360), (346.0, 0.009860),
)
MB_TO_GEV2 = 2.56819  # proton area sigma_0 / 2 in mb -> GeV^-2


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="Calculate DIS structure-function grids (MI355X)")
    p.add_argument("--name", default="nlodis")
    p.add_argument("--datafile", required=True)
    p.add_argument("--C2", type=float, required=True)
    p.add_argument("--charm-mass", type=float, required=True)
    p.add_argument("--light-mass", type=float, default=0.01)
    p.add_argument("--proton-area", type=float, required=True, help="sigma_0/2 in mb")
    p.add_argument("--rc-scheme", choices=("PARENT", "SMALLEST"), required=True)
    p.add_argument("--order", choices=("LO", "NLO"), default="NLO")
    p.add_argument("--runmode", choices=("F2FL_GRID", "HERA_FL"), default="HERA_FL")
    p.add_argument("--epsrel", type=float, default=0.01)
    p.add_argument("--mcintpoints", type=int, default=1_000_000)
    p.add_argument("--nf", type=int, default=-1)
    p.add_argument("--maxr", type=float, default=30.0)
    p.add_argument("--quadrature-points", type=int, default=24)
    p.add_argument("--batch-size", type=int, default=65_536)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--device", default="cuda")
    p.add_argument("--dtype", choices=("float32", "float64"), default="float32")
    p.add_argument("--cuda-nested-points", type=int, default=48)
    p.add_argument("--no-header", action="store_true")
    return p


def _report(contribution: str, Q2: float, xbj: float, result: IntegralResult) -> None:
    print(f"# {contribution} Q2={Q2:.15g} x={xbj:.15g} seed={result.seed} n_eval={result.n_eval} "
          f"stderr={result.error.item():.8e} converged={result.converged}", file=sys.stderr)


def main(argv: list[str] | None = None) -> int:
    args = build_parser().parse_args(argv)
    device = torch.device(args.device)
    dtype = torch.float32 if args.dtype == "float32" else torch.float64
    table = load_bk_table(args.datafile, dtype=torch.float64)
    dipole = BKDipole(table, interpolation=AmplitudeInterpolation.LINEAR_LINEAR, device=device)
    light = Quark(QuarkType.LIGHT, args.light_mass)
    charm = Quark(QuarkType.C, args.charm_mass)
    config = DISConfig(
        order=DISOrder[args.order],
        rc_scheme=RunningCouplingScheme[args.rc_scheme],
        rc_ir_scheme=RunningCouplingIRScheme.SMOOTH,
        maxr=args.maxr,
        C2_alpha=args.C2,
        nf_alphas=args.nf,
        transverse_area=args.proton_area * MB_TO_GEV2,
        maxeval=args.mcintpoints,
        epsrel=args.epsrel,
        cuda_nested_points=args.cuda_nested_points,
        quarks=(light, charm),
    )
    opts = {"quadrature_points": args.quadrature_points, "batch_size": args.batch_size, "seed": args.seed}
    if args.runmode == "F2FL_GRID":
        if not args.no_header:
            print("fit,x,Q2,F2 light,FL light,F2 charm,FL charm,F2_LO")
        for Q2v in (4.5, 45.0, 100.0):
            xv = 1e-5
            while xv <= 1e-2:
                Q2 = torch.tensor(Q2v, dtype=dtype)
                xbj = torch.tensor(xv, dtype=dtype)
                lc, cc = replace(config, quarks=(light,)), replace(config, quarks=(charm,))
                ftl, fll = FT(Q2, xbj, dipole, lc, **opts), FL(Q2, xbj, dipole, lc, **opts)
                ftc, flc = FT(Q2, xbj, dipole, cc, **opts), FL(Q2, xbj, dipole, cc, **opts)
                f2lo = F2(Q2, xbj, dipole, replace(config, order=DISOrder.LO, quarks=(light, charm)), **opts)
                print(f"{args.name},{xv:.15g},{Q2v:.15g},{(ftl.value + fll.value).item():.15e},"
                      f"{fll.value.item():.15e},{(ftc.value + flc.value).item():.15e},{flc.value.item():.15e},"
                      f"{f2lo.value.item():.15e}")
                for label, res in (("FT_light", ftl), ("FL_light", fll), ("FT_charm", ftc), ("FL_charm", flc)):
                    _report(label, Q2v, xv, res)
                xv *= 1.5
    else:
        if not args.no_header:
            print("fit,x,Q2,FL")
        for Q2v, xv in HERA_FL_POINTS:
            res = FL(torch.tensor(Q2v, dtype=dtype), torch.tensor(xv, dtype=dtype), dipole, config, **opts)
            print(f"{args.name},{xv:.15g},{Q2v:.15g},{res.value.item():.15e}")
            _report("FL", Q2v, xv, res)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
