"""MI355X (gfx950) implementation of the NLO BK right-hand side and rapidity
driver and of the NLO DIS structure functions, drop-in for the reference's BK
and DIS entry points.

The numerics run in ``libnlo_gfx950.so`` (hand-written HIP kernels behind a C
ABI, ``include/nlo_gfx950.h``); this package is the Python surface of the
reference (S/__init__.py of torch_nlo_clean/src/nlo_torch).
"""

from nlo_bk_optimized_amd.bk import (
    BKConfig,
    BKOrder,
    DNDYResult,
    ICDataFile,
    IntegrationMethod,
    MV,
    ResummationCoupling,
    RunningCouplingLO,
    RunningCouplingNLO,
    compute_dndy,
    solve_bk,
    solve_bk_batch,
)
from nlo_bk_optimized_amd.dipole import GBW, BKDipole, DipoleTable, load_bk_table, save_bk_table
from nlo_bk_optimized_amd.dis import F2, FL, FT, DISConfig, DISOrder, Polarization, Quark, QuarkType, sigma_r

__all__ = [
    "BKConfig",
    "BKDipole",
    "BKOrder",
    "DISConfig",
    "DISOrder",
    "F2",
    "FL",
    "FT",
    "GBW",
    "Polarization",
    "Quark",
    "QuarkType",
    "sigma_r",
    "solve_bk_batch",
    "DNDYResult",
    "DipoleTable",
    "ICDataFile",
    "IntegrationMethod",
    "MV",
    "ResummationCoupling",
    "RunningCouplingLO",
    "RunningCouplingNLO",
    "compute_dndy",
    "load_bk_table",
    "save_bk_table",
    "solve_bk",
]
