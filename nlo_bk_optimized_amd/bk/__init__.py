from nlo_bk_optimized_amd.bk.config import (
    BKConfig,
    BKOrder,
    IntegrationMethod,
    ResummationCoupling,
    RunningCouplingLO,
    RunningCouplingNLO,
)
from nlo_bk_optimized_amd.bk.evolution import DNDYResult, compute_dndy, solve_bk, solve_bk_batch
from nlo_bk_optimized_amd.bk.initial_conditions import MV, ICDataFile

__all__ = [
    "BKConfig",
    "BKOrder",
    "DNDYResult",
    "ICDataFile",
    "IntegrationMethod",
    "MV",
    "ResummationCoupling",
    "RunningCouplingLO",
    "RunningCouplingNLO",
    "compute_dndy",
    "solve_bk",
    "solve_bk_batch",
]
