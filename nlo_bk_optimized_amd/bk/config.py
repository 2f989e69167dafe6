"""BK evolution choices (mirrors S/bk/config.py:9-124 of the reference).

The enum members, field names, defaults and ``__post_init__`` validation are
those of the reference's ``BKConfig`` so that callers can switch without
changes.  Fields after ``VEGAS_REUSE_WARMUP_FRACTION`` are this
implementation's own sampling and precision knobs (Sobol points, frozen VEGAS
grids, FP32/FP64 integrand, RK method).  The reference's backend switches
``CUDA_FUSION``, ``K1_FIXED``, ``K1_FIXED_REFINE`` and ``VEGAS_REUSE_GRID``
are accepted and validated but select nothing here: there is one GPU path.
"""

from __future__ import annotations

from dataclasses import dataclass
from enum import Enum, auto


class RunningCouplingLO(Enum):
    FIXED_LO = auto()
    PARENT_LO = auto()
    SMALLEST_LO = auto()
    BALITSKY_LO = auto()
    FAC_LO = auto()
    BEUF_LO = auto()


class RunningCouplingNLO(Enum):
    FIXED_NLO = auto()
    PARENT_NLO = auto()
    SMALLEST_NLO = auto()


class ResummationCoupling(Enum):
    RESUM_RC_PARENT = auto()
    RESUM_RC_SMALLEST = auto()
    RESUM_RC_FIXED = auto()


class IntegrationMethod(Enum):
    VEGAS = auto()
    MISER = auto()
    MULTIPLE = auto()


class BKOrder(Enum):
    LO = auto()
    LO_RESUM_DLOG = auto()
    LO_RESUM_DLOG_SLOG = auto()
    NLO = auto()
    NLO_RESUM_DLOG = auto()
    NLO_RESUM_DLOG_SLOG = auto()

    @property
    def has_nlo_kernels(self) -> bool:
        return self in {BKOrder.NLO, BKOrder.NLO_RESUM_DLOG, BKOrder.NLO_RESUM_DLOG_SLOG}

    @property
    def resum_dlog(self) -> bool:
        return self in {
            BKOrder.LO_RESUM_DLOG,
            BKOrder.LO_RESUM_DLOG_SLOG,
            BKOrder.NLO_RESUM_DLOG,
            BKOrder.NLO_RESUM_DLOG_SLOG,
        }

    @property
    def resum_slog(self) -> bool:
        return self in {BKOrder.LO_RESUM_DLOG_SLOG, BKOrder.NLO_RESUM_DLOG_SLOG}


PRECISIONS = ("fp32", "fp64", "fp64_stable")
RK_METHODS = ("auto", "euler", "rk2", "kutta3", "rk4", "adaptive_rk2")
K1_RULES = ("auto", "fixed", "adaptive")
MAX_K1_INTERVALS = 100  # device limit of the adaptive K1 interval lists


@dataclass(frozen=True, slots=True)
class BKConfig:
    NC: float = 3.0
    NF: int = 3
    LambdaQCD: float = 0.241
    RINTPOINTS: int = 85
    THETAINTPOINTS: int = 85
    INTACCURACY: float = 0.001
    MCINTACCURACY: float = 0.2
    MAXR: float = 30.0
    MINR: float = 1e-6
    RPOINTS: int = 100
    MCINTPOINTS: int = 100_000
    DE_SOLVER_STEP: float = 0.2
    DE_SOLVER_ABSERR: float = 1e-6
    DE_SOLVER_RELERR: float = 1e-4
    FIXED_AS: float = 0.2
    RC_LO: RunningCouplingLO = RunningCouplingLO.BALITSKY_LO
    RC_NLO: RunningCouplingNLO = RunningCouplingNLO.SMALLEST_NLO
    RESUM_RC: ResummationCoupling = ResummationCoupling.RESUM_RC_SMALLEST
    INTMETHOD_NLO: IntegrationMethod = IntegrationMethod.VEGAS
    FORCE_POSITIVE_N: bool = True
    SYMMETRIZE_Z_Z2_INTEGRATION: bool = True
    DNDY: bool = False
    Order: BKOrder = BKOrder.NLO_RESUM_DLOG_SLOG
    KSUB: float = 0.65
    KINEMATICAL_CONSTRAINT: bool = False
    EULER_METHOD: bool = False
    C2: float = 1.0
    CUDA_FUSION: bool = True
    K1_FIXED: bool = True
    K1_FIXED_REFINE: bool = False
    VEGAS_REUSE_GRID: bool = True
    VEGAS_REUSE_WARMUP_FRACTION: float = 0.25
    # --- sampling and precision of this implementation -----------------
    # VEGAS bins per dimension of the 4-D K2 + Kf box.
    MC_BINS: int = 50
    # Adaptation iterations on the initial condition; the grid is then frozen.
    VEGAS_ADAPT_ITERATIONS: int = 5
    # Sobol points per adaptation iteration (None: MCINTPOINTS).
    VEGAS_ADAPT_POINTS: int | None = None
    # "fp32": FP32 integrand (cancellation-free K2/Kf) with FP64 sums;
    # "fp64": the reference's K2/Kf formula in FP64; "fp64_stable": FP64 with
    # the cancellation-free kernels.
    PRECISION: str = "fp32"
    # Rapidity stepping: "auto" chooses as the reference's _bk_steps does
    # (S/bk/evolution.py:108-146): Euler for EULER_METHOD, the three-stage
    # Kutta step (S/bk/evolution.py:190-236) where the reference takes its
    # fixed RK2(3) step, otherwise the adaptive midpoint RK2 with step
    # doubling (S/bk/evolution.py:239-326).  Or explicitly "euler", "rk2"
    # (midpoint), "kutta3", "rk4", "adaptive_rk2".
    RK_METHOD: str = "auto"
    # K1 Gauss-Legendre orders per log-r panel (ln z) and in v (theta = pi v^2).
    K1_RADIAL_ORDER: int = 8
    K1_ANGULAR_ORDER: int = 24
    # K1 quadrature: "fixed" (the Gauss-Legendre rule above), "adaptive" (the
    # nested adaptive Gauss-Kronrod 21 rule of rapidity_derivative_lo,
    # S/bk/derivatives.py:43-110, to INTACCURACY with at most RINTPOINTS /
    # THETAINTPOINTS intervals) or "auto": the reference's choice
    # (S/bk/evolution.py:160-177, :491-533): the fixed rule for float32 where
    # it is validated (production physics, RPOINTS = 100, RINTPOINTS =
    # THETAINTPOINTS = 85, INTACCURACY = 0.001), the adaptive rule otherwise.
    K1_RULE: str = "auto"

    def __post_init__(self) -> None:
        if self.NF not in {0, 3, 5}:
            raise ValueError("NF must be 0, 3, or 5")
        if self.MINR <= 0 or self.MAXR <= self.MINR or self.RPOINTS < 2:
            raise ValueError("the BK r grid requires 0 < MINR < MAXR and RPOINTS >= 2")
        if self.RINTPOINTS < 1 or self.THETAINTPOINTS < 1 or self.MCINTPOINTS < 1:
            raise ValueError("integration limits must be positive")
        if self.INTACCURACY <= 0 or self.MCINTACCURACY <= 0:
            raise ValueError("integration accuracies must be positive")
        if not 0 <= self.VEGAS_REUSE_WARMUP_FRACTION <= 1:
            raise ValueError("VEGAS_REUSE_WARMUP_FRACTION must lie between zero and one")
        if self.DE_SOLVER_STEP <= 0:
            raise ValueError("DE_SOLVER_STEP must be positive")
        if self.C2 <= 0 or self.KSUB <= 0:
            raise ValueError("C2 and KSUB must be positive")

        fixed_LO = self.RC_LO is RunningCouplingLO.FIXED_LO
        fixed_NLO = self.RC_NLO is RunningCouplingNLO.FIXED_NLO
        if fixed_LO != fixed_NLO:
            raise ValueError("fixed coupling must be selected for both RC_LO and RC_NLO")
        if self.KINEMATICAL_CONSTRAINT and not self.EULER_METHOD:
            raise ValueError("KINEMATICAL_CONSTRAINT requires EULER_METHOD")

        if not 2 <= self.MC_BINS <= 255:
            raise ValueError("MC_BINS must lie in [2, 255]")
        if self.VEGAS_ADAPT_ITERATIONS < 0:
            raise ValueError("VEGAS_ADAPT_ITERATIONS must be non-negative")
        if self.VEGAS_ADAPT_POINTS is not None and self.VEGAS_ADAPT_POINTS < 1:
            raise ValueError("VEGAS_ADAPT_POINTS must be positive")
        if self.PRECISION not in PRECISIONS:
            raise ValueError(f"PRECISION must be one of {PRECISIONS}")
        if self.RK_METHOD not in RK_METHODS:
            raise ValueError(f"RK_METHOD must be one of {RK_METHODS}")
        if not 1 <= self.K1_RADIAL_ORDER <= 32 or not 1 <= self.K1_ANGULAR_ORDER <= 64:
            raise ValueError("K1 Gauss-Legendre orders must lie in [1, 32] x [1, 64]")
        if self.K1_RULE not in K1_RULES:
            raise ValueError(f"K1_RULE must be one of {K1_RULES}")


def require_gpu_supported(config: BKConfig) -> None:
    """Reject configurations the device path does not implement: the NLO
    terms are integrated by VEGAS only (MISER and the nested adaptive
    MULTIPLE rule of S/bk/derivatives.py:313-420, :635-709 are not built)."""

    if config.INTMETHOD_NLO is not IntegrationMethod.VEGAS:
        raise NotImplementedError(
            f"the MI355X BK path integrates K2 + Kf by VEGAS only; unsupported: "
            f"INTMETHOD_NLO={config.INTMETHOD_NLO.name}"
        )


def fixed_k1_validated(config: BKConfig) -> bool:
    """_fixed_k1_supported (S/bk/derivatives.py:145-166): the configuration
    the reference validated its fixed K1 rule on."""

    return (
        not config.KINEMATICAL_CONSTRAINT
        and config.RC_LO is RunningCouplingLO.BALITSKY_LO
        and config.RESUM_RC is ResummationCoupling.RESUM_RC_SMALLEST
        and config.Order is BKOrder.NLO_RESUM_DLOG_SLOG
        and config.NF <= 3
        and config.RPOINTS == 100
        and config.RINTPOINTS == 85
        and config.THETAINTPOINTS == 85
        and config.INTACCURACY == 0.001
    )


def resolve_rk_method(config: BKConfig, dtype) -> str:
    """The stepping method RK_METHOD="auto" stands for (S/bk/evolution.py:108-146,
    :180-187): the reference takes its fixed RK2(3) step only for float32 with
    the validated fixed K1 rule (production physics, RPOINTS = 100,
    RINTPOINTS = THETAINTPOINTS = 85, INTACCURACY = 0.001); else the adaptive RK2."""

    import torch

    if config.EULER_METHOD:
        return "euler"
    if config.RK_METHOD != "auto":
        return config.RK_METHOD
    if dtype is torch.float32 and config.INTMETHOD_NLO is IntegrationMethod.VEGAS and fixed_k1_validated(config):
        return "kutta3"
    return "adaptive_rk2"


def resolve_k1_rule(config: BKConfig, dtype) -> str:
    """The K1 rule K1_RULE="auto" stands for: on the GPU with float32 the
    reference switches K1_FIXED on (S/bk/evolution.py:160-177) and takes the
    fixed rule where it is validated (:491-533); everything else goes through
    rapidity_derivative_lo's adaptive rule.  The adaptive rule's interval
    limits must fit the device lists."""

    import torch

    rule = config.K1_RULE
    if rule == "auto":
        rule = "fixed" if dtype is torch.float32 and fixed_k1_validated(config) else "adaptive"
    if rule == "adaptive" and not (
        1 <= config.RINTPOINTS <= MAX_K1_INTERVALS and 1 <= config.THETAINTPOINTS <= MAX_K1_INTERVALS
    ):
        raise NotImplementedError(
            f"the adaptive K1 rule keeps at most {MAX_K1_INTERVALS} intervals per list on the device; "
            f"got RINTPOINTS={config.RINTPOINTS}, THETAINTPOINTS={config.THETAINTPOINTS}"
        )
    return rule


__all__ = [
    "BKConfig",
    "BKOrder",
    "IntegrationMethod",
    "K1_RULES",
    "PRECISIONS",
    "RK_METHODS",
    "fixed_k1_validated",
    "resolve_k1_rule",
    "resolve_rk_method",
    "ResummationCoupling",
    "RunningCouplingLO",
    "RunningCouplingNLO",
    "require_gpu_supported",
]
