"""DIS constants, quarks and configuration (S/dis/config.py:1-153).

The fields, defaults and validation follow the reference's DISConfig so that
existing configurations carry over.  ``cuda_fusion`` has no effect here (every
integrand runs fused in libnlo_gfx950); ``cuda_nested`` selects the nested
I2/I3 integrals with ``cuda_nested_points`` inner Gauss-Legendre nodes.
Two fields are added: ``bins`` (VEGAS bins per dimension, the reference's
vegas() default 32) and ``sobol_first_index`` (first Sobol index of every
VEGAS integral).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from enum import Enum, auto

LambdaQCD = 0.241
AlphaEM = 1 / 137.035999
NC = 3
CF = (NC**2 - 1) / (2 * NC)
Q0sqr = 1.0


class Polarization(Enum):
    T = auto()
    L = auto()


class DISOrder(Enum):
    LO = auto()
    NLO = auto()


class NcScheme(Enum):
    FiniteNC = auto()
    LargeNC = auto()


class RunningCouplingScheme(Enum):
    SMALLEST = auto()
    PARENT = auto()


class RunningCouplingIRScheme(Enum):
    FREEZE = auto()
    SMOOTH = auto()


class HeavyQuarkX(Enum):
    MassDependentX = auto()
    MassIndependentX = auto()


class QuarkType(Enum):
    U = auto()
    D = auto()
    S = auto()
    C = auto()
    B = auto()
    T = auto()
    LIGHT = auto()


_DEFAULT_MASS = {QuarkType.U: 0.002, QuarkType.D: 0.002, QuarkType.S: 0.002, QuarkType.LIGHT: 0.002,
                 QuarkType.C: 1.4, QuarkType.B: 4.18, QuarkType.T: 173.0}
_NOT_GIVEN = object()


@dataclass(frozen=True, slots=True, init=False)
class Quark:
    """A quark flavour: Quark() is a light quark of mass 0.001, Quark(type) takes
    the type's default mass (S/dis/config.py:60-85)."""

    type: QuarkType
    mass: float

    def __init__(self, type: QuarkType | object = _NOT_GIVEN, mass: float | None = None) -> None:
        plain = type is _NOT_GIVEN
        qt = QuarkType.LIGHT if plain else type
        if not isinstance(qt, QuarkType):
            raise TypeError("type must be a QuarkType")
        if mass is None:
            mass = 0.001 if plain else _DEFAULT_MASS[qt]
        object.__setattr__(self, "type", qt)
        object.__setattr__(self, "mass", float(mass))

    @property
    def charge(self) -> float:
        if self.type in (QuarkType.U, QuarkType.C):
            return 2 / 3
        if self.type in (QuarkType.D, QuarkType.S, QuarkType.B):
            return -1 / 3
        if self.type is QuarkType.LIGHT:
            return math.sqrt(2 / 3)
        raise ValueError("the source does not define an electric charge for the top quark")


@dataclass(frozen=True, slots=True)
class DISConfig:
    order: DISOrder = DISOrder.LO
    nc_scheme: NcScheme = NcScheme.FiniteNC
    rc_scheme: RunningCouplingScheme = RunningCouplingScheme.SMALLEST
    rc_ir_scheme: RunningCouplingIRScheme = RunningCouplingIRScheme.FREEZE
    maxr: float = 30.0
    C2_alpha: float = 1.0
    nf_alphas: int = -1
    max_alpha_s_freeze: float = 0.7
    heavy_quark_x_scheme: HeavyQuarkX = HeavyQuarkX.MassIndependentX
    transverse_area: float = 1.0
    maxeval: int = 2_000_000
    epsrel: float = 0.001
    cuda_fusion: bool = True
    cuda_nested: bool = True
    cuda_nested_points: int = 48
    quarks: tuple[Quark, ...] = field(default_factory=lambda: (Quark(QuarkType.LIGHT), Quark(QuarkType.C)))
    bins: int = 32
    sobol_first_index: int = 0

    def __post_init__(self) -> None:
        if self.maxr <= 0 or self.C2_alpha <= 0:
            raise ValueError("maxr and C2_alpha must be positive")
        if self.max_alpha_s_freeze <= 0 or self.transverse_area <= 0:
            raise ValueError("coupling freeze and transverse area must be positive")
        if self.maxeval < 1 or self.epsrel <= 0:
            raise ValueError("integration budget and relative accuracy must be positive")
        if not 8 <= self.cuda_nested_points <= 128:
            raise ValueError("nested DIS integration requires 8 to 128 inner points")
        if not self.quarks:
            raise ValueError("at least one quark is required")
        if not 2 <= self.bins <= 255:
            raise ValueError("bins must lie in [2, 255]")
        if self.sobol_first_index < 0 or self.sobol_first_index % 64:
            raise ValueError("sobol_first_index must be a non-negative multiple of 64")

    @property
    def active_flavors(self) -> int:
        if self.nf_alphas >= 0:
            return self.nf_alphas
        return sum(3 if q.type is QuarkType.LIGHT else 1 for q in self.quarks)


__all__ = [
    "AlphaEM",
    "CF",
    "DISConfig",
    "DISOrder",
    "HeavyQuarkX",
    "LambdaQCD",
    "NC",
    "NcScheme",
    "Polarization",
    "Q0sqr",
    "Quark",
    "QuarkType",
    "RunningCouplingIRScheme",
    "RunningCouplingScheme",
]
