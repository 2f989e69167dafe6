"""NLO DIS impact factors and observables on the device (S/dis/__init__.py)."""

from nlo_bk_optimized_amd.dis.config import DISConfig, DISOrder, Polarization, Quark, QuarkType
from nlo_bk_optimized_amd.dis.observables import (
    F2,
    F2_batch,
    FL,
    FT,
    sigma_r,
    sigma_r_batch,
    structure_functions_batch,
)

__all__ = ["DISConfig", "DISOrder", "F2", "F2_batch", "FL", "FT", "Polarization", "Quark", "QuarkType", "sigma_r",
           "sigma_r_batch", "structure_functions_batch"]
