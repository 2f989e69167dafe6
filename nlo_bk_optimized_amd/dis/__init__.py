"""NLO DIS impact factors and observables on the device (S/dis/__init__.py)."""

from nlo_bk_optimized_amd.dis.config import DISConfig, DISOrder, Polarization, Quark, QuarkType
from nlo_bk_optimized_amd.dis.observables import F2, FL, FT, sigma_r

__all__ = ["DISConfig", "DISOrder", "F2", "FL", "FT", "Polarization", "Quark", "QuarkType", "sigma_r"]
