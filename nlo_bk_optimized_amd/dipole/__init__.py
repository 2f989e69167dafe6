from nlo_bk_optimized_amd.dipole.amplitude import GBW, AmplitudeInterpolation, BKDipole
from nlo_bk_optimized_amd.dipole.table import DipoleTable, load_bk_table, save_bk_table

__all__ = ["AmplitudeInterpolation", "BKDipole", "DipoleTable", "GBW", "load_bk_table", "save_bk_table"]
