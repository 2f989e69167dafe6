from nlo_bk_optimized_amd.dipole.table import DipoleTable, load_bk_table, save_bk_table

__all__ = ["DipoleTable", "load_bk_table", "save_bk_table"]
