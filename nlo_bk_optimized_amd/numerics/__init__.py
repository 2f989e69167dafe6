from nlo_bk_optimized_amd.numerics.interpolation import MIN_LOG_VALUE, NaturalCubicSpline

__all__ = ["MIN_LOG_VALUE", "NaturalCubicSpline"]
