"""Special functions on the device (S/numerics/special.py:10-97).

Inputs are converted to float64 device tensors; results come back in the
input's dtype and device.  float32 inputs are evaluated with the FP32 device
code, float64 with the FP64 code.
"""

from __future__ import annotations

import torch


def _apply(fn: str, x: torch.Tensor) -> torch.Tensor:
    from nlo_bk_optimized_amd import device as dev

    x = torch.as_tensor(x)
    target = x.device if x.is_cuda else torch.device("cuda")
    xd = x.detach().to(device=target, dtype=torch.float64).contiguous()
    out = dev.dis_special(fn, xd.reshape(-1), precision="fp32" if x.dtype is torch.float32 else "fp64")
    return out.reshape(x.shape).to(device=x.device, dtype=x.dtype if x.is_floating_point() else torch.float64)


def bessel_K0(x: torch.Tensor) -> torch.Tensor:
    return _apply("K0", x)


def bessel_K1(x: torch.Tensor) -> torch.Tensor:
    return _apply("K1", x)


def bessel_K2(x: torch.Tensor) -> torch.Tensor:
    """K2 = K0 + 2 K1 / x, +inf at 0."""

    return _apply("K2", x)


def Li2(x: torch.Tensor) -> torch.Tensor:
    """Real dilogarithm for x <= 1."""

    if not torch.as_tensor(x).is_floating_point():
        raise TypeError("Li2 expects a floating-point tensor")
    return _apply("Li2", x)


__all__ = ["Li2", "bessel_K0", "bessel_K1", "bessel_K2"]
