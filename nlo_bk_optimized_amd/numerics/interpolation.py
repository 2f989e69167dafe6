"""Natural cubic splines for initial-condition files and table I/O.

Host-side interpolation outside the hot path, with the semantics of
S/numerics/interpolation.py:10-92: natural end conditions, Thomas solve,
interval = searchsorted(x, v, right=True) - 1 clamped to [0, n - 2].  The
per-step log-log splines of the evolution are built on the GPU
(``nlo_bk_spline_build``).
"""

from __future__ import annotations

import torch

MIN_LOG_VALUE = -40.0


class NaturalCubicSpline:
    """Natural cubic spline of y(x) on a 1-D strictly increasing grid."""

    def __init__(self, x: torch.Tensor, y: torch.Tensor) -> None:
        if x.ndim != 1 or y.ndim != 1 or x.shape != y.shape:
            raise ValueError("x and y must be one-dimensional tensors of equal shape")
        if x.numel() < 3:
            raise ValueError("a cubic spline requires at least three points")
        if x.device != y.device or x.dtype != y.dtype:
            raise ValueError("x and y must have the same device and dtype")
        if not x.is_floating_point():
            raise TypeError("x and y must be floating-point tensors")
        if not bool(torch.isfinite(x).all()) or not bool(torch.isfinite(y).all()):
            raise ValueError("x and y must contain only finite values")
        h = x[1:] - x[:-1]
        if not bool((h > 0).all()):
            raise ValueError("x must be strictly increasing")

        m = x.numel() - 2
        hc = h.double().cpu().tolist()
        yc = y.double().cpu().tolist()
        upper = [0.0] * max(m, 1)
        sol = [0.0] * max(m, 1)
        for i in range(m):
            diag = 2.0 * (hc[i] + hc[i + 1])
            rhs = 6.0 * ((yc[i + 2] - yc[i + 1]) / hc[i + 1] - (yc[i + 1] - yc[i]) / hc[i])
            if i == 0:
                if m > 1:
                    upper[0] = hc[1] / diag
                sol[0] = rhs / diag
            else:
                den = diag - hc[i] * upper[i - 1]
                if i < m - 1:
                    upper[i] = hc[i + 1] / den
                sol[i] = (rhs - hc[i] * sol[i - 1]) / den
        for i in range(m - 2, -1, -1):
            sol[i] = sol[i] - upper[i] * sol[i + 1]
        sec = torch.tensor([0.0] + sol[:m] + [0.0], dtype=y.dtype, device=y.device)
        self.x = x
        self.a = y[:-1]
        self.b = (y[1:] - y[:-1]) / h - h * (2 * sec[:-1] + sec[1:]) / 6
        self.c = sec[:-1] / 2
        self.d = (sec[1:] - sec[:-1]) / (6 * h)

    def __call__(self, v: torch.Tensor) -> torch.Tensor:
        if v.device != self.x.device or v.dtype != self.x.dtype:
            raise ValueError("evaluation points must have the spline device and dtype")
        k = torch.searchsorted(self.x, v, right=True) - 1
        k = k.clamp(0, self.x.numel() - 2)
        dx = v - self.x[k]
        return self.a[k] + dx * (self.b[k] + dx * (self.c[k] + dx * self.d[k]))


__all__ = ["MIN_LOG_VALUE", "NaturalCubicSpline"]
