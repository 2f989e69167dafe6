"""GBW and BK-table dipole amplitudes evaluated on the device
(S/dipole/amplitude.py:1-302).

Both classes keep the reference's interface: ``dipole_amplitude(r, Y)``,
``min_r`` / ``max_r``, ``X0()``, ``saturation_scale(Y, Ns)``, and for the
table ``initialize_interpolation(Y)`` with the same initialised / local-spline
switching.  The numbers come from libnlo_gfx950 (nlo_dis_dipole_amplitude,
nlo_dis_bk_initialize); the host only keeps state and validates.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from enum import Enum, auto

import torch

from nlo_bk_optimized_amd.dipole.table import DipoleTable


class AmplitudeInterpolation(Enum):
    SPLINE_LINEAR = auto()
    LINEAR_LINEAR = auto()


def _as_device_f64(t: torch.Tensor, device: torch.device) -> torch.Tensor:
    return t.detach().to(device=device, dtype=torch.float64).contiguous()


def _evaluate(dipole, r: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
    from nlo_bk_optimized_amd import device as dev

    r, Y = torch.broadcast_tensors(torch.as_tensor(r), torch.as_tensor(Y))
    out_dtype = r.dtype if r.is_floating_point() else torch.float64
    target = dipole.device if hasattr(dipole, "device") else (r.device if r.is_cuda else torch.device("cuda"))
    rr = _as_device_f64(r, target).reshape(-1)
    yy = _as_device_f64(Y, target).reshape(-1)
    out = dev.dis_dipole_amplitude(dipole, rr, yy)
    return out.reshape(r.shape).to(device=r.device if r.is_cuda else target, dtype=out_dtype)


@dataclass(frozen=True, slots=True)
class GBW:
    """N = 1 - exp(-(r^2 Qs^2(Y))^gamma / 4), Qs^2 = Qs0^2 e^(lambda max(Y, ln 1/x0))."""

    Qs0sqr: float = 0.1
    lambda_: float = 0.3
    gamma: float = 1.0
    x0: float = 1.0

    kind = "gbw"

    @property
    def min_r(self) -> float:
        return 1e-30

    @property
    def max_r(self) -> float:
        return 1e3

    def X0(self) -> float:
        return self.x0

    def dipole_amplitude(self, r: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        return _evaluate(self, r, Y)

    def saturation_scale(self, Y: torch.Tensor, Ns: float) -> torch.Tensor:
        return _saturation_scale(self, Y, Ns)


class BKDipole:
    """N(r, Y) from a BK table (S/dipole/amplitude.py:49-270).

    The table is held on the device in FP64.  Until ``initialize_interpolation``
    is called every query uses the local 7-point (9 at the top of the grid)
    natural spline in r of the Y-linear table rows (SPLINE_LINEAR) or bilinear
    interpolation in r and Y (LINEAR_LINEAR); afterwards queries whose Y
    matches the initialised rapidity (BKDipole.interpolator_initialized) use
    the global spline at that Y, set to 1 beyond maxr_interpolate.
    """

    kind = "bk"

    def __init__(self, table: DipoleTable,
                 interpolation: AmplitudeInterpolation = AmplitudeInterpolation.SPLINE_LINEAR,
                 *, device: torch.device | str | None = None) -> None:
        if table.r.numel() < 9:
            raise ValueError("BKDipole requires at least nine r points")
        ratio = table.r[1] / table.r[0]
        expected = table.r[0] * torch.pow(ratio, torch.arange(table.r.numel(), dtype=table.r.dtype,
                                                                device=table.r.device))
        if not bool((torch.abs(expected / table.r - 1) <= 0.01).all()):
            raise ValueError("BKDipole requires a logarithmic r grid")
        self.table = table
        self.interpolation = interpolation
        self.bilinear = interpolation is AmplitudeInterpolation.LINEAR_LINEAR
        self.device = torch.device(device) if device is not None else (
            table.r.device if table.r.is_cuda else torch.device("cuda"))
        y0 = max(0.0, math.log(1 / table.x0))
        self.Y = table.y + table.y.new_tensor(y0)
        self.device_r = _as_device_f64(table.r, self.device)
        self.device_Y = _as_device_f64(self.Y, self.device)
        self.device_N = _as_device_f64(table.N, self.device)
        self._r_host = table.r.detach().double().cpu()
        self._y_host = self.Y.detach().double().cpu()
        self.device_coef: torch.Tensor | None = None
        self._interpolator_Y = -1.0
        self._maxr_interpolate = -1.0
        # the local splines of every row and r interval, computed once: a
        # query combines two precomputed cubics instead of solving a spline
        self.device_local: torch.Tensor | None = None
        if not self.bilinear:
            from nlo_bk_optimized_amd import device as dev

            self.device_local = dev.dis_bk_local_splines(self)

    @property
    def min_r(self) -> float:
        return float(self._r_host[0])

    @property
    def max_r(self) -> float:
        return float(self._r_host[-1])

    @property
    def min_Y(self) -> float:
        return float(self._y_host[0])

    @property
    def max_Y(self) -> float:
        return float(self._y_host[-1])

    @property
    def maxr_interpolate(self) -> float:
        return self._maxr_interpolate

    @property
    def interpolator_Y(self) -> float:
        return self._interpolator_Y

    def X0(self) -> float:
        return 1.0

    def interpolator_initialized(self, Y: torch.Tensor) -> torch.Tensor:
        """S/dipole/amplitude.py:95-103."""

        Y = torch.as_tensor(Y, dtype=torch.float64)
        near_zero = (torch.abs(Y) < 0.001) & (abs(self._interpolator_Y) < 0.001)
        den = torch.minimum(Y, Y.new_tensor(self._interpolator_Y))
        den = torch.where(den != 0, den, torch.ones_like(Y))
        match = (self._interpolator_Y >= 0) & (torch.abs(Y - self._interpolator_Y) / den < 0.001)
        return near_zero | match

    def initialize_interpolation(self, Y: float | torch.Tensor) -> None:
        """S/dipole/amplitude.py:111-131."""

        from nlo_bk_optimized_amd import device as dev

        y = float(Y.item()) if isinstance(Y, torch.Tensor) else float(Y)
        if 0 <= y <= self.min_Y:
            y = self.min_Y
        if y < 0:
            raise ValueError(f"cannot initialize interpolation at negative Y={y}")
        if y > self.max_Y:
            raise ValueError(f"cannot initialize interpolation at Y={y}; maximum is {self.max_Y}")
        if self.device_coef is not None and bool(self.interpolator_initialized(torch.tensor(y))):
            return
        self.device_coef = None
        self._interpolator_Y = -1.0
        self._maxr_interpolate = -1.0
        coef, maxr = dev.dis_bk_initialize(self, y)
        self.device_coef = coef
        self._interpolator_Y = y
        self._maxr_interpolate = maxr

    def dipole_amplitude(self, r: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        """S/dipole/amplitude.py:133-168 (Y outside the table after freezing raises)."""

        r, Y = torch.broadcast_tensors(torch.as_tensor(r), torch.as_tensor(Y))
        if not bool(torch.isfinite(r).all()) or not bool(torch.isfinite(Y).all()):
            raise ValueError("r and Y must contain only finite values")
        Yc = Y.detach().double()
        frozen = torch.where((Yc >= 0) & (Yc < self.min_Y), torch.full_like(Yc, self.min_Y), Yc)
        if bool(((frozen < self.min_Y) | (frozen > self.max_Y)).any()):
            raise ValueError(f"Y must lie in [{self.min_Y}, {self.max_Y}] after initial-rapidity freezing")
        return _evaluate(self, r, Y)

    def saturation_scale(self, Y: torch.Tensor, Ns: float) -> torch.Tensor:
        return _saturation_scale(self, Y, Ns)


def _saturation_scale(dipole, Y: torch.Tensor, Ns: float) -> torch.Tensor:
    """Bisection for N(r, Y) = Ns, Q_s^2 = 2 / r^2 (S/dipole/amplitude.py:263-293)."""

    Y = torch.as_tensor(Y, dtype=torch.float64)
    if Y.numel() != 1:
        raise ValueError("saturation_scale requires a scalar Y")

    def f(r: float) -> float:
        return float(dipole.dipole_amplitude(torch.tensor([r], dtype=torch.float64), Y.reshape(1)).item()) - Ns

    lower, upper = dipole.min_r * 1.0001, dipole.max_r * 0.999
    lv, uv = f(lower), f(upper)
    if lv * uv > 0:
        raise ValueError("the saturation-scale bracket does not contain a root")
    root = 0.5 * (lower + upper)
    for it in range(1, 1001):
        root = 0.5 * (lower + upper)
        v = f(root)
        if lv * v <= 0:
            upper = root
        else:
            lower, lv = root, v
        if abs(upper - lower) < 1e-5 * min(abs(lower), abs(upper)) and it < 1000:
            return torch.tensor(2 / root**2, dtype=torch.float64)
    raise RuntimeError(f"saturation-scale root finding failed at Y={float(Y)}")


__all__ = ["AmplitudeInterpolation", "BKDipole", "GBW"]
