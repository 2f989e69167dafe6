"""Initial dipole amplitudes N(r, Y = 0) for the BK evolution.

Interface of S/bk/initial_conditions.py:10-91: ``MV`` (fields qs0sqr, x0,
ec, lambdaqcd, gamma; ``dipole_amplitude``; ``min_r`` / ``max_r``) and
``ICDataFile`` (a two-column ``r N`` text file interpolated by a natural
cubic spline in r, 0 below the first node and 1 above the last).

The McLerran-Venugopalan form is

    E(r) = (r^2 Qs0^2)^gamma / 4 * ln(1 / (r Lambda) + e_c * 2.7182818),
    N(r) = E                 if E < 1e-5,
           1 - exp(-E)       otherwise,

with the truncated e and the linear small-E branch the reference uses
(S/bk/initial_conditions.py:29-35), so that the amplitude a caller gets is
the reference's to the last bit of the same floating-point operations.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path

import torch

from nlo_bk_optimized_amd.numerics.interpolation import NaturalCubicSpline

# e as the reference writes it: 2.7182818, not math.e.  The difference
# (2.8e-8) shifts N by ~1e-9 relative at r ~ 1/Lambda; parity needs it.
E_TRUNCATED = 2.7182818
# Below this exponent N is the exponent itself (the first Taylor term).
LINEAR_BELOW = 1e-5
# An initial-condition file needs at least this many nodes.
MIN_TABLE_ROWS = 10


def mv_amplitude(r: torch.Tensor, qs0sqr: float, gamma: float, ec: float, lambdaqcd: float) -> torch.Tensor:
    """The MV amplitude at radii r, in r's dtype and on r's device."""

    saturation = torch.pow(r.square() * qs0sqr, gamma)
    logarithm = torch.log(torch.reciprocal(r * lambdaqcd) + ec * E_TRUNCATED)
    exponent = saturation / 4 * logarithm
    return torch.where(exponent < LINEAR_BELOW, exponent, 1 - torch.exp(-exponent))


@dataclass(frozen=True, slots=True)
class MV:
    """McLerran-Venugopalan initial condition; ``x0`` is the Bjorken x of
    the evolution's starting point (carried into the output table)."""

    qs0sqr: float = 0.10
    x0: float = 0.01
    ec: float = 1.0
    lambdaqcd: float = 0.241
    gamma: float = 1.0

    @property
    def min_r(self) -> float:
        return 1e-99  # defined for every positive r

    @property
    def max_r(self) -> float:
        return 1e99

    def dipole_amplitude(self, r: torch.Tensor) -> torch.Tensor:
        return mv_amplitude(r, self.qs0sqr, self.gamma, self.ec, self.lambdaqcd)


def read_rN_columns(path: Path) -> tuple[list[float], list[float]]:
    """(r, N) from the first two whitespace-separated columns of every line
    that is neither blank nor a '#' comment; extra columns are ignored."""

    rs: list[float] = []
    ns: list[float] = []
    with path.open(encoding="utf-8") as handle:
        for number, text in enumerate(handle, start=1):
            fields = text.split()
            if not fields or fields[0].startswith("#"):
                continue
            if len(fields) < 2:
                raise ValueError(f"{path}:{number}: expected the two columns 'r N', found {len(fields)}")
            try:
                r_value, n_value = float(fields[0]), float(fields[1])
            except ValueError as error:
                raise ValueError(f"{path}:{number}: r and N must be numbers, got {fields[:2]}") from error
            rs.append(r_value)
            ns.append(n_value)
    if len(rs) < MIN_TABLE_ROWS:
        raise ValueError(f"{path}: {len(rs)} data rows; an initial condition needs {MIN_TABLE_ROWS} or more")
    return rs, ns


class ICDataFile:
    """Initial condition read from a two-column ``r N`` file.

    ``r`` and ``N`` hold the file's nodes as tensors of ``dtype`` on
    ``device``; the amplitude between nodes is the natural cubic spline of N
    in r, and outside [r_first, r_last] it is 0 (below) or 1 (above)."""

    def __init__(
        self,
        path: str | Path,
        *,
        x0: float = 0.01,
        dtype: torch.dtype = torch.float64,
        device: torch.device | str | None = None,
    ) -> None:
        self.path = Path(path)
        self.x0 = x0
        rs, ns = read_rN_columns(self.path)
        if not all(math.isfinite(v) for v in rs + ns):
            raise ValueError(f"{self.path}: r and N must be finite")
        self.r = torch.tensor(rs, dtype=dtype, device=device)
        self.N = torch.tensor(ns, dtype=dtype, device=device)
        self._spline = NaturalCubicSpline(self.r, self.N)

    @property
    def min_r(self) -> float:
        return float(self.r[0].item())

    @property
    def max_r(self) -> float:
        return float(self.r[-1].item())

    def dipole_amplitude(self, r: torch.Tensor) -> torch.Tensor:
        inside = self._spline(r)
        below, above = r < self.r[0], r > self.r[-1]
        return inside.masked_fill(below, 0.0).masked_fill(above, 1.0)


__all__ = ["E_TRUNCATED", "ICDataFile", "MV", "mv_amplitude", "read_rN_columns"]
