"""In-memory BK table and the legacy ``nlobk`` text format.

``DipoleTable`` has the fields and validation of S/dipole/table.py:12-45
(r[R], y[T], N[T, R], x0).  The text format is the four-header layout of
S/dipole/table.py:48-144: ``###minr``, ``###r_multiplier``, ``###points``,
``###x0``, then per rapidity a ``###y`` line followed by ``points`` amplitude
lines; lines starting with a single ``#`` are comments.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path

import torch


@dataclass(frozen=True, slots=True)
class DipoleTable:
    """BK amplitudes ``N(y, r)`` at evolution rapidities ``y``."""

    r: torch.Tensor
    y: torch.Tensor
    N: torch.Tensor
    x0: float

    def __post_init__(self) -> None:
        if self.r.ndim != 1 or self.y.ndim != 1 or self.N.ndim != 2:
            raise ValueError("r and y must be one-dimensional and N must be two-dimensional")
        if self.r.numel() < 2 or self.y.numel() < 1:
            raise ValueError("a dipole table requires at least two r points and one y point")
        if self.N.shape != (self.y.numel(), self.r.numel()):
            raise ValueError("N must have shape (len(y), len(r))")
        if self.r.device != self.y.device or self.r.device != self.N.device:
            raise ValueError("r, y, and N must be on the same device")
        if self.r.dtype != self.y.dtype or self.r.dtype != self.N.dtype:
            raise ValueError("r, y, and N must have the same dtype")
        if not self.r.is_floating_point():
            raise TypeError("r, y, and N must be floating-point tensors")
        finite = torch.stack(
            (
                torch.isfinite(self.r).all(),
                torch.isfinite(self.y).all(),
                torch.isfinite(self.N).all(),
                (self.r > 0).all(),
                (self.r[1:] > self.r[:-1]).all(),
                (self.y[1:] > self.y[:-1]).all() if self.y.numel() > 1 else self.r.new_ones((), dtype=torch.bool),
            )
        ).tolist()
        if not finite[0]:
            raise ValueError("r must contain only finite values")
        if not finite[1]:
            raise ValueError("y must contain only finite values")
        if not finite[2]:
            raise ValueError("N must contain only finite values")
        if not (finite[3] and finite[4]):
            raise ValueError("r must be positive and strictly increasing")
        if not finite[5]:
            raise ValueError("y must be strictly increasing")
        if not math.isfinite(self.x0) or self.x0 <= 0:
            raise ValueError("x0 must be positive and finite")


def load_bk_table(
    path: str | Path,
    *,
    dtype: torch.dtype = torch.float64,
    device: torch.device | str | None = None,
) -> DipoleTable:
    """Read a legacy table; the stored rapidities are kept as written."""

    filename = Path(path)
    header: list[str] = []
    blocks: list[tuple[float, list[float]]] = []
    for raw in filename.read_text(encoding="utf-8").splitlines():
        line = raw.strip()
        if not line:
            continue
        if line.startswith("###"):
            value = line[3:].strip()
            if len(header) < 4:
                header.append(value)
                continue
            try:
                blocks.append((float(value), []))
            except ValueError as error:
                raise ValueError(f"{filename} has an invalid rapidity header") from error
            continue
        if line.startswith("#"):
            continue
        if len(header) < 4:
            continue
        if not blocks:
            raise ValueError(f"{filename} contains amplitude data before a rapidity header")
        try:
            blocks[-1][1].append(float(line))
        except ValueError as error:
            raise ValueError(f"{filename} contains an invalid amplitude value") from error

    if len(header) != 4:
        raise ValueError(f"{filename} does not contain four BK configuration headers")
    try:
        minr, multiplier, points, x0 = (
            float(header[0]),
            float(header[1]),
            int(header[2]),
            float(header[3]),
        )
    except ValueError as error:
        raise ValueError(f"{filename} has an invalid BK configuration header") from error
    if minr <= 0 or multiplier <= 1 or points < 2 or x0 <= 0:
        raise ValueError(f"{filename} has invalid BK grid metadata")
    if not blocks:
        raise ValueError(f"{filename} contains no complete rapidity blocks")
    for _, values in blocks:
        if len(values) != points:
            raise ValueError(
                f"{filename} has {len(values)} amplitudes in a block; expected {points}"
            )

    index = torch.arange(points, dtype=dtype, device=device)
    r = minr * torch.pow(torch.tensor(multiplier, dtype=dtype, device=device), index)
    y = torch.tensor([b[0] for b in blocks], dtype=dtype, device=device)
    N = torch.tensor([b[1] for b in blocks], dtype=dtype, device=device)
    return DipoleTable(r=r, y=y, N=N, x0=x0)


def save_bk_table(path: str | Path, table: DipoleTable) -> None:
    """Write ``table`` in the legacy format (log-uniform r grids only)."""

    r = table.r.detach().cpu()
    multiplier = r[1] / r[0]
    expected = r[0] * torch.pow(multiplier, torch.arange(r.numel(), dtype=r.dtype))
    tolerance = 5e-6 if r.dtype == torch.float32 else 2e-12
    if not bool(torch.allclose(r, expected, rtol=tolerance, atol=0)):
        raise ValueError("the legacy BK format requires a logarithmically uniform r grid")
    lines = [
        "# nlo_bk_optimized_amd BK dipole table",
        f"###{r[0].item():.15e}",
        f"###{multiplier.item():.15e}",
        f"###{r.numel()}",
        f"###{table.x0:.15e}",
    ]
    for y_value, row in zip(table.y.detach().cpu().tolist(), table.N.detach().cpu().tolist()):
        lines.append(f"###{y_value:.15e}")
        lines.extend(f"{v:.15e}" for v in row)
    Path(path).write_text("\n".join(lines) + "\n", encoding="utf-8")


__all__ = ["DipoleTable", "load_bk_table", "save_bk_table"]
