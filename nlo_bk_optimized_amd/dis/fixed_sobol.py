"""The GBW fixed-Sobol qqbar-g estimator on the device (S/dis/fixed_sobol.py).

``gbw_qg_fixed_sobol`` gives one quark's complete qqbar-g term at one point:
three training point sets (65 536 scrambled Sobol points each) learn one
frozen 32-bin grid from the normalised magnitudes of I1, I2 and I3, then seven
independently scrambled replicates (32 768 points each) through that grid give
I1, I2, I3 and their correlated sum as replicate means with standard errors.
Every point is mapped through the grid and split into the two nearest-endpoint
sectors of the gluon plane; the integrands are evaluated at both.  The whole
estimator, for any number of (point, polarisation, quark) entries, is one
``nlo_dis_fixed_sobol`` call: training passes, fixed-order histogram sums, grid
refinement, replicate passes and statistics all run on the GPU.

The point sets are the reference's (SobolEngine(5, scramble=True,
seed=seed + k) for training round k, seed + 10000 + k for replicate k), drawn
in the library bit for bit as SobolEngine.draw_base2 gives them.

Differences from the reference, on purpose:

* the training histograms are FP64 sums in point order (the reference sums
  FP32 scores with ``scatter_add_``, whose order CUDA does not fix); the grid
  is refined in FP64 with the point dtype's epsilon, and on the float32 path
  the map and endpoint split run in float32 through the grid rounded to
  float32, as the reference's;
* I1 is evaluated in FP64 at every point (the reference evaluates an FP32
  regular region and an FP64 sensitive region, K/dis/i1/kernel.cu:412-433);
* the generic forms ``learn_triple_importance_grid`` and
  ``integrate_triple_fixed_sobol`` over Python callables are not provided: the
  estimator runs the device integrands;
* float64 scalars are accepted too (every step in FP64), beside the
  reference's float32.
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache

import torch

from nlo_bk_optimized_amd.dipole.amplitude import GBW
from nlo_bk_optimized_amd.dis.config import DISConfig, Polarization, Quark
from nlo_bk_optimized_amd.numerics.integration import IntegralResult

TRAINING_SAMPLES = 65_536
TRAINING_ROUNDS = 3
SAMPLES_PER_REPLICATE = 32_768
REPLICATES = 7
GRID_BINS = 32
FIXED_SOBOL_EVALUATIONS = 2 * (TRAINING_ROUNDS * TRAINING_SAMPLES + REPLICATES * SAMPLES_PER_REPLICATE)


@dataclass(frozen=True, slots=True)
class SobolPointSet:
    training: tuple[torch.Tensor, ...]
    replicates: tuple[torch.Tensor, ...]
    seed: int


@dataclass(frozen=True, slots=True)
class FixedSobolResult:
    contributions: tuple[IntegralResult, IntegralResult, IntegralResult]
    total: IntegralResult


def _check_sizes(training_samples: int, training_rounds: int, samples_per_replicate: int, replicates: int,
                 dimensions: int = 5) -> None:
    """S/dis/fixed_sobol.py:79-85."""

    if dimensions < 1 or training_rounds < 1 or replicates < 2:
        raise ValueError("dimensions and training rounds must be positive; replicates must exceed one")
    for count in (training_samples, samples_per_replicate):
        if count < 2 or count & (count - 1):
            raise ValueError("Sobol sample counts must be powers of two greater than one")


def make_sobol_points(dimensions: int, *, training_samples: int, training_rounds: int, samples_per_replicate: int,
                      replicates: int, device, dtype: torch.dtype, seed: int) -> SobolPointSet:
    """Independently scrambled Sobol point sets (S/dis/fixed_sobol.py:66-97),
    drawn on the device: training round k is SobolEngine(dimensions,
    scramble=True, seed=seed + k).draw_base2(...), replicate k the same with
    seed + 10000 + k."""

    from nlo_bk_optimized_amd import device as dev

    _check_sizes(training_samples, training_rounds, samples_per_replicate, replicates, dimensions)
    if dtype not in (torch.float32, torch.float64):
        raise TypeError("Sobol points are float32 or float64")
    kind = "f32" if dtype == torch.float32 else "f64"

    def draw(count: int, offset: int) -> torch.Tensor:
        return dev.sobol_draw(dev.sobol_table(dimensions, seed + offset), count, kind=kind, device=device)

    training = tuple(draw(training_samples, k) for k in range(training_rounds))
    estimates = tuple(draw(samples_per_replicate, 10_000 + k) for k in range(replicates))
    return SobolPointSet(training, estimates, seed)


def dis_sobol_endpoint_map(points: torch.Tensor, edges: torch.Tensor, maxr: float, fold_angle: bool = False):
    """Map points [n, 5] through a frozen grid [5, bins + 1] and split them into
    the two nearest-endpoint sectors (dis_sobol_endpoint_map_cuda,
    K/dis/sampling/kernel.cu:15-81): (first, first_weight, second,
    second_weight), the weights including the grid's inverse density."""

    from nlo_bk_optimized_amd import device as dev

    if maxr <= 0 or maxr != maxr or maxr == float("inf"):
        raise ValueError("DIS Sobol maxr must be positive and finite")
    return dev.dis_endpoint_map(points.contiguous(), edges.contiguous(), maxr, fold_angle=fold_angle)


def fixed_sobol_supported(Q2, dipole, config: DISConfig) -> bool:
    """_fixed_sobol_supported (S/dis/observables.py:640-658): float32 with
    fused kernels and the 48-point nested rule, a GBW dipole and a budget of at
    least FIXED_SOBOL_EVALUATIONS evaluations per quark."""

    return (isinstance(Q2, torch.Tensor) and Q2.dtype is torch.float32 and config.cuda_fusion
            and config.cuda_nested and config.cuda_nested_points == 48 and isinstance(dipole, GBW)
            and config.maxeval >= FIXED_SOBOL_EVALUATIONS)


@lru_cache(maxsize=64)
def _point_set_tables(seed: int, training_rounds: int, replicates: int) -> tuple:
    """The scrambled direction tables of one seed's point sets (torch's
    SobolEngine scramble runs on the host once per seed)."""

    from nlo_bk_optimized_amd import device as dev

    return tuple(dev.fixed_sobol_tables(seed, training_rounds, replicates))


def _scalar_pair(Q2, xbj) -> str:
    ok = (isinstance(Q2, torch.Tensor) and isinstance(xbj, torch.Tensor) and Q2.numel() == 1 and xbj.numel() == 1
          and Q2.dtype == xbj.dtype and Q2.dtype in (torch.float32, torch.float64))
    if not ok:
        raise ValueError("fixed-Sobol DIS requires matching float32 (or float64) scalars")
    return "fp32" if Q2.dtype == torch.float32 else "fp64"


def gbw_qg_fixed_sobol_batch(entries, dipole: GBW, config: DISConfig, *, training_samples: int = TRAINING_SAMPLES,
                             training_rounds: int = TRAINING_ROUNDS,
                             samples_per_replicate: int = SAMPLES_PER_REPLICATE, replicates: int = REPLICATES,
                             bins: int = GRID_BINS, i1_weight: float = 1.0, fold_angle: bool = True,
                             return_edges: bool = False, grid: torch.Tensor | None = None, device=None):
    """gbw_qg_fixed_sobol for every (Q2, xbj, polarization, quark, seed) of
    ``entries`` in one device call; scalars of one dtype.  With
    ``return_edges`` also the learned grids [n, 5, bins + 1] (float64).  With
    ``grid`` (float64 [n, 5, bins + 1] on the device) the grids are frozen:
    integrate_triple_fixed_sobol alone (S/dis/fixed_sobol.py:213-271), no
    training, evaluations counting the replicates only."""

    from nlo_bk_optimized_amd import _lib
    from nlo_bk_optimized_amd import device as dev

    entries = list(entries)
    if not entries:
        return ([], None) if return_edges else []
    if not isinstance(dipole, GBW):
        raise TypeError("the fixed-Sobol estimator takes a GBW dipole")
    if grid is not None:
        training_rounds = 0
        _check_sizes(training_samples, 1, samples_per_replicate, replicates)
    else:
        _check_sizes(training_samples, training_rounds, samples_per_replicate, replicates)
    precisions = {_scalar_pair(Q2, xbj) for Q2, xbj, *_ in entries}
    if len(precisions) != 1:
        raise ValueError("all scalars of a fixed-Sobol batch must have one dtype")
    precision = precisions.pop()
    if device is None:
        q0 = entries[0][0]
        device = q0.device if q0.is_cuda else torch.device("cuda")
    sets: dict[int, int] = {}
    tables: list = []
    items = []
    seeds = []
    for Q2, xbj, polarization, quark, seed in entries:
        if polarization not in (Polarization.T, Polarization.L):
            raise ValueError(f"unknown polarization: {polarization}")
        if not isinstance(quark, Quark):
            raise TypeError("quark must be a Quark")
        seed = int(seed)
        if seed not in sets:
            sets[seed] = len(sets)
            tables.extend(_point_set_tables(seed, training_rounds, replicates))
        item = _lib.FixedSobolItem()
        item.desc = dev.dis_desc(float(Q2.item()), float(xbj.item()), polarization is Polarization.T, (quark,),
                                 dipole, config)
        item.table_set = sets[seed]
        items.append(item)
        seeds.append(seed)
    res = dev.dis_fixed_sobol(items, tables, training_samples=training_samples, training_rounds=training_rounds,
                              samples_per_replicate=samples_per_replicate, replicates=replicates, bins=bins,
                              precision=precision, epsrel=config.epsrel, i1_weight=i1_weight,
                              fold_angle=fold_angle, device=device, return_edges=return_edges, grid=grid)
    rows, edges = res if return_edges else (res, None)
    rows = rows.cpu().tolist()
    out = []
    for per_item, seed in zip(rows, seeds):
        results = [IntegralResult(torch.tensor(v, dtype=torch.float64), torch.tensor(e, dtype=torch.float64),
                                  int(n), bool(c), seed) for v, e, n, c in per_item]
        out.append(FixedSobolResult(tuple(results[:3]), results[3]))
    return (out, edges) if return_edges else out


def gbw_qg_fixed_sobol(Q2, xbj, polarization: Polarization, quark: Quark, dipole: GBW, config: DISConfig, *,
                       seed: int) -> FixedSobolResult:
    """One quark's complete qqbar-g contribution by the fixed-Sobol estimator
    (S/dis/fixed_sobol.py:274-368)."""

    return gbw_qg_fixed_sobol_batch([(Q2, xbj, polarization, quark, seed)], dipole, config)[0]


__all__ = [
    "FIXED_SOBOL_EVALUATIONS",
    "FixedSobolResult",
    "GRID_BINS",
    "REPLICATES",
    "SAMPLES_PER_REPLICATE",
    "SobolPointSet",
    "TRAINING_ROUNDS",
    "TRAINING_SAMPLES",
    "dis_sobol_endpoint_map",
    "fixed_sobol_supported",
    "gbw_qg_fixed_sobol",
    "gbw_qg_fixed_sobol_batch",
    "make_sobol_points",
]
