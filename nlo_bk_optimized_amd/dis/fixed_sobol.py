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
* the generic forms over Python callables (``map_importance_grid``,
  ``paired_endpoint_inputs``, ``learn_triple_importance_grid``,
  ``integrate_triple_fixed_sobol``) keep the reference's signatures and run
  their map, endpoint split, histogram and refinement in the library's kernels,
  the callables on the caller's device tensors; ``gbw_qg_fixed_sobol`` itself
  does not go through them (its integrands are fused into the passes);
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


def map_importance_grid(points: torch.Tensor, edges: torch.Tensor):
    """S/dis/fixed_sobol.py:100-113 on the device (nlo_vegas_map): (source,
    inverse density, bin index) of points [n, dim] through edges
    [dim, bins + 1], in the points' dtype (the map runs in FP64)."""

    from nlo_bk_optimized_amd import device as dev

    u = points.to(torch.float64).contiguous()
    e = edges.to(device=u.device, dtype=torch.float64).contiguous()
    dim = u.shape[1]
    zero = torch.zeros(dim, dtype=torch.float64, device=u.device)
    x, jac, b = dev.vegas_map(u, e, zero, torch.ones_like(zero))
    return x.to(points.dtype), jac.to(points.dtype), b.to(torch.long)


def _unit_grid(dtype: torch.dtype, device) -> torch.Tensor:
    # two bins on [0, 1]: the map is the identity bit for bit and the inverse
    # density is 1
    return torch.tensor([0.0, 0.5, 1.0], dtype=dtype, device=device).expand(5, -1).contiguous()


def paired_endpoint_inputs(source: torch.Tensor, maxr: float, *, fold_angle: bool = False):
    """S/dis/fixed_sobol.py:116-158 on the device: (first, first weight,
    second, second weight) of mapped points [n, 5]."""

    return dis_sobol_endpoint_map(source, _unit_grid(source.dtype, source.device), maxr, fold_angle=fold_angle)


@torch.no_grad()
def learn_triple_importance_grid(function, training_points, *, maxr: float, bins: int = GRID_BINS,
                                 fold_angle: bool = True, i1_weight: float = 1.0) -> torch.Tensor:
    """S/dis/fixed_sobol.py:161-210 with a caller's ``function(x) -> (regular
    I1, sensitive I1, I2, I3)``: one grid learned from the normalised
    magnitudes over the training sets; the histograms (nlo_vegas_histogram,
    fixed-order sums) and the refinement (nlo_vegas_adapt, the points' epsilon)
    in the library."""

    from nlo_bk_optimized_amd import device as dev

    if not training_points:
        raise ValueError("at least one independent training point set is required")
    reference = training_points[0]
    dims = reference.shape[1]
    edges = torch.linspace(0, 1, bins + 1, device=reference.device, dtype=torch.float64).expand(dims, -1).contiguous()
    eps = float(torch.finfo(reference.dtype).eps)
    tiny = torch.finfo(reference.dtype).tiny
    for points in training_points:
        source, inverse_density, bin_index = map_importance_grid(points, edges)
        first, first_weight, second, second_weight = paired_endpoint_inputs(source, maxr, fold_angle=fold_angle)
        first_regular, first_sensitive, first_I2, first_I3 = function(first)
        second_regular, second_sensitive, second_I2, second_I3 = function(second)
        first_values = first_regular + first_sensitive.float(), first_I2, first_I3
        second_values = second_regular + second_sensitive.float(), second_I2, second_I3
        weighted = tuple((a * first_weight + b * second_weight) * inverse_density
                         for a, b in zip(first_values, second_values, strict=True))
        score = sum(weight * value.abs() / value.abs().mean().clamp_min(tiny)
                    for weight, value in zip((i1_weight, 1.0, 1.0), weighted, strict=True))
        histogram, bin_count = dev.vegas_histogram(bin_index.to(torch.uint8).contiguous(),
                                                   score.to(torch.float64).contiguous(), bins)
        edges = dev.vegas_adapt(edges, histogram, bin_count, eps)
    return edges.to(reference.dtype)


@torch.no_grad()
def integrate_triple_fixed_sobol(i1_sum_function, nested_function, edges: torch.Tensor, replicate_points, *,
                                 maxr: float, epsrel: float, seed: int, training_evaluations: int = 0):
    """S/dis/fixed_sobol.py:213-271 with a caller's ``i1_sum_function(x,
    weights) -> (regular sums, sensitive sums)`` and ``nested_function(x) ->
    (I2, I3)``: the replicate estimates of I1, I2, I3 and of their correlated
    sum; the endpoint map of every replicate in the library."""

    import math

    if len(replicate_points) < 2:
        raise ValueError("at least two scrambled Sobol replicates are required")
    estimates = []
    for points in replicate_points:
        first, first_weight, second, second_weight = dis_sobol_endpoint_map(points.contiguous(), edges, maxr,
                                                                            fold_angle=True)
        first_regular_sums, first_sensitive_sums = i1_sum_function(first, first_weight)
        second_regular_sums, second_sensitive_sums = i1_sum_function(second, second_weight)
        first_I2, first_I3 = nested_function(first)
        second_I2, second_I3 = nested_function(second)
        sample_count = points.shape[0]
        regular_I1 = (first_regular_sums.sum() + second_regular_sums.sum()) / sample_count
        sensitive_I1 = ((first_sensitive_sums.sum() + second_sensitive_sums.sum()) / sample_count).float()
        estimates.append(torch.stack((
            (regular_I1 + sensitive_I1).double(),
            (first_I2 * first_weight + second_I2 * second_weight).double().mean(),
            (first_I3 * first_weight + second_I3 * second_weight).double().mean(),
        )))
    stacked = torch.stack(estimates)
    samples = training_evaluations + 2 * sum(p.shape[0] for p in replicate_points)

    def result(values: torch.Tensor) -> IntegralResult:
        value = values.mean()
        error = values.std(unbiased=True) / math.sqrt(values.numel())
        return IntegralResult(value, error, samples, bool(error <= epsrel * value.abs()), seed)

    contributions = tuple(result(stacked[:, k]) for k in range(3))
    return FixedSobolResult(contributions, result(stacked.sum(dim=1)))


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
    "integrate_triple_fixed_sobol",
    "learn_triple_importance_grid",
    "make_sobol_points",
    "map_importance_grid",
    "paired_endpoint_inputs",
]
