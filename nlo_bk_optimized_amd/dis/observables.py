"""DIS cross sections and structure functions on the device
(S/dis/observables.py:1-742).

The composition follows the reference: sigma = sigma_LO(x0) + sigma_dip +
sigma_qg at NLO (sigma_LO at xbj at LO); F_{T,L} = Q^2/(4 pi^2 alpha_em) sigma
x transverse_area; F2 = FT + FL; errors add in quadrature.  Every integrand
value comes from libnlo_gfx950: LO and the two-dimensional dipole terms on the
tensor Gauss-Legendre rule, the others through the device VEGAS driver.  The
arithmetic follows the dtype of Q2 as in the reference: float32 -> the FP32
integrands, float64 -> FP64 (sums are FP64 in both).

The qqbar-g term takes the reference's GBW fixed-Sobol estimator where it
does (float32, fused kernels, the 48-point nested rule, a GBW dipole and
maxeval >= FIXED_SOBOL_EVALUATIONS; S/dis/observables.py:470-491) and falls
back to VEGAS for every quark of the term when one quark's fixed-Sobol total
misses epsrel; otherwise VEGAS over I1 plus, for float32 with cuda_nested, the
nested I2/I3, else the 6-/7-dimensional I2/I3.  A batch runs every
fixed-Sobol estimator in one device call first, then every VEGAS and
Gauss-Legendre integral (fallbacks included) in one device batch.

Differences from the reference, on purpose:
* VEGAS draws scrambled Sobol points (seeded, deterministic) instead of
  torch.Generator samples;
* the fixed-Sobol selection reads the dtype of Q2 only (every path here runs
  on the GPU).
"""

from __future__ import annotations

import math

import torch

from nlo_bk_optimized_amd.dipole.amplitude import GBW, BKDipole
from nlo_bk_optimized_amd.dis.config import (
    CF,
    NC,
    AlphaEM,
    DISConfig,
    DISOrder,
    HeavyQuarkX,
    NcScheme,
    Polarization,
    Q0sqr,
    Quark,
    RunningCouplingScheme,
)
from nlo_bk_optimized_amd.numerics.integration import IntegralResult

ZMIN = 1e-6
Dipole = GBW | BKDipole

_DIP_KINDS = {"T0": "T0", "T1": "T1", "T2": "T2", "Omega_L_const": "Omega_L_const", "ab": "ab", "cd": "cd"}
_DIP_DIMENSIONS = {(Polarization.L, "Omega_L_const"): 2, (Polarization.L, "ab"): 3, (Polarization.L, "cd"): 4,
                   (Polarization.T, "T0"): 2, (Polarization.T, "T1"): 3, (Polarization.T, "T2"): 4}


# ------------------------------------------------------------- kinematics


def evolution_rapidity_qqg(xbj: torch.Tensor, Q2: torch.Tensor, z2: torch.Tensor) -> torch.Tensor:
    """ln(W^2 z2 / Q0^2), W^2 = Q^2 / xbj (S/dis/observables.py:49-52)."""

    return torch.log((Q2 / xbj) * z2 / Q0sqr)


def evolution_rapidity_dipole(xbj: torch.Tensor, Q2: torch.Tensor, mf: float,
                              heavy_quark_x_scheme: HeavyQuarkX) -> torch.Tensor:
    """S/dis/observables.py:55-68."""

    W2 = (1 - xbj) * Q2 / xbj
    if heavy_quark_x_scheme is HeavyQuarkX.MassIndependentX:
        X = Q2 / (Q2 + W2)
    elif heavy_quark_x_scheme is HeavyQuarkX.MassDependentX:
        X = (Q2 + 4 * mf**2) / (Q2 + W2)
    else:
        raise ValueError(f"unknown heavy-quark x scheme: {heavy_quark_x_scheme}")
    return torch.log(1 / X)


def z2_lower_bound(xbj: torch.Tensor, Q2: torch.Tensor) -> torch.Tensor:
    """Q0^2 / W^2 (S/dis/observables.py:71-73)."""

    return Q0sqr / (Q2 / xbj)


def running_coupling_scale(x01: torch.Tensor, x02: torch.Tensor, x21: torch.Tensor,
                           rc_scheme: RunningCouplingScheme) -> torch.Tensor:
    """S/dis/observables.py:76-83."""

    if rc_scheme is RunningCouplingScheme.SMALLEST:
        return torch.minimum(x01, torch.minimum(x02, x21))
    if rc_scheme is RunningCouplingScheme.PARENT:
        return x01
    raise ValueError(f"unknown running-coupling scheme: {rc_scheme}")


def tripole_amplitude(dipole: Dipole, x01: torch.Tensor, x02: torch.Tensor, x21: torch.Tensor, Y: torch.Tensor,
                      nc_scheme: NcScheme) -> torch.Tensor:
    """S/dis/observables.py:86-104 (dipole amplitudes from the device)."""

    S01 = 1 - dipole.dipole_amplitude(x01, Y)
    S02 = 1 - dipole.dipole_amplitude(x02, Y)
    S12 = 1 - dipole.dipole_amplitude(x21, Y)
    if nc_scheme is NcScheme.LargeNC:
        return 1 - S02 * S12
    if nc_scheme is NcScheme.FiniteNC:
        return 1 - NC / (2 * CF) * (S02 * S12 - S01 / NC**2)
    raise ValueError(f"unknown Nc scheme: {nc_scheme}")


# ------------------------------------------------------------- integrands


def _scalar(t) -> float:
    return float(t.item()) if isinstance(t, torch.Tensor) else float(t)


def _precision(reference) -> str:
    return "fp32" if isinstance(reference, torch.Tensor) and reference.dtype is torch.float32 else "fp64"


def _desc(Q2, xbj, polarization: Polarization, quarks, dipole: Dipole, config: DISConfig):
    from nlo_bk_optimized_amd import device as dev

    if polarization not in (Polarization.T, Polarization.L):
        raise ValueError(f"unknown polarization: {polarization}")
    return dev.dis_desc(_scalar(Q2), _scalar(xbj), polarization is Polarization.T, quarks, dipole, config)


def _device_of(dipole: Dipole, reference) -> torch.device:
    if isinstance(dipole, BKDipole):
        return dipole.device
    if isinstance(reference, torch.Tensor) and reference.is_cuda:
        return reference.device
    return torch.device("cuda")


def _points(x: torch.Tensor, device: torch.device) -> torch.Tensor:
    return x.detach().to(device=device, dtype=torch.float64).contiguous()


def _evaluate(kind: str, x, Q2, xbj, polarization, quarks, dipole, config) -> torch.Tensor:
    from nlo_bk_optimized_amd import device as dev

    device = _device_of(dipole, x)
    out = dev.dis_integrand(_desc(Q2, xbj, polarization, quarks, dipole, config), kind, _points(x, device),
                            precision=_precision(Q2))
    return out.to(dtype=x.dtype if x.is_floating_point() else torch.float64)


def integrand_photon_target_LO_unit(x: torch.Tensor, Q2, xbj, polarization: Polarization, quarks,
                                    dipole: Dipole, config: DISConfig) -> torch.Tensor:
    """2 pi r maxr |Psi|^2 N at unit-square points (z, r / maxr)
    (S/dis/observables.py:345-352 with S/dis/lo.py:12-51)."""

    return _evaluate("LO", x, Q2, xbj, polarization, tuple(quarks), dipole, config)


def integrand_dip_massive(x: torch.Tensor, Q2, xbj, polarization: Polarization, quark: Quark, contribution: str,
                          dipole: Dipole, config: DISConfig) -> torch.Tensor:
    """S/dis/observables.py:107-172."""

    dims = _DIP_DIMENSIONS.get((polarization, contribution))
    if dims is None or x.shape[1] != dims:
        raise ValueError(f"invalid contribution/dimension pair: {polarization}, {contribution}, {x.shape[1]}")
    return _evaluate(contribution, x, Q2, xbj, polarization, (quark,), dipole, config)


def integrand_qgunsub_massive(x: torch.Tensor, Q2, xbj, polarization: Polarization, quark: Quark, contribution: str,
                              dipole: Dipole, config: DISConfig) -> torch.Tensor:
    """S/dis/observables.py:175-250."""

    dims = {"I1": 5, "I2": 6, "I3": 7}.get(contribution)
    if dims is None or x.shape[1] != dims:
        raise ValueError(f"invalid qqg contribution/dimension pair: {contribution}, {x.shape[1]}")
    return _evaluate(contribution, x, Q2, xbj, polarization, (quark,), dipole, config)


def integrand_qg_nested_massive(x: torch.Tensor, Q2, xbj, polarization: Polarization, quark: Quark,
                                contribution: str, dipole: Dipole, config: DISConfig) -> torch.Tensor:
    """S/dis/observables.py:253-331: I2 or I3 with the y_t integrals done on
    config.cuda_nested_points mapped Gauss-Legendre nodes."""

    if contribution not in ("I2", "I3") or x.shape[1] != 5:
        raise ValueError(f"invalid nested qqg contribution/dimension pair: {contribution}, {x.shape[1]}")
    return _evaluate(contribution + "_nested", x, Q2, xbj, polarization, (quark,), dipole, config)


def integrand_qg_nested_pair(x: torch.Tensor, Q2, xbj, polarization: Polarization, quark: Quark, dipole: Dipole,
                             config: DISConfig) -> tuple[torch.Tensor, torch.Tensor]:
    """(I2, I3) of integrand_qg_nested_massive at the same points from one pass
    over the inner rule (the pair dis_nested_gbw_i2_i3_integrand_cuda returns,
    K/dis/i2_i3/extension.py)."""

    from nlo_bk_optimized_amd import device as dev

    if x.ndim != 2 or x.shape[1] != 5:
        raise ValueError(f"the nested I2 / I3 need points of shape (n, 5), got {tuple(x.shape)}")
    device = _device_of(dipole, x)
    i2, i3 = dev.dis_qg_i23(_desc(Q2, xbj, polarization, (quark,), dipole, config), _points(x, device),
                            precision=_precision(Q2))
    dtype = x.dtype if x.is_floating_point() else torch.float64
    return i2.to(dtype), i3.to(dtype)


# --------------------------------------------------------- cross sections


def _validate_kinematics(Q2, xbj) -> None:
    """S/dis/observables.py:726-735."""

    for name, v in (("Q2", Q2), ("xbj", xbj)):
        if isinstance(v, torch.Tensor) and v.numel() != 1:
            raise ValueError("Q2 and xbj must be scalar tensors")
    q, x = _scalar(Q2), _scalar(xbj)
    if not (math.isfinite(q) and math.isfinite(x)):
        raise ValueError("Q2 and xbj must be finite")
    if not (q > 0 and 0 < x <= 1):
        raise ValueError("Q2 must be positive and xbj must lie in (0, 1]")


def _scale_result(result: IntegralResult, factor: float) -> IntegralResult:
    return IntegralResult(result.value * factor, result.error * abs(factor), result.n_eval, result.converged,
                          result.seed)


def _sum_results(results, seed) -> IntegralResult:
    """Values add, errors add in quadrature (S/dis/observables.py:706-723)."""

    if not results:
        raise ValueError("at least one integration result is required")
    value = torch.stack([r.value for r in results]).sum()
    error = torch.sqrt(torch.stack([r.error.square() for r in results]).sum())
    return IntegralResult(value, error, sum(r.n_eval for r in results), all(r.converged for r in results), seed)


class _Plan:
    """The integrals of one or more cross sections, collected in the order the
    per-point composition of the reference makes them and run as one device
    batch (nlo_dis_batch_run): every VEGAS integral keeps its own grid, seed,
    sample budget and iteration schedule, so the results do not depend on
    what else is in the batch."""

    def __init__(self, dipole: Dipole, config: DISConfig, precision: str, device: torch.device) -> None:
        self.dipole = dipole
        self.config = config
        self.precision = precision
        self.device = device
        self.items: list = []
        self.seeds: list[int | None] = []
        self.tables: list = []
        self._table_index: dict[tuple[int, int], int] = {}
        self._keep: list[torch.Tensor] = []  # device arrays the items point to (BK initialisations)
        self.fixed_entries: list[tuple] = []  # fixed-Sobol estimators (Q2, xbj, polarization, quark, seed)
        self._deferred: list = []  # called with the fixed-Sobol results before the batch runs

    def _desc(self, Q2, xbj, polarization: Polarization, quarks):
        desc = _desc(Q2, xbj, polarization, tuple(quarks), self.dipole, self.config)
        coef = getattr(self.dipole, "device_coef", None)
        if coef is not None:
            self._keep.append(coef)
        return desc

    def _add(self, item, seed) -> int:
        self.items.append(item)
        self.seeds.append(seed)
        return len(self.items) - 1

    def gauss_legendre(self, Q2, xbj, polarization, quarks, kind: str, points: int) -> int:
        """tensor_gauss_legendre (S/numerics/integration.py:95-117)."""

        from nlo_bk_optimized_amd import _lib

        item = _lib.DisItem()
        item.desc = self._desc(Q2, xbj, polarization, quarks)
        item.contribution = _lib.DIS_KINDS[kind]
        item.method = _lib.NLO_DIS_METHOD_GAUSS_LEGENDRE
        item.n = int(points)
        return self._add(item, None)

    def vegas(self, Q2, xbj, polarization, quarks, kind: str, maxeval: int, seed: int) -> int:
        """_vegas_with_budget (S/dis/observables.py:661-686): up to 8 iterations
        of maxeval / iterations samples, at least three before convergence."""

        from nlo_bk_optimized_amd import _lib
        from nlo_bk_optimized_amd import device as dev

        if maxeval < 2:
            raise ValueError("NLO integration requires maxeval >= 2")
        max_iterations = min(8, max(1, maxeval // 2))
        samples = max(2, maxeval // max_iterations)
        dim = _lib.DIS_DIMENSIONS[kind]
        key = (dim, int(seed))
        if key not in self._table_index:
            self._table_index[key] = len(self.tables)
            self.tables.append(dev.sobol_table(dim, int(seed)))
        stride = (samples + 63) // 64 * 64
        first = self.config.sobol_first_index
        if first + stride * max_iterations > 1 << _lib.SOBOL_BITS:
            raise ValueError("the Sobol sequence (2^30 points) is too short for this budget")
        item = _lib.DisItem()
        item.desc = self._desc(Q2, xbj, polarization, quarks)
        item.contribution = _lib.DIS_KINDS[kind]
        item.method = _lib.NLO_DIS_METHOD_VEGAS
        item.n = samples
        item.sobol_table = self._table_index[key]
        item.max_iterations = max_iterations
        item.min_iterations = min(3, max_iterations)
        item.first_index = first
        return self._add(item, int(seed))

    def fixed_sobol(self, Q2, xbj, polarization, quark, seed: int) -> int:
        """gbw_qg_fixed_sobol (S/dis/fixed_sobol.py:274-368) for one quark."""

        self.fixed_entries.append((Q2, xbj, polarization, quark, int(seed)))
        return len(self.fixed_entries) - 1

    def defer(self, decide) -> None:
        """Call decide(fixed results) once the fixed-Sobol estimators have run,
        before the batch; it may add integrals (the VEGAS fallback)."""

        self._deferred.append(decide)

    def run(self, workspace_budget: int = 8 << 30) -> "_Results":
        """Every collected integral as an IntegralResult, in collection order,
        with the fixed-Sobol results in ``.fixed``."""

        from nlo_bk_optimized_amd import device as dev
        from nlo_bk_optimized_amd.dis.fixed_sobol import gbw_qg_fixed_sobol_batch

        fixed = (gbw_qg_fixed_sobol_batch(self.fixed_entries, self.dipole, self.config, device=self.device)
                 if self.fixed_entries else [])
        for decide in self._deferred:
            decide(fixed)
        if not self.items:
            return _Results([], fixed)
        rows = dev.dis_batch(self.items, self.tables, bins=self.config.bins, precision=self.precision,
                             epsrel=self.config.epsrel, epsabs=0.0, device=self.device,
                             workspace_budget=workspace_budget).cpu().tolist()
        return _Results([IntegralResult(torch.tensor(v, dtype=torch.float64), torch.tensor(e, dtype=torch.float64),
                                        int(n), bool(c), seed) for (v, e, n, c), seed in zip(rows, self.seeds)],
                        fixed)


class _Results(list):
    """The batch's IntegralResults in collection order; ``fixed`` holds the
    FixedSobolResults in the order the estimators were added."""

    def __init__(self, results, fixed) -> None:
        super().__init__(results)
        self.fixed = fixed


def _lo_recipe(plan: _Plan, Q2, xbj, polarization: Polarization, quadrature_points: int):
    """S/dis/observables.py:334-385."""

    dipole, config = plan.dipole, plan.config
    _validate_kinematics(Q2, xbj)
    Y = max(math.log(dipole.X0() / _scalar(xbj)), 0.0)
    if isinstance(dipole, BKDipole):
        dipole.initialize_interpolation(Y)
    k = plan.gauss_legendre(Q2, xbj, polarization, tuple(config.quarks), "LO", quadrature_points)
    return lambda R: _scale_result(R[k], 4 * AlphaEM * NC / (2 * math.pi) ** 2)


def _dip_recipe(plan: _Plan, Q2, xbj, polarization: Polarization, quadrature_points: int, seed: int):
    """S/dis/observables.py:388-460."""

    config = plan.config
    _validate_kinematics(Q2, xbj)
    contributions = ((("Omega_L_const", 2), ("ab", 3), ("cd", 4)) if polarization is Polarization.L
                     else (("T0", 2), ("T1", 3), ("T2", 4)))
    per_quark = []
    offset = 0
    for quark in config.quarks:
        maxeval = config.maxeval * (10 if quark.mass < 0.4 else 1)
        terms = []
        for contribution, dims in contributions:
            if dims == 2:
                k = plan.gauss_legendre(Q2, xbj, polarization, (quark,), contribution, quadrature_points)
            else:
                k = plan.vegas(Q2, xbj, polarization, (quark,), contribution, maxeval, seed + offset)
                offset += 1
            terms.append((k, quark.charge**2))
        per_quark.append(terms)
    factor = 4 * NC * AlphaEM / (2 * math.pi) ** 2 * 2 * math.pi

    def compose(R):
        results = [_sum_results([_scale_result(R[k], c) for k, c in terms], seed) for terms in per_quark]
        return _scale_result(_sum_results(results, seed), factor)

    return compose


def _fixed_sobol_supported(Q2, dipole: Dipole, config: DISConfig) -> bool:
    """S/dis/observables.py:644-658 (see fixed_sobol.fixed_sobol_supported)."""

    from nlo_bk_optimized_amd.dis.fixed_sobol import fixed_sobol_supported

    return fixed_sobol_supported(Q2, dipole, config)


def _qg_recipe(plan: _Plan, Q2, xbj, polarization: Polarization, seed: int):
    """sigma_qg_d2b (S/dis/observables.py:459-532): the fixed-Sobol estimator
    per quark (seed + quark index) where it applies, its totals scaled by the
    squared charges and summed when every quark's total converged; else, and
    where it does not apply, the VEGAS path."""

    config = plan.config
    _validate_kinematics(Q2, xbj)
    if not _fixed_sobol_supported(Q2, plan.dipole, config):
        return _qg_vegas_recipe(plan, Q2, xbj, polarization, seed)
    handles = [plan.fixed_sobol(Q2, xbj, polarization, quark, seed + index)
               for index, quark in enumerate(config.quarks)]
    fallback = []

    def decide(fixed) -> None:
        # the reference stops at the first quark whose total misses epsrel and
        # integrates every quark of the term with VEGAS instead
        if not all(fixed[h].total.converged for h in handles):
            fallback.append(_qg_vegas_recipe(plan, Q2, xbj, polarization, seed))

    plan.defer(decide)
    factor = 4 * NC * AlphaEM / (2 * math.pi) ** 3 * 2 * math.pi

    def compose(R):
        if fallback:
            return fallback[0](R)
        results = [_scale_result(R.fixed[h].total, quark.charge**2) for h, quark in zip(handles, config.quarks)]
        return _scale_result(_sum_results(results, seed), factor)

    return compose


def _qg_vegas_recipe(plan: _Plan, Q2, xbj, polarization: Polarization, seed: int):
    """S/dis/observables.py:493-532 (VEGAS path)."""

    config = plan.config
    # the reference nests the I2 / I3 inner integral only on its float32 CUDA
    # path (use_nested, S/dis/observables.py:492); float64 integrates the
    # 6- / 7-dimensional forms
    kinds = (("I1", "I2_nested", "I3_nested") if config.cuda_nested and plan.precision == "fp32"
             else ("I1", "I2", "I3"))
    per_quark = []
    offset = 0
    for quark in config.quarks:
        terms = []
        for kind in kinds:
            k = plan.vegas(Q2, xbj, polarization, (quark,), kind, config.maxeval, seed + offset)
            offset += 1
            terms.append((k, quark.charge**2))
        per_quark.append(terms)
    factor = 4 * NC * AlphaEM / (2 * math.pi) ** 3 * 2 * math.pi

    def compose(R):
        results = [_sum_results([_scale_result(R[k], c) for k, c in terms], seed) for terms in per_quark]
        return _scale_result(_sum_results(results, seed), factor)

    return compose


def _cross_section_recipe(plan: _Plan, Q2, xbj, polarization: Polarization, quadrature_points: int, seed: int):
    """S/dis/observables.py:541-578."""

    config, dipole = plan.config, plan.dipole
    if config.order is DISOrder.LO:
        return _lo_recipe(plan, Q2, xbj, polarization, quadrature_points)
    x0 = torch.tensor(dipole.X0(), dtype=Q2.dtype) if isinstance(Q2, torch.Tensor) else dipole.X0()
    lo = _lo_recipe(plan, Q2, x0, polarization, quadrature_points)
    dip = _dip_recipe(plan, Q2, xbj, polarization, quadrature_points, seed)
    qg = _qg_recipe(plan, Q2, xbj, polarization, seed + 100)
    return lambda R: _sum_results((lo(R), dip(R), qg(R)), seed)


def _structure_recipe(plan: _Plan, Q2, xbj, polarization: Polarization, quadrature_points: int, seed: int):
    """F_{T,L} = Q^2 / (4 pi^2 alpha_em) sigma x transverse area (S/dis/observables.py:581-610)."""

    sigma = _cross_section_recipe(plan, Q2, xbj, polarization, quadrature_points, seed)
    factor = _scalar(Q2) / (4 * math.pi**2 * AlphaEM) * plan.config.transverse_area
    return lambda R: _scale_result(sigma(R), factor)


def _plan(Q2, dipole: Dipole, config: DISConfig) -> _Plan:
    return _Plan(dipole, config, _precision(Q2), _device_of(dipole, Q2))


def photon_proton_cross_section_LO_d2b(Q2, xbj, polarization: Polarization, dipole: Dipole, config: DISConfig, *,
                                       quadrature_points: int = 24, batch_size: int = 65_536) -> IntegralResult:
    """S/dis/observables.py:334-385."""

    plan = _plan(Q2, dipole, config)
    recipe = _lo_recipe(plan, Q2, xbj, polarization, quadrature_points)
    return recipe(plan.run())


def sigma_dip_d2b(Q2, xbj, polarization: Polarization, dipole: Dipole, config: DISConfig, *,
                  quadrature_points: int = 24, batch_size: int = 65_536, seed: int = 0) -> IntegralResult:
    """S/dis/observables.py:388-460."""

    plan = _plan(Q2, dipole, config)
    recipe = _dip_recipe(plan, Q2, xbj, polarization, quadrature_points, seed)
    return recipe(plan.run())


def sigma_qg_d2b(Q2, xbj, polarization: Polarization, dipole: Dipole, config: DISConfig, *,
                 batch_size: int = 65_536, seed: int = 0) -> IntegralResult:
    """S/dis/observables.py:459-532."""

    plan = _plan(Q2, dipole, config)
    recipe = _qg_recipe(plan, Q2, xbj, polarization, seed)
    return recipe(plan.run())


def photon_proton_cross_section_d2b(Q2, xbj, polarization: Polarization, dipole: Dipole, config: DISConfig, *,
                                    quadrature_points: int = 24, batch_size: int = 65_536,
                                    seed: int = 0) -> IntegralResult:
    """S/dis/observables.py:541-578."""

    plan = _plan(Q2, dipole, config)
    recipe = _cross_section_recipe(plan, Q2, xbj, polarization, quadrature_points, seed)
    return recipe(plan.run())


def FL(Q2, xbj, dipole: Dipole, config: DISConfig, *, quadrature_points: int = 24, batch_size: int = 65_536,
       seed: int = 0) -> IntegralResult:
    """S/dis/observables.py:581-594."""

    plan = _plan(Q2, dipole, config)
    recipe = _structure_recipe(plan, Q2, xbj, Polarization.L, quadrature_points, seed)
    return recipe(plan.run())


def FT(Q2, xbj, dipole: Dipole, config: DISConfig, *, quadrature_points: int = 24, batch_size: int = 65_536,
       seed: int = 0) -> IntegralResult:
    """S/dis/observables.py:597-610."""

    plan = _plan(Q2, dipole, config)
    recipe = _structure_recipe(plan, Q2, xbj, Polarization.T, quadrature_points, seed)
    return recipe(plan.run())


def F2(Q2, xbj, dipole: Dipole, config: DISConfig, **integration_options) -> IntegralResult:
    """FT + FL, the L integrals seeded 1000 later (S/dis/observables.py:613-624)."""

    return F2_batch([(Q2, xbj)], dipole, config, **integration_options)[0]


def sigma_r(Q2, xbj, sqrt_s: float, dipole: Dipole, config: DISConfig, **integration_options) -> IntegralResult:
    """Reduced cross section F2 - y^2 / (1 + (1 - y)^2) FL, y = Q^2 / (xbj s).
    Not in the reference; built from its F2 and FL."""

    return sigma_r_batch([(Q2, xbj)], sqrt_s, dipole, config, **integration_options)[0]


def structure_functions_batch(points, dipole: Dipole, config: DISConfig, *, quadrature_points: int = 24,
                              batch_size: int = 65_536, seed: int | None = None,
                              workspace_budget: int = 8 << 30) -> list[tuple[IntegralResult, IntegralResult]]:
    """(FT, FL) at every (Q^2, xbj) of ``points`` with F2's seeds (a given
    seed for FT and seed + 1000 for FL; without one, both take the default 0),
    all integrals of all points in one device batch.  Each point's results
    equal FT / FL called for that point alone.  Q^2 values of one dtype:
    float32 tensors take the FP32 integrands."""

    points = list(points)
    if not points:
        return []
    s0 = 0 if seed is None else seed
    s_l = 0 if seed is None else seed + 1_000
    plan = _plan(points[0][0], dipole, config)
    recipes = []
    for Q2, xbj in points:
        if _precision(Q2) != plan.precision:
            raise ValueError("all Q2 of a batch must have one dtype")
        ft = _structure_recipe(plan, Q2, xbj, Polarization.T, quadrature_points, s0)
        fl = _structure_recipe(plan, Q2, xbj, Polarization.L, quadrature_points, s_l)
        recipes.append((ft, fl))
    R = plan.run(workspace_budget)
    return [(ft(R), fl(R)) for ft, fl in recipes]


def F2_batch(points, dipole: Dipole, config: DISConfig, **integration_options) -> list[IntegralResult]:
    """F2 = FT + FL at every (Q^2, xbj) of ``points`` in one device batch
    (S/dis/observables.py:613-624 per point)."""

    seed = integration_options.get("seed")
    return [_sum_results((ft, fl), seed)
            for ft, fl in structure_functions_batch(points, dipole, config, **integration_options)]


def sigma_r_batch(points, sqrt_s: float, dipole: Dipole, config: DISConfig,
                  **integration_options) -> list[IntegralResult]:
    """sigma_r = FT + (1 - y^2 / (1 + (1 - y)^2)) FL at every point, y = Q^2 / (xbj s),
    with F2's seeds, in one device batch."""

    points = list(points)
    ys = []
    for Q2, xbj in points:
        y = _scalar(Q2) / (_scalar(xbj) * sqrt_s**2)
        if not 0 < y <= 1:
            raise ValueError(f"inelasticity y = {y} outside (0, 1]")
        ys.append(y)
    out = []
    for (ft, fl), y in zip(structure_functions_batch(points, dipole, config, **integration_options), ys):
        c = 1 - y * y / (1 + (1 - y) ** 2)
        out.append(IntegralResult(ft.value + c * fl.value, torch.sqrt(ft.error**2 + (c * fl.error) ** 2),
                                  ft.n_eval + fl.n_eval, ft.converged and fl.converged,
                                  integration_options.get("seed")))
    return out


__all__ = [
    "F2",
    "F2_batch",
    "FL",
    "FT",
    "evolution_rapidity_dipole",
    "evolution_rapidity_qqg",
    "integrand_dip_massive",
    "integrand_photon_target_LO_unit",
    "integrand_qg_nested_massive",
    "integrand_qg_nested_pair",
    "integrand_qgunsub_massive",
    "photon_proton_cross_section_LO_d2b",
    "photon_proton_cross_section_d2b",
    "running_coupling_scale",
    "sigma_dip_d2b",
    "sigma_qg_d2b",
    "sigma_r",
    "sigma_r_batch",
    "structure_functions_batch",
    "tripole_amplitude",
    "z2_lower_bound",
]
