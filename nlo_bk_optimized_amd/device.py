"""Torch-facing wrappers of the C ABI (one function per entry point).

Each wrapper checks shapes and dtypes on the host, allocates outputs and
workspace with torch on the tensors' device, and launches on the current
stream.  Reference prefixes: S/ = torch_nlo_clean/src/nlo_torch/,
K/ = S/custom_kernels/.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from nlo_bk_optimized_amd import _lib
from nlo_bk_optimized_amd._lib import check, lib, ptr, require_cuda, stream_handle

# ----------------------------------------------------------------- Sobol


def sobol_table(dim: int, seed: int) -> _lib.SobolTable:
    """Scrambled direction numbers of ``SobolEngine(dim, scramble=True, seed=seed)``.

    torch computes the scramble (S/dis/fixed_sobol.py:87-93 draws from the
    same engine); the library runs the gray-code recurrence.  ``first`` is the
    integer form of the engine's float64 first point, float32(shift / 2^30)
    widened (torch/quasirandom.py ``_first_point``).
    """

    if not 1 <= dim <= _lib.SOBOL_MAXDIM:
        raise ValueError(f"dim must lie in [1, {_lib.SOBOL_MAXDIM}]")
    engine = torch.quasirandom.SobolEngine(dim, scramble=True, seed=seed)
    direction = engine.sobolstate.numpy().astype(np.uint32)
    shift = engine.shift.numpy().astype(np.uint32)
    first = (shift.astype(np.float64) / 2.0**30).astype(np.float32).astype(np.float64)
    first_q = np.minimum(np.round(first * 2.0**30), 2.0**30).astype(np.uint32)
    t = _lib.SobolTable()
    t.dim = dim
    for d in range(dim):
        for b in range(_lib.SOBOL_BITS):
            t.direction[d][b] = int(direction[d, b])
        t.shift[d] = int(shift[d])
        t.first[d] = int(first_q[d])
    return t


def sobol_draw(
    table: _lib.SobolTable,
    n: int,
    *,
    first_index: int = 0,
    kind: str = "f64",
    device: torch.device | str = "cuda",
) -> torch.Tensor:
    """Points [first_index, first_index + n) as SobolEngine draws them."""

    kinds = {"u32": (_lib.NLO_SOBOL_U32, torch.int32), "f32": (_lib.NLO_SOBOL_F32, torch.float32),
             "f64": (_lib.NLO_SOBOL_F64, torch.float64)}
    if kind not in kinds:
        raise ValueError(f"kind must be one of {sorted(kinds)}")
    code, dtype = kinds[kind]
    out = torch.empty((n, table.dim), dtype=dtype, device=device)
    check(
        lib().nlo_sobol_draw(ctypes.byref(table), first_index, n, code, ptr(out), stream_handle(out.device)),
        "nlo_sobol_draw",
    )
    return out


# ----------------------------------------------------------------- VEGAS


def vegas_map(
    u: torch.Tensor, edges: torch.Tensor, lo: torch.Tensor, width: torch.Tensor
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """(x, jacobian, bins) of unit-cube points u[n, dim] through edges[dim, B + 1]."""

    require_cuda(u, torch.float64, "u")
    require_cuda(edges, torch.float64, "edges")
    require_cuda(lo, torch.float64, "lo")
    require_cuda(width, torch.float64, "width")
    n, dim = u.shape
    bins = edges.shape[1] - 1
    if edges.shape != (dim, bins + 1) or lo.shape != (dim,) or width.shape != (dim,):
        raise ValueError("edges must be [dim, bins + 1] and lo, width [dim]")
    x = torch.empty_like(u)
    jac = torch.empty(n, dtype=torch.float64, device=u.device)
    b = torch.empty((n, dim), dtype=torch.uint8, device=u.device)
    check(
        lib().nlo_vegas_map(dim, bins, n, ptr(u), ptr(edges), ptr(lo), ptr(width), ptr(x), ptr(jac),
                            ptr(b), stream_handle(u.device)),
        "nlo_vegas_map",
    )
    return x, jac, b


def vegas_histogram(bins_idx: torch.Tensor, absw: torch.Tensor, bins: int) -> tuple[torch.Tensor, torch.Tensor]:
    """Deterministic per-bin sum |w| and hit counts, each [dim, bins]."""

    require_cuda(bins_idx, torch.uint8, "bins_idx")
    require_cuda(absw, torch.float64, "absw")
    n, dim = bins_idx.shape
    if absw.shape != (n,):
        raise ValueError("absw must hold one weight per sample")
    hist = torch.empty((dim, bins), dtype=torch.float64, device=absw.device)
    count = torch.empty_like(hist)
    nbytes = lib().nlo_vegas_histogram_workspace(dim, bins, n)
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=absw.device)
    check(
        lib().nlo_vegas_histogram(dim, bins, n, ptr(bins_idx), ptr(absw), ptr(hist), ptr(count),
                                  ptr(ws), nbytes, stream_handle(absw.device)),
        "nlo_vegas_histogram",
    )
    return hist, count


def vegas_adapt(edges: torch.Tensor, hist: torch.Tensor, count: torch.Tensor, eps: float) -> torch.Tensor:
    """Refined edges for grids [..., dim, B + 1] (S/numerics/integration.py:804-866)."""

    require_cuda(edges, torch.float64, "edges")
    require_cuda(hist, torch.float64, "hist")
    require_cuda(count, torch.float64, "count")
    bins = edges.shape[-1] - 1
    dim = edges.shape[-2]
    rows = edges.numel() // (dim * (bins + 1))
    if hist.shape != edges.shape[:-1] + (bins,) or count.shape != hist.shape:
        raise ValueError("hist and count must be [..., dim, bins] matching edges")
    out = torch.empty_like(edges)
    check(
        lib().nlo_vegas_adapt(rows, dim, bins, ptr(edges), ptr(hist), ptr(count), eps, ptr(out),
                              stream_handle(edges.device)),
        "nlo_vegas_adapt",
    )
    return out


# ------------------------------------------------------------------- BK


def physics(config) -> _lib.Physics:
    p = _lib.Physics()
    p.nc = float(config.NC)
    p.nf = float(config.NF)
    p.lambda_qcd = float(config.LambdaQCD)
    p.c2 = float(config.C2)
    p.ksub = float(config.KSUB)
    p.minr = float(config.MINR)
    p.symmetrize = int(bool(config.SYMMETRIZE_Z_Z2_INTEGRATION))
    p.order = _lib.ORDER_CODES[config.Order.name]
    p.rc_lo = _lib.RC_LO_CODES[config.RC_LO.name]
    p.rc_nlo = _lib.RC_NLO_CODES[config.RC_NLO.name]
    p.resum_rc = _lib.RESUM_RC_CODES[config.RESUM_RC.name]
    p.fixed_as = float(config.FIXED_AS)
    return p


def spline_build(r: torch.Tensor, N: torch.Tensor, force_positive: bool = True) -> tuple[torch.Tensor, torch.Tensor]:
    """FP64 and FP32 log-log spline tables of N and S = 1 - N."""

    require_cuda(r, torch.float64, "r")
    require_cuda(N, torch.float64, "N")
    R = r.numel()
    t64 = torch.empty(_lib.spline_words(R), dtype=torch.float64, device=r.device)
    t32 = torch.empty(_lib.spline_words(R), dtype=torch.float32, device=r.device)
    check(
        lib().nlo_bk_spline_build(R, ptr(r), ptr(N), int(force_positive), ptr(t64), ptr(t32),
                                  stream_handle(r.device)),
        "nlo_bk_spline_build",
    )
    return t64, t32


K1_RULE_CODES = {"fixed": _lib.NLO_K1_FIXED, "adaptive": _lib.NLO_K1_ADAPTIVE}


def _history_rows(history, R: int, device) -> torch.Tensor:
    hy, hN = history
    require_cuda(hy, torch.float64, "history rapidities")
    require_cuda(hN, torch.float64, "history amplitudes")
    rows = hy.numel()
    if rows < 1 or hN.shape != (rows, R):
        raise ValueError("history must be rapidities [rows] and amplitudes [rows, R]")
    if not bool((hy[1:] > hy[:-1]).all()):
        raise ValueError("history rapidities must increase")
    return torch.tensor([rows], dtype=torch.int32, device=device)


def bk_k1(r: torch.Tensor, N: torch.Tensor, t64: torch.Tensor, config,
          history: tuple[torch.Tensor, torch.Tensor] | None = None, rule: str = "fixed",
          stats: dict | None = None) -> torch.Tensor:
    """K1 for every r node (0 for index 0 and N > 0.99999).  ``history``
    (rapidities [rows], amplitudes [rows, R], the last row equal to N) turns
    the kinematical constraint on at the last row's rapidity.  ``rule``:
    "fixed" (Gauss-Legendre per log-r panel) or "adaptive" (the nested
    adaptive Gauss-Kronrod rule with RINTPOINTS, THETAINTPOINTS, INTACCURACY);
    for the adaptive rule ``stats`` (a dict) receives the per-node error
    estimates and integrand evaluations."""

    R = r.numel()
    out = torch.empty(R, dtype=torch.float64, device=r.device)
    ph = physics(config)
    if rule not in K1_RULE_CODES:
        raise ValueError(f"unknown K1 rule {rule!r}")
    if rule == "adaptive":
        err = torch.empty(R, dtype=torch.float64, device=r.device)
        evals = torch.empty(R, dtype=torch.int64, device=r.device)
        hy = hN = d_rows = None
        cap = 0
        if history is not None:
            d_rows = _history_rows(history, R, r.device)
            hy, hN = history
            cap = hy.numel()
        check(
            lib().nlo_bk_k1_adaptive(R, ptr(r), ptr(N), ptr(t64), ctypes.byref(ph), config.RINTPOINTS,
                                     config.THETAINTPOINTS, float(config.INTACCURACY), ptr(hy), ptr(hN),
                                     ptr(d_rows), cap, ptr(out), ptr(err), ptr(evals),
                                     stream_handle(r.device)),
            "nlo_bk_k1_adaptive",
        )
        if stats is not None:
            stats["error"] = err
            stats["evaluations"] = evals
        return out
    if history is None:
        check(
            lib().nlo_bk_k1(R, ptr(r), ptr(N), ptr(t64), ctypes.byref(ph), config.K1_RADIAL_ORDER,
                            config.K1_ANGULAR_ORDER, ptr(out), stream_handle(r.device)),
            "nlo_bk_k1",
        )
        return out
    d_rows = _history_rows(history, R, r.device)
    hy, hN = history
    rows = hy.numel()
    check(
        lib().nlo_bk_k1_history(R, ptr(r), ptr(N), ptr(t64), ctypes.byref(ph), config.K1_RADIAL_ORDER,
                                config.K1_ANGULAR_ORDER, ptr(hy), ptr(hN), ptr(d_rows), rows, ptr(out),
                                stream_handle(r.device)),
        "nlo_bk_k1_history",
    )
    return out


PRECISION_CODES = {"fp64": _lib.NLO_PREC_FP64, "fp32": _lib.NLO_PREC_FP32,
                   "fp64_stable": _lib.NLO_PREC_FP64_STABLE}


def bk_nlo(
    r: torch.Tensor,
    N: torch.Tensor,
    t64: torch.Tensor,
    t32: torch.Tensor,
    edges: torch.Tensor,
    table: _lib.SobolTable,
    n: int,
    config,
    *,
    first_index: int = 0,
    precision: str = "fp64",
    histogram: bool = False,
    sum_only: bool = False,
):
    """Per-parent (sum w, sum w^2, sum w K2, sum w Kf) [R, 4] (+ hist, count).

    sum_only: only sum w is formed (the other columns read 0), the kernel
    instance the evolver runs."""

    R = r.numel()
    bins = edges.shape[-1] - 1
    require_cuda(edges, torch.float64, "edges")
    if edges.shape != (R, 4, bins + 1):
        raise ValueError("edges must be [R, 4, bins + 1]")
    sums = torch.empty((R, 4), dtype=torch.float64, device=r.device)
    hist = torch.empty((R, 4, bins), dtype=torch.float64, device=r.device) if histogram else None
    count = torch.empty_like(hist) if histogram else None
    nbytes = lib().nlo_bk_nlo_workspace(R, bins, n, int(histogram))
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=r.device)
    ph = physics(config)
    check(
        lib().nlo_bk_nlo(R, ptr(r), ptr(N), ptr(t64), ptr(t32), ptr(edges), bins, ctypes.byref(table),
                         first_index, n, PRECISION_CODES[precision] | (_lib.NLO_BK_SUM_ONLY if sum_only else 0),
                         ctypes.byref(ph), ptr(sums),
                         ptr(hist), ptr(count), ptr(ws), nbytes, stream_handle(r.device)),
        "nlo_bk_nlo",
    )
    return sums, hist, count


def rhs_finalize(N: torch.Tensor, k1: torch.Tensor, sums: torch.Tensor, n: int) -> tuple[torch.Tensor, torch.Tensor]:
    """dN/dy and the K2 + Kf part (S/bk/evolution.py:340-351)."""

    R = N.numel()
    dndy = torch.empty(R, dtype=torch.float64, device=N.device)
    k2kf = torch.empty(R, dtype=torch.float64, device=N.device)
    check(
        lib().nlo_bk_rhs_finalize(R, ptr(N), ptr(k1), ptr(sums), n, ptr(dndy), ptr(k2kf),
                                  stream_handle(N.device)),
        "nlo_bk_rhs_finalize",
    )
    return dndy, k2kf


RK_CODES = {"euler": _lib.NLO_RK_EULER, "rk2": _lib.NLO_RK_MIDPOINT, "kutta3": _lib.NLO_RK_KUTTA3,
            "rk4": _lib.NLO_RK_CLASSIC4, "adaptive_rk2": _lib.NLO_RK_ADAPTIVE2}


def _set_k1_rule(d, config, rule: str) -> None:
    if rule not in K1_RULE_CODES:
        raise ValueError(f"unknown K1 rule {rule!r}")
    d.k1_rule = K1_RULE_CODES[rule]
    d.k1_rint = int(config.RINTPOINTS)
    d.k1_tint = int(config.THETAINTPOINTS)
    d.k1_epsrel = float(config.INTACCURACY)


class Evolver:
    """One captured hipGraph per RK step (C ABI ``nlo_bk_evolver_*``)."""

    def __init__(
        self,
        r: torch.Tensor,
        edges: torch.Tensor,
        table: _lib.SobolTable,
        config,
        *,
        n_samples: int,
        precision: str,
        rk_method: str,
        step: float,
        nlo: bool = True,
        first_index: int = 0,
        eps: float = float(torch.finfo(torch.float64).eps),
        history: tuple[torch.Tensor, torch.Tensor] | None = None,
        k1_rule: str = "fixed",
    ) -> None:
        """``history``: (rapidities [capacity], amplitudes [capacity, R]) with
        row 0 filled, for the kinematical constraint (Euler only); each step
        appends a row.  ``eps``: the step-underflow epsilon of adaptive RK2.
        ``k1_rule``: "fixed" or "adaptive" (see bk_k1)."""

        require_cuda(r, torch.float64, "r")
        require_cuda(edges, torch.float64, "edges")
        self.r = r
        self.edges = edges
        self.history = history
        d = _lib.EvolverDesc()
        d.R = r.numel()
        d.bins = edges.shape[-1] - 1
        d.n_samples = n_samples
        d.first_index = first_index
        d.precision = PRECISION_CODES[precision]
        d.rk_method = RK_CODES[rk_method]
        d.step = step
        d.force_positive = int(bool(config.FORCE_POSITIVE_N))
        d.nr = config.K1_RADIAL_ORDER
        d.nt = config.K1_ANGULAR_ORDER
        d.nlo = int(nlo)
        d.physics = physics(config)
        d.sobol = table
        d.abserr = float(config.DE_SOLVER_ABSERR)
        d.relerr = float(config.DE_SOLVER_RELERR)
        d.eps = float(eps)
        _set_k1_rule(d, config, k1_rule)
        if history is not None:
            hy, hN = history
            require_cuda(hy, torch.float64, "history rapidities")
            require_cuda(hN, torch.float64, "history amplitudes")
            if hN.shape != (hy.numel(), r.numel()):
                raise ValueError("history must be rapidities [capacity] and amplitudes [capacity, R]")
            d.kinematical = 1
            d.history_capacity = hy.numel()
            d.d_history_y = hy.data_ptr()
            d.d_history_N = hN.data_ptr()
        self.desc = d
        nbytes = lib().nlo_bk_evolver_workspace(ctypes.byref(d))
        if nbytes == 0:
            raise ValueError("invalid evolver description")
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=r.device)
        handle = ctypes.c_void_p()
        check(
            lib().nlo_bk_evolver_create(ctypes.byref(d), ptr(r), ptr(edges), ptr(self.workspace), nbytes,
                                        stream_handle(r.device), ctypes.byref(handle)),
            "nlo_bk_evolver_create",
        )
        self.handle = handle
        base = self.workspace.data_ptr()
        R = d.R

        def view(address: int, count: int, dtype: torch.dtype) -> torch.Tensor:
            offset = address - base
            size = torch.empty((), dtype=dtype).element_size()
            if offset < 0 or offset % size or offset + count * size > nbytes:
                raise RuntimeError("evolver buffer outside its workspace")
            return self.workspace[offset : offset + count * size].view(dtype)

        self.state = view(lib().nlo_bk_evolver_state(handle), R, torch.float64)
        self.k1 = view(lib().nlo_bk_evolver_k1(handle), R, torch.float64)
        self.k2kf = view(lib().nlo_bk_evolver_k2kf(handle), R, torch.float64)
        self.flag = view(lib().nlo_bk_evolver_nonfinite_flag(handle), 1, torch.int32)
        self.active_evaluations = view(lib().nlo_bk_evolver_active_evaluations(handle), 1, torch.int64)

    def step(self) -> None:
        check(lib().nlo_bk_evolver_step(self.handle, stream_handle(self.r.device)), "nlo_bk_evolver_step")

    def adaptive_state(self) -> tuple[float, float, int, int]:
        """(y, next step size, accepted, rejected attempts) of the controller."""

        y, h = ctypes.c_double(), ctypes.c_double()
        acc, rej = ctypes.c_int64(), ctypes.c_int64()
        check(lib().nlo_bk_evolver_adaptive_state(self.handle, ctypes.byref(y), ctypes.byref(h),
                                                   ctypes.byref(acc), ctypes.byref(rej)),
              "nlo_bk_evolver_adaptive_state")
        return y.value, h.value, acc.value, rej.value

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            torch.cuda.current_stream(self.r.device).synchronize()
            lib().nlo_bk_evolver_destroy(self.handle)
            self.handle = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


class BatchEvolver:
    """Many evolutions of uneven grid sizes stepped together (C ABI ``nlo_bk_batch_*``).

    ``r`` and ``edges`` concatenate the runs' grids ([sum R], [sum R, 4, B + 1]);
    ``sizes`` lists R per run.  ``spt`` / ``grid_blocks`` fix the K2 + Kf launch
    shape; left at 0 the first step times the candidates and keeps the fastest.
    """

    def __init__(self, r: torch.Tensor, edges: torch.Tensor, sizes, table: _lib.SobolTable, config, *,
                 n_samples: int, precision: str, rk_method: str, step: float, nlo: bool = True,
                 first_index: int = 0, spt: int = 0, grid_blocks: int = 0, k1_rule: str = "fixed") -> None:
        require_cuda(r, torch.float64, "r")
        require_cuda(edges, torch.float64, "edges")
        sizes = [int(v) for v in sizes]
        if sum(sizes) != r.numel() or edges.shape[0] != r.numel() or edges.dim() != 3 or edges.shape[1] != 4:
            raise ValueError("r [sum R] and edges [sum R, 4, bins + 1] must concatenate the runs")
        if rk_method not in ("euler", "rk2", "kutta3", "rk4"):
            raise ValueError("the batched driver takes fixed-step methods only")
        self.r, self.edges, self.sizes = r, edges, sizes
        d = _lib.BatchDesc()
        d.runs = len(sizes)
        d.bins = edges.shape[-1] - 1
        d.n_samples = n_samples
        d.first_index = first_index
        d.precision = PRECISION_CODES[precision]
        d.rk_method = RK_CODES[rk_method]
        d.step = step
        d.force_positive = int(bool(config.FORCE_POSITIVE_N))
        d.nr = config.K1_RADIAL_ORDER
        d.nt = config.K1_ANGULAR_ORDER
        d.nlo = int(nlo)
        d.spt = int(spt)
        d.grid_blocks = int(grid_blocks)
        d.physics = physics(config)
        d.sobol = table
        _set_k1_rule(d, config, k1_rule)
        self.desc = d
        self._sizes = (ctypes.c_int32 * len(sizes))(*sizes)
        nbytes = lib().nlo_bk_batch_workspace(ctypes.byref(d), self._sizes)
        if nbytes == 0:
            raise ValueError("invalid batch description")
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=r.device)
        handle = ctypes.c_void_p()
        check(lib().nlo_bk_batch_create(ctypes.byref(d), self._sizes, ptr(r), ptr(edges), ptr(self.workspace),
                                        nbytes, stream_handle(r.device), ctypes.byref(handle)),
              "nlo_bk_batch_create")
        self.handle = handle
        base = self.workspace.data_ptr()

        def view(address: int, count: int, dtype: torch.dtype) -> torch.Tensor:
            offset = address - base
            size = torch.empty((), dtype=dtype).element_size()
            if offset < 0 or offset % size or offset + count * size > nbytes:
                raise RuntimeError("batch buffer outside its workspace")
            return self.workspace[offset : offset + count * size].view(dtype)

        self.state = view(lib().nlo_bk_batch_state(handle), r.numel(), torch.float64)
        self.flag = view(lib().nlo_bk_batch_nonfinite_flag(handle), 1, torch.int32)
        self.active_evaluations = view(lib().nlo_bk_batch_active_evaluations(handle), 1, torch.int64)

    def step(self) -> None:
        check(lib().nlo_bk_batch_step(self.handle, stream_handle(self.r.device)), "nlo_bk_batch_step")

    def tuning(self) -> tuple[int, int, float]:
        """(samples per lane, persistent grid blocks, ms per K2 + Kf launch) in use."""

        spt, grid, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        check(lib().nlo_bk_batch_tuning(self.handle, ctypes.byref(spt), ctypes.byref(grid), ctypes.byref(ms)),
              "nlo_bk_batch_tuning")
        return spt.value, grid.value, ms.value

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            torch.cuda.current_stream(self.r.device).synchronize()
            lib().nlo_bk_batch_destroy(self.handle)
            self.handle = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------- DIS


def dis_special(fn: str, x: torch.Tensor, *, precision: str = "fp64") -> torch.Tensor:
    """K0, K1, K2 or Li2 of a float64 device tensor (S/numerics/special.py:18-97)."""

    codes = {"K0": _lib.NLO_FN_K0, "K1": _lib.NLO_FN_K1, "K2": _lib.NLO_FN_K2, "Li2": _lib.NLO_FN_LI2}
    require_cuda(x, torch.float64, "x")
    out = torch.empty_like(x)
    check(lib().nlo_dis_special(codes[fn], x.numel(), ptr(x), PRECISION_CODES[precision], ptr(out),
                                stream_handle(x.device)), "nlo_dis_special")
    return out


def dis_dipole_struct(dipole) -> _lib.DisDipole:
    """The C view of a GBW or BKDipole (nlo_bk_optimized_amd.dipole.amplitude)."""

    d = _lib.DisDipole()
    if dipole.kind == "gbw":
        d.kind = _lib.NLO_DIPOLE_GBW
        d.qs0sqr, d.lambda_, d.gamma, d.x0 = dipole.Qs0sqr, dipole.lambda_, dipole.gamma, dipole.x0
        return d
    d.kind = _lib.NLO_DIPOLE_BK_TABLE
    d.R, d.T = dipole.device_r.numel(), dipole.device_Y.numel()
    d.d_r, d.d_Y, d.d_N = ptr(dipole.device_r), ptr(dipole.device_Y), ptr(dipole.device_N)
    d.x0 = 1.0
    d.bilinear = 1 if dipole.bilinear else 0
    local = getattr(dipole, "device_local", None)
    if local is not None and not dipole.bilinear:
        d.d_local_coef = ptr(local)
    if dipole.device_coef is not None:
        d.initialized = 1
        d.y_init = dipole.interpolator_Y
        d.maxr_interpolate = dipole.maxr_interpolate
        d.d_init_coef = ptr(dipole.device_coef)
    return d


def dis_dipole_amplitude(dipole, r: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
    """N(r, Y) on the device for float64 tensors of one shape."""

    require_cuda(r, torch.float64, "r")
    require_cuda(Y, torch.float64, "Y")
    if r.shape != Y.shape:
        raise ValueError("r and Y must have the same shape")
    out = torch.empty_like(r)
    d = dis_dipole_struct(dipole)
    check(lib().nlo_dis_dipole_amplitude(ctypes.byref(d), r.numel(), ptr(r), ptr(Y), ptr(out),
                                         stream_handle(r.device)), "nlo_dis_dipole_amplitude")
    return out


def dis_bk_local_splines(dipole) -> torch.Tensor:
    """The local-spline table of a BK-table dipole (nlo_dis_bk_local_splines):
    T (R - 1) x 4 doubles on the dipole's device."""

    d = dis_dipole_struct(dipole)
    d.d_local_coef = None
    T, R = dipole.device_Y.numel(), dipole.device_r.numel()
    coef = torch.empty(4 * T * (R - 1), dtype=torch.float64, device=dipole.device_r.device)
    check(lib().nlo_dis_bk_local_splines(ctypes.byref(d), ptr(coef), stream_handle(coef.device)),
          "nlo_dis_bk_local_splines")
    return coef


def dis_bk_initialize(dipole, Y: float) -> tuple[torch.Tensor, float]:
    """Coefficients (4 (R - 1)) and maxr_interpolate of BKDipole.initialize_interpolation at Y."""

    d = dis_dipole_struct(dipole)
    d.initialized = 0
    coef = torch.empty(4 * (dipole.device_r.numel() - 1), dtype=torch.float64, device=dipole.device_r.device)
    maxr = torch.empty(1, dtype=torch.float64, device=dipole.device_r.device)
    check(lib().nlo_dis_bk_initialize(ctypes.byref(d), float(Y), ptr(coef), ptr(maxr),
                                      stream_handle(coef.device)), "nlo_dis_bk_initialize")
    return coef, float(maxr.item())


def dis_desc(Q2: float, xbj: float, transverse: bool, quarks, dipole, config) -> _lib.DisDesc:
    """nlo_dis_desc from DISConfig fields (S/dis/config.py:101-145)."""

    from nlo_bk_optimized_amd.dis.config import (
        HeavyQuarkX,
        NcScheme,
        RunningCouplingIRScheme,
        RunningCouplingScheme,
    )

    if not 1 <= len(quarks) <= _lib.DIS_MAX_QUARKS:
        raise ValueError(f"between 1 and {_lib.DIS_MAX_QUARKS} quarks are supported")
    d = _lib.DisDesc()
    d.Q2, d.xbj = float(Q2), float(xbj)
    d.transverse = 1 if transverse else 0
    d.n_quarks = len(quarks)
    for i, q in enumerate(quarks):
        d.mass[i] = q.mass
        d.charge[i] = q.charge
    d.maxr = config.maxr
    d.finite_nc = 1 if config.nc_scheme is NcScheme.FiniteNC else 0
    d.parent_coupling = 1 if config.rc_scheme is RunningCouplingScheme.PARENT else 0
    d.smooth_coupling = 1 if config.rc_ir_scheme is RunningCouplingIRScheme.SMOOTH else 0
    d.mass_dependent_x = 1 if config.heavy_quark_x_scheme is HeavyQuarkX.MassDependentX else 0
    d.c2_alpha, d.max_alpha = config.C2_alpha, config.max_alpha_s_freeze
    d.nf_alphas = config.active_flavors
    d.nested_points = config.cuda_nested_points
    d.dipole = dis_dipole_struct(dipole)
    return d


def dis_integrand(desc: _lib.DisDesc, kind: str, x: torch.Tensor, *, precision: str = "fp64") -> torch.Tensor:
    """The contribution's integrand at unit-cube points x[n, dim] (float64 device)."""

    require_cuda(x, torch.float64, "x")
    dim = _lib.DIS_DIMENSIONS[kind]
    if x.ndim != 2 or x.shape[1] != dim:
        raise ValueError(f"{kind} needs points of shape (n, {dim})")
    out = torch.empty(x.shape[0], dtype=torch.float64, device=x.device)
    check(lib().nlo_dis_integrand(ctypes.byref(desc), _lib.DIS_KINDS[kind], x.shape[0], ptr(x),
                                  PRECISION_CODES[precision], ptr(out), stream_handle(x.device)), "nlo_dis_integrand")
    return out


def dis_qg_i23(desc: _lib.DisDesc, x: torch.Tensor, *, precision: str = "fp64") -> tuple[torch.Tensor, torch.Tensor]:
    """The nested I2 and I3 at unit-cube points x[n, 5] (float64 device) from
    one pass over the inner rule."""

    require_cuda(x, torch.float64, "x")
    if x.ndim != 2 or x.shape[1] != 5:
        raise ValueError("the nested I2 / I3 need points of shape (n, 5)")
    i2 = torch.empty(x.shape[0], dtype=torch.float64, device=x.device)
    i3 = torch.empty_like(i2)
    check(lib().nlo_dis_qg_i23(ctypes.byref(desc), x.shape[0], ptr(x), PRECISION_CODES[precision], ptr(i2),
                               ptr(i3), stream_handle(x.device)), "nlo_dis_qg_i23")
    return i2, i3


def dis_vegas(desc: _lib.DisDesc, kind: str, edges: torch.Tensor, table: _lib.SobolTable, n: int, *,
              first_index: int = 0, precision: str = "fp64", histogram: bool = True):
    """One VEGAS pass: (sums[2] = (sum w, sum w^2), hist[dim, B], count[dim, B])."""

    require_cuda(edges, torch.float64, "edges")
    dim = _lib.DIS_DIMENSIONS[kind]
    bins = edges.shape[-1] - 1
    if edges.shape != (dim, bins + 1):
        raise ValueError(f"edges must be ({dim}, bins + 1)")
    dev = edges.device
    sums = torch.empty(2, dtype=torch.float64, device=dev)
    hist = torch.empty((dim, bins), dtype=torch.float64, device=dev) if histogram else None
    count = torch.empty((dim, bins), dtype=torch.float64, device=dev) if histogram else None
    ws_bytes = lib().nlo_dis_vegas_workspace(_lib.DIS_KINDS[kind], bins, n)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    check(
        lib().nlo_dis_vegas(ctypes.byref(desc), _lib.DIS_KINDS[kind], bins, ptr(edges), ctypes.byref(table),
                            first_index, n, PRECISION_CODES[precision], ptr(sums), ptr(hist), ptr(count), ptr(ws),
                            ws_bytes, stream_handle(dev)),
        "nlo_dis_vegas",
    )
    return sums, hist, count


def dis_gauss_legendre(desc: _lib.DisDesc, kind: str, points: int, *, precision: str = "fp64",
                       device: torch.device | str = "cuda") -> torch.Tensor:
    """[fine, coarse] tensor Gauss-Legendre sums of a two-dimensional contribution."""

    out = torch.empty(2, dtype=torch.float64, device=device)
    check(lib().nlo_dis_gauss_legendre(ctypes.byref(desc), _lib.DIS_KINDS[kind], points, PRECISION_CODES[precision],
                                       ptr(out), stream_handle(out.device)), "nlo_dis_gauss_legendre")
    return out


def dis_batch(items, tables, *, bins: int, precision: str, epsrel: float, epsabs: float = 0.0,
              device: torch.device | str = "cuda", workspace_budget: int = 8 << 30) -> torch.Tensor:
    """nlo_dis_batch_run over ``items`` (nlo_dis_item) with Sobol ``tables``:
    [n_items, 4] float64 (value, error, evaluations, converged) on the device.
    Items are independent, so a batch whose workspace would exceed
    ``workspace_budget`` bytes runs as consecutive sub-batches."""

    device = torch.device(device)
    edges0 = torch.linspace(0, 1, bins + 1, dtype=torch.float64, device=device)
    tarr = (_lib.SobolTable * max(1, len(tables)))(*tables)
    code = PRECISION_CODES[precision]

    def describe(part) -> _lib.DisBatchDesc:
        d = _lib.DisBatchDesc()
        d.n_items = len(part)
        d.n_tables = len(tables)
        d.items = (_lib.DisItem * len(part))(*part)
        d.tables = tarr
        d.d_initial_edges = ptr(edges0)
        d.bins = bins
        d.precision = code
        d.epsrel, d.epsabs = float(epsrel), float(epsabs)
        return d

    out = torch.empty((len(items), 4), dtype=torch.float64, device=device)
    start = 0
    while start < len(items):
        count = len(items) - start
        while True:
            d = describe(items[start:start + count])
            nbytes = lib().nlo_dis_batch_workspace(ctypes.byref(d))
            if nbytes <= workspace_budget or count == 1:
                break
            count = max(1, count // 2)
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)
        view = out[start:start + count]
        check(lib().nlo_dis_batch_run(ctypes.byref(d), ptr(view), ptr(ws), nbytes, stream_handle(device)),
              "nlo_dis_batch_run")
        del ws
        start += count
    return out


def fixed_sobol_tables(seed: int, training_rounds: int, replicates: int) -> list[_lib.SobolTable]:
    """The 5-D point sets of make_sobol_points (S/dis/fixed_sobol.py:87-97):
    training round k scrambled with seed + k, replicate k with seed + 10000 + k."""

    return ([sobol_table(5, seed + k) for k in range(training_rounds)]
            + [sobol_table(5, seed + 10_000 + k) for k in range(replicates)])


def dis_fixed_sobol(items, tables, *, training_samples: int, training_rounds: int, samples_per_replicate: int,
                    replicates: int, bins: int, precision: str, epsrel: float, i1_weight: float = 1.0,
                    fold_angle: bool = True, device: torch.device | str = "cuda",
                    return_edges: bool = False, grid: torch.Tensor | None = None,
                    workspace_budget: int = 8 << 30):
    """nlo_dis_fixed_sobol over ``items`` (nlo_dis_fixed_sobol_item) with the
    flat list of point-set ``tables``: [n_items, 4, 4] float64 on the device
    (rows I1, I2, I3, total; columns value, error, evaluations, converged),
    and with ``return_edges`` the learned grids [n_items, 5, bins + 1].  With
    ``grid`` (float64 [n_items, 5, bins + 1] on the device) the grids are frozen
    and no training runs (training_rounds must be 0).  Items are independent: a
    batch whose workspace would exceed ``workspace_budget`` bytes runs as
    consecutive sub-batches."""

    device = torch.device(device)
    tarr = (_lib.SobolTable * len(tables))(*tables)
    if grid is not None:
        require_cuda(grid, torch.float64, "grid")
        if tuple(grid.shape) != (len(items), 5, bins + 1):
            raise ValueError(f"grid must have shape [{len(items)}, 5, {bins + 1}]")

    def describe(part, at: int = 0) -> _lib.FixedSobolDesc:
        d = _lib.FixedSobolDesc()
        d.n_items = len(part)
        d.n_tables = len(tables)
        d.items = (_lib.FixedSobolItem * len(part))(*part)
        d.tables = tarr
        d.training_samples = int(training_samples)
        d.training_rounds = int(training_rounds)
        d.samples_per_replicate = int(samples_per_replicate)
        d.replicates = int(replicates)
        d.bins = int(bins)
        d.fold_angle = 1 if fold_angle else 0
        d.precision = PRECISION_CODES[precision]
        d.i1_weight = float(i1_weight)
        d.epsrel = float(epsrel)
        d.d_grid = ptr(grid[at:at + len(part)]) if grid is not None else None
        return d

    out = torch.empty((len(items), 4, 4), dtype=torch.float64, device=device)
    edges = torch.empty((len(items), 5, bins + 1), dtype=torch.float64, device=device) if return_edges else None
    start = 0
    while start < len(items):
        count = len(items) - start
        while True:
            d = describe(items[start:start + count], start)
            nbytes = lib().nlo_dis_fixed_sobol_workspace(ctypes.byref(d))
            if nbytes <= workspace_budget or count == 1:
                break
            count = max(1, count // 2)
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)
        check(lib().nlo_dis_fixed_sobol(ctypes.byref(d), ptr(out[start:start + count]),
                                        ptr(edges[start:start + count]) if edges is not None else None,
                                        ptr(ws), nbytes, stream_handle(device)), "nlo_dis_fixed_sobol")
        del ws
        start += count
    return (out, edges) if return_edges else out


def dis_endpoint_map(points: torch.Tensor, edges: torch.Tensor, maxr: float, *, fold_angle: bool = True):
    """dis_sobol_endpoint_map (K/dis/sampling/kernel.cu:15-81): (first,
    first_weight, second, second_weight) for device points [n, 5] and a grid
    [5, bins + 1] of one dtype (float32 or float64); the weights include the
    grid's inverse density."""

    if points.dtype not in (torch.float32, torch.float64):
        raise TypeError("points must be float32 or float64")
    require_cuda(points, points.dtype, "points")
    require_cuda(edges, points.dtype, "edges")
    if points.dim() != 2 or points.shape[1] != 5:
        raise ValueError("points must have shape [n, 5]")
    if edges.dim() != 2 or edges.shape[0] != 5 or edges.shape[1] < 3:
        raise ValueError("edges must have shape [5, bins + 1] with at least two bins")
    n = points.shape[0]
    first = torch.empty_like(points)
    second = torch.empty_like(points)
    w1 = torch.empty(n, dtype=points.dtype, device=points.device)
    w2 = torch.empty_like(w1)
    code = _lib.NLO_PREC_FP32 if points.dtype == torch.float32 else _lib.NLO_PREC_FP64
    check(lib().nlo_dis_endpoint_map(code, n, ptr(points), ptr(edges), edges.shape[1] - 1, float(maxr),
                                     1 if fold_angle else 0, ptr(first), ptr(w1), ptr(second), ptr(w2),
                                     stream_handle(points.device)), "nlo_dis_endpoint_map")
    return first, w1, second, w2


def dtype_epsilon(dtype: torch.dtype) -> float:
    return float(torch.finfo(dtype).eps)


__all__ = [
    "Evolver",
    "dis_bk_initialize",
    "dis_desc",
    "dis_dipole_amplitude",
    "dis_dipole_struct",
    "dis_gauss_legendre",
    "dis_integrand",
    "dis_special",
    "dis_vegas",
    "PRECISION_CODES",
    "RK_CODES",
    "BatchEvolver",
    "bk_k1",
    "bk_nlo",
    "dtype_epsilon",
    "physics",
    "rhs_finalize",
    "sobol_draw",
    "sobol_table",
    "spline_build",
    "vegas_adapt",
    "vegas_histogram",
    "vegas_map",
]

