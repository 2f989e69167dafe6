"""ctypes binding of the FP64 CPU oracle (``libnlo_oracle.so``).

TEST INFRASTRUCTURE ONLY: the reference the HIP path is tested against.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module.  The product package never imports it.

Functions restate the reference (S/ = torch_nlo_clean/src/nlo_torch/) as cited
in ``nlo_oracle.c``.  Parity status of each piece:

* MV, r grid, coupling, Kernel_lo/_lo_integrand, Kernel_nlo, Kernel_nlo_fermion,
  _nlo_integrand, _adapt_vegas_edges: pinned by vectors from the reference's own
  FP64 CPU code (tests/golden/bk_reference_vectors.npz, made by
  tests/golden/make_bk_golden.py).
* natural / log-log spline: pinned by scipy.interpolate.CubicSpline(natural).
* Bessel J1 / I1: pinned by scipy.special.j1 / i1.
* Sobol: pinned by torch.quasirandom.SobolEngine (bit-exact).
* VEGAS map: pinned by the uniform-grid identity (S/dis/fixed_sobol.py:100-113)
  and brute-force bin counts.
* fixed-rule K1: pinned by an adaptive scipy integral of the oracle integrand.
* RK stages: pinned by the reference's known answer (constant derivative) and
  exact polynomial ODEs.
"""

from __future__ import annotations

import ctypes
import math
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libnlo_oracle.so"

_dp = ctypes.POINTER(ctypes.c_double)
_GK_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_void_p)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)


class BKParams(ctypes.Structure):
    _fields_ = [
        ("nc", ctypes.c_double),
        ("nf", ctypes.c_double),
        ("lambda_qcd", ctypes.c_double),
        ("c2", ctypes.c_double),
        ("ksub", ctypes.c_double),
        ("minr", ctypes.c_double),
        ("symmetrize", ctypes.c_int),
        ("geometry_cos", ctypes.c_int),
        ("stable", ctypes.c_int),
        ("rc_lo", ctypes.c_int),
        ("rc_nlo", ctypes.c_int),
        ("resum_rc", ctypes.c_int),
        ("order", ctypes.c_int),
        ("reserved", ctypes.c_int),
        ("fixed_as", ctypes.c_double),
    ]


class _History(ctypes.Structure):
    _fields_ = [
        ("rows", ctypes.c_int),
        ("n", ctypes.c_int),
        ("r", _dp),
        ("y", _dp),
        ("N", _dp),
        ("rapidity", ctypes.c_double),
    ]


class _Spline(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("r_min", ctypes.c_double),
        ("r_max", ctypes.c_double),
        ("x", _dp),
        ("a", _dp),
        ("b", _dp),
        ("c", _dp),
        ("d", _dp),
    ]


def build() -> Path:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""

    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB_PATH))
        L = _lib
        L.orc_alpha_s.restype = ctypes.c_double
        L.orc_alpha_s.argtypes = [ctypes.c_double, ctypes.POINTER(BKParams)]
        L.orc_mv.restype = ctypes.c_double
        L.orc_mv.argtypes = [ctypes.c_double] * 5
        L.orc_r_grid.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_natural_spline.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_evolution_splines.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, _dp]
        L.orc_spline_eval.restype = ctypes.c_double
        L.orc_spline_eval.argtypes = [ctypes.POINTER(_Spline), ctypes.c_double]
        L.orc_loglog_eval.restype = ctypes.c_double
        L.orc_loglog_eval.argtypes = [
            ctypes.POINTER(_Spline),
            ctypes.c_double,
            ctypes.c_double,
            ctypes.c_double,
        ]
        L.orc_bessel_j1.restype = ctypes.c_double
        L.orc_bessel_j1.argtypes = [ctypes.c_double]
        L.orc_bessel_i1.restype = ctypes.c_double
        L.orc_bessel_i1.argtypes = [ctypes.c_double]
        L.orc_kernel_lo.restype = ctypes.c_double
        L.orc_kernel_lo.argtypes = [ctypes.c_double] * 3 + [ctypes.POINTER(BKParams)]
        L.orc_k1_integrand.restype = ctypes.c_double
        L.orc_k1_integrand.argtypes = [ctypes.c_double] * 3 + [
            ctypes.POINTER(_Spline),
            ctypes.POINTER(BKParams),
        ]
        L.orc_kernel_lo_magnitude.restype = ctypes.c_double
        L.orc_kernel_lo_magnitude.argtypes = [ctypes.c_double] * 3 + [ctypes.POINTER(BKParams)]
        L.orc_k1_integrand_magnitude.restype = ctypes.c_double
        L.orc_k1_integrand_magnitude.argtypes = [ctypes.c_double] * 3 + [
            ctypes.POINTER(_Spline),
            ctypes.POINTER(BKParams),
        ]
        L.orc_nlo_integrand_magnitude.restype = ctypes.c_double
        L.orc_nlo_integrand_magnitude.argtypes = [ctypes.c_double] * 5 + [
            ctypes.POINTER(_Spline),
            ctypes.POINTER(BKParams),
        ]
        L.orc_kernel_nlo.restype = ctypes.c_double
        L.orc_kernel_nlo.argtypes = [ctypes.c_double] * 6
        L.orc_kernel_nlo_fermion.restype = ctypes.c_double
        L.orc_kernel_nlo_fermion.argtypes = [ctypes.c_double] * 8
        L.orc_nlo_integrand_log.restype = ctypes.c_double
        L.orc_nlo_integrand_log.argtypes = [ctypes.c_double] * 5 + [
            ctypes.POINTER(_Spline),
            ctypes.POINTER(BKParams),
            _dp,
            _dp,
        ]
        L.orc_nlo_integrand_stable.restype = ctypes.c_double
        L.orc_nlo_integrand_stable.argtypes = [ctypes.c_double] * 5 + [
            ctypes.POINTER(_Spline),
            ctypes.POINTER(BKParams),
            _dp,
            _dp,
        ]
        L.orc_nlo_kernels_stable.argtypes = [ctypes.c_double] * 7 + [_dp]
        L.orc_k1_fixed.argtypes = [
            ctypes.c_int,
            _dp,
            ctypes.POINTER(_Spline),
            ctypes.c_int,
            _dp,
            _dp,
            ctypes.c_int,
            _dp,
            _dp,
            ctypes.POINTER(BKParams),
            _dp,
        ]
        L.orc_k1_fixed_magnitude.argtypes = L.orc_k1_fixed.argtypes
        L.orc_k1_fixed_history.argtypes = L.orc_k1_fixed.argtypes[:-1] + [
            ctypes.POINTER(_History),
            _dp,
        ]
        L.orc_history_N.restype = ctypes.c_double
        L.orc_history_N.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.POINTER(_Spline),
                                    ctypes.POINTER(_History)]
        L.orc_k1_integrand_history.restype = ctypes.c_double
        L.orc_k1_integrand_history.argtypes = [ctypes.c_double] * 3 + [
            ctypes.POINTER(_Spline),
            ctypes.POINTER(BKParams),
            ctypes.POINTER(_History),
        ]
        L.orc_adaptive_gk21.restype = ctypes.c_double
        L.orc_adaptive_gk21.argtypes = [_GK_FN, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, _dp,
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_long)]
        L.orc_k1_adaptive.argtypes = [
            ctypes.c_int, _dp, ctypes.POINTER(_Spline), ctypes.c_int, ctypes.c_int, ctypes.c_double,
            ctypes.POINTER(BKParams), ctypes.POINTER(_History), _dp, _dp,
            ctypes.POINTER(ctypes.c_long),
        ]
        L.orc_sobol_u32.argtypes = [
            ctypes.c_int,
            _u32p,
            _u32p,
            ctypes.c_int64,
            ctypes.c_int64,
            _u32p,
        ]
        L.orc_vegas_map.argtypes = [
            ctypes.c_int,
            ctypes.c_int,
            ctypes.c_int64,
            _u32p,
            _dp,
            _dp,
            _dp,
            _dp,
            _dp,
            _i32p,
        ]
        L.orc_vegas_adapt.argtypes = [
            ctypes.c_int,
            ctypes.c_int,
            _dp,
            _dp,
            _dp,
            ctypes.c_double,
            _dp,
        ]
        L.orc_bk_nlo_sums.argtypes = [
            ctypes.c_int,
            _dp,
            ctypes.POINTER(_Spline),
            ctypes.c_int,
            _dp,
            _dp,
            _dp,
            _u32p,
            _u32p,
            _u32p,
            ctypes.c_int64,
            ctypes.c_int64,
            ctypes.POINTER(BKParams),
            _dp,
            _dp,
            _dp,
            _dp,
        ]
        L.orc_num_threads.restype = ctypes.c_int
    return _lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _u(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


def params(
    *,
    nc: float = 3.0,
    nf: float = 3.0,
    lambda_qcd: float = 0.241,
    c2: float = 1.0,
    ksub: float = 0.65,
    minr: float = 1e-6,
    symmetrize: bool = True,
    geometry_cos: bool = False,
    stable: bool = False,
    rc_lo: int = 0,
    rc_nlo: int = 0,
    resum_rc: int = 0,
    order: int = 0,
    fixed_as: float = 0.2,
) -> BKParams:
    """Physics choices as the ORC_* codes of nlo_oracle.h (0 = production)."""

    return BKParams(
        nc, nf, lambda_qcd, c2, ksub, minr, int(symmetrize), int(geometry_cos), int(stable),
        int(rc_lo), int(rc_nlo), int(resum_rc), int(order), 0, float(fixed_as),
    )


@dataclass
class History:
    """Amplitude history rows N(r_grid, y_k) for the kinematical constraint."""

    r: np.ndarray
    y: np.ndarray
    N: np.ndarray  # rows x n
    rapidity: float

    def struct(self) -> _History:
        self.r = np.ascontiguousarray(self.r, dtype=np.float64)
        self.y = np.ascontiguousarray(self.y, dtype=np.float64)
        self.N = np.ascontiguousarray(self.N, dtype=np.float64)
        return _History(self.y.size, self.r.size, _d(self.r), _d(self.y), _d(self.N), float(self.rapidity))


@dataclass
class Spline:
    """A log-log natural spline held in numpy arrays (keeps them alive)."""

    r_min: float
    r_max: float
    x: np.ndarray
    coef: np.ndarray  # a|b|c|d, 4 x (n - 1)

    def struct(self) -> _Spline:
        m = self.x.size - 1
        c = self.coef
        return _Spline(
            self.x.size,
            self.r_min,
            self.r_max,
            _d(self.x),
            _d(c[0:m]),
            _d(c[m : 2 * m]),
            _d(c[2 * m : 3 * m]),
            _d(c[3 * m : 4 * m]),
        )


def r_grid(minr: float, maxr: float, rpoints: int) -> np.ndarray:
    out = np.empty(rpoints + 1)
    lib().orc_r_grid(minr, maxr, rpoints, _d(out))
    return out


def mv(r: np.ndarray, qs0sqr=0.1, gamma=1.0, ec=1.0, lambda_qcd=0.241) -> np.ndarray:
    return np.array([lib().orc_mv(float(v), qs0sqr, gamma, ec, lambda_qcd) for v in r])


def alpha_s(r: np.ndarray, p: BKParams | None = None) -> np.ndarray:
    p = p or params()
    return np.array([lib().orc_alpha_s(float(v), ctypes.byref(p)) for v in np.ravel(r)])


def natural_spline(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    m = x.size - 1
    out = np.empty((4, m))
    lib().orc_natural_spline(x.size, _d(x), _d(y), _d(out[0]), _d(out[1]), _d(out[2]), _d(out[3]))
    return out


def evolution_splines(r: np.ndarray, N: np.ndarray, force_positive: bool = True):
    r = np.ascontiguousarray(r, dtype=np.float64)
    N = np.ascontiguousarray(N, dtype=np.float64)
    n = r.size
    x = np.empty(n)
    cn = np.empty(4 * (n - 1))
    cs = np.empty(4 * (n - 1))
    lib().orc_evolution_splines(n, _d(r), _d(N), int(force_positive), _d(x), _d(cn), _d(cs))
    return Spline(float(r[0]), float(r[-1]), x, cn), Spline(float(r[0]), float(r[-1]), x, cs)


def loglog_eval(s: Spline, values, underflow: float, overflow: float) -> np.ndarray:
    st = s.struct()
    return np.array(
        [lib().orc_loglog_eval(ctypes.byref(st), float(v), underflow, overflow) for v in np.ravel(values)]
    )


def bessel_j1(x) -> np.ndarray:
    return np.array([lib().orc_bessel_j1(float(v)) for v in np.ravel(x)])


def bessel_i1(x) -> np.ndarray:
    return np.array([lib().orc_bessel_i1(float(v)) for v in np.ravel(x)])


def kernel_lo(r, z, theta, p: BKParams | None = None) -> np.ndarray:
    p = p or params()
    return np.array(
        [
            lib().orc_kernel_lo(float(a), float(b), float(c), ctypes.byref(p))
            for a, b, c in zip(np.ravel(r), np.ravel(z), np.ravel(theta))
        ]
    )


def k1_integrand(r, z, theta, sN: Spline, p: BKParams | None = None) -> np.ndarray:
    p = p or params()
    st = sN.struct()
    return np.array(
        [
            lib().orc_k1_integrand(float(a), float(b), float(c), ctypes.byref(st), ctypes.byref(p))
            for a, b, c in zip(np.ravel(r), np.ravel(z), np.ravel(theta))
        ]
    )


def kernel_lo_magnitude(r, z, theta, p: BKParams | None = None) -> np.ndarray:
    p = p or params()
    return np.array(
        [
            lib().orc_kernel_lo_magnitude(float(a), float(b), float(c), ctypes.byref(p))
            for a, b, c in zip(np.ravel(r), np.ravel(z), np.ravel(theta))
        ]
    )


def k1_integrand_magnitude(r, z, theta, sN: Spline, p: BKParams | None = None) -> np.ndarray:
    p = p or params()
    st = sN.struct()
    return np.array(
        [
            lib().orc_k1_integrand_magnitude(
                float(a), float(b), float(c), ctypes.byref(st), ctypes.byref(p)
            )
            for a, b, c in zip(np.ravel(r), np.ravel(z), np.ravel(theta))
        ]
    )


def nlo_integrand_magnitude(r, log_z, log_z2, theta_z, theta_z2, sS: Spline, p: BKParams | None = None):
    p = p or params()
    st = sS.struct()
    return np.array(
        [
            lib().orc_nlo_integrand_magnitude(*map(float, v), ctypes.byref(st), ctypes.byref(p))
            for v in zip(
                np.ravel(r), np.ravel(log_z), np.ravel(log_z2), np.ravel(theta_z), np.ravel(theta_z2)
            )
        ]
    )


def kernel_nlo(r, X, Y, X2, Y2, zz) -> np.ndarray:
    return np.array(
        [lib().orc_kernel_nlo(*map(float, v)) for v in zip(r, X, Y, X2, Y2, zz)]
    )


def kernel_nlo_fermion(r, X, Y, X2, Y2, zz, nf=3.0, nc=3.0) -> np.ndarray:
    return np.array(
        [lib().orc_kernel_nlo_fermion(*map(float, v), nf, nc) for v in zip(r, X, Y, X2, Y2, zz)]
    )


def nlo_integrand_log(r, log_z, log_z2, theta_z, theta_z2, sS: Spline, p: BKParams | None = None):
    """Returns (total, k2 part, kf part), each with the exp(2 log z + 2 log z') Jacobian."""

    p = p or params()
    st = sS.struct()
    k2 = ctypes.c_double()
    kf = ctypes.c_double()
    tot, a, b = [], [], []
    for v in zip(np.ravel(r), np.ravel(log_z), np.ravel(log_z2), np.ravel(theta_z), np.ravel(theta_z2)):
        tot.append(
            lib().orc_nlo_integrand_log(
                *map(float, v), ctypes.byref(st), ctypes.byref(p), ctypes.byref(k2), ctypes.byref(kf)
            )
        )
        a.append(k2.value)
        b.append(kf.value)
    return np.array(tot), np.array(a), np.array(b)


def nlo_integrand_stable(r, log_z, log_z2, theta_z, theta_z2, sN: Spline, p: BKParams | None = None):
    """Cancellation-free K2 + Kf integrand on the N spline: (total, k2 part, kf part)."""

    p = p or params()
    st = sN.struct()
    k2 = ctypes.c_double()
    kf = ctypes.c_double()
    tot, a, b = [], [], []
    for v in zip(np.ravel(r), np.ravel(log_z), np.ravel(log_z2), np.ravel(theta_z), np.ravel(theta_z2)):
        tot.append(
            lib().orc_nlo_integrand_stable(
                *map(float, v), ctypes.byref(st), ctypes.byref(p), ctypes.byref(k2), ctypes.byref(kf)
            )
        )
        a.append(k2.value)
        b.append(kf.value)
    return np.array(tot), np.array(a), np.array(b)


def k1_fixed(
    parent_r: np.ndarray, sN: Spline, nr=8, nt=24, p: BKParams | None = None, magnitude: bool = False,
    history: History | None = None,
) -> np.ndarray:
    """K1 on the fixed rule; ``magnitude`` integrates |integrand| (rounding scale);
    ``history`` switches the kinematical constraint on."""

    p = p or params()
    parent_r = np.ascontiguousarray(parent_r, dtype=np.float64)
    gx, gw = np.polynomial.legendre.leggauss(nr)
    tx, tw = np.polynomial.legendre.leggauss(nt)
    gx, gw, tx, tw = (np.ascontiguousarray(a) for a in (gx, gw, tx, tw))
    out = np.empty(parent_r.size)
    st = sN.struct()
    args = (parent_r.size, _d(parent_r), ctypes.byref(st), nr, _d(gx), _d(gw), nt, _d(tx), _d(tw),
            ctypes.byref(p))
    if history is not None:
        if magnitude:
            raise ValueError("no magnitude rule with the kinematical constraint")
        hs = history.struct()
        lib().orc_k1_fixed_history(*args, ctypes.byref(hs), _d(out))
        return out
    fn = lib().orc_k1_fixed_magnitude if magnitude else lib().orc_k1_fixed
    fn(*args, _d(out))
    return out


def history_N(r, rapidity, sN: Spline, history: History) -> np.ndarray:
    st, hs = sN.struct(), history.struct()
    return np.array([lib().orc_history_N(float(a), float(b), ctypes.byref(st), ctypes.byref(hs))
                     for a, b in zip(np.ravel(r), np.ravel(rapidity))])


def adaptive_gk21(f, a: float, b: float, epsrel: float = 1e-3, max_intervals: int = 85):
    """adaptive_gauss_kronrod_1d (S/numerics/integration.py:43-92), rule 21, of a
    Python callable: (value, error, converged, evaluations)."""

    cb = _GK_FN(lambda x, _ctx: float(f(x)))
    err, conv, ev = ctypes.c_double(), ctypes.c_int(), ctypes.c_long()
    v = lib().orc_adaptive_gk21(cb, None, a, b, epsrel, max_intervals, ctypes.byref(err),
                                ctypes.byref(conv), ctypes.byref(ev))
    return v, err.value, bool(conv.value), ev.value


def k1_adaptive(parent_r: np.ndarray, sN: Spline, rint: int = 85, tint: int = 85, epsrel: float = 1e-3,
                p: BKParams | None = None, history: History | None = None):
    """K1 by the nested adaptive rule of rapidity_derivative_lo
    (S/bk/derivatives.py:43-110): (K1, outer error estimate, integrand evaluations)."""

    p = p or params()
    parent_r = np.ascontiguousarray(parent_r, dtype=np.float64)
    out, err = np.empty(parent_r.size), np.empty(parent_r.size)
    ev = (ctypes.c_long * parent_r.size)()
    st = sN.struct()
    hs = history.struct() if history is not None else None
    lib().orc_k1_adaptive(parent_r.size, _d(parent_r), ctypes.byref(st), rint, tint, epsrel,
                          ctypes.byref(p), ctypes.byref(hs) if hs is not None else None, _d(out),
                          _d(err), ev)
    return out, err, np.array(ev[:], dtype=np.int64)


def k1_integrand_history(r, z, theta, sN: Spline, history: History, p: BKParams | None = None):
    p = p or params()
    st, hs = sN.struct(), history.struct()
    return np.array([
        lib().orc_k1_integrand_history(float(a), float(b), float(c), ctypes.byref(st), ctypes.byref(p),
                                       ctypes.byref(hs))
        for a, b, c in zip(np.ravel(r), np.ravel(z), np.ravel(theta))])


@dataclass
class SobolTable:
    """Scrambled direction numbers and digital shift of a SobolEngine."""

    dim: int
    direction: np.ndarray  # uint32 [dim, 30]
    shift: np.ndarray  # uint32 [dim]
    first_q: np.ndarray  # uint32 [dim]: integer form of the float64 first point


def sobol_table(dim: int, seed: int) -> SobolTable:
    """Direction numbers of ``SobolEngine(dim, scramble=True, seed=seed)``.

    torch is third-party here and computes the scramble; the oracle only runs
    the gray-code recurrence.  SobolEngine returns its float64 first point as
    float32(shift / 2^30) widened (torch/quasirandom.py ``_first_point``); that
    value is again an integer multiple of 2^-30 and is kept as ``first_q``.
    """

    import torch

    eng = torch.quasirandom.SobolEngine(dim, scramble=True, seed=seed)
    direction = eng.sobolstate.numpy().astype(np.uint32).copy()
    shift = eng.shift.numpy().astype(np.uint32).copy()
    first = (shift.astype(np.float64) / 2.0**30).astype(np.float32).astype(np.float64)
    first_q = np.round(first * 2.0**30).astype(np.uint64)
    first_q = np.minimum(first_q, np.uint64(2**30)).astype(np.uint32)
    return SobolTable(dim, direction, shift, first_q)


def sobol_u32(t: SobolTable, first: int, n: int) -> np.ndarray:
    out = np.empty((n, t.dim), dtype=np.uint32)
    lib().orc_sobol_u32(t.dim, _u(t.direction), _u(t.shift), first, n, _u(out))
    return out


def sobol_f64(t: SobolTable, first: int, n: int) -> np.ndarray:
    q = sobol_u32(t, first, n).astype(np.float64)
    if first == 0 and n > 0:
        q[0] = t.first_q
    return q / 2.0**30


def vegas_map(q: np.ndarray, edges: np.ndarray, lo: np.ndarray, width: np.ndarray):
    q = np.ascontiguousarray(q, dtype=np.uint32)
    n, dim = q.shape
    bins = edges.shape[1] - 1
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    width = np.ascontiguousarray(width, dtype=np.float64)
    x = np.empty((n, dim))
    jac = np.empty(n)
    b = np.empty((n, dim), dtype=np.int32)
    lib().orc_vegas_map(
        dim, bins, n, _u(q), _d(edges), _d(lo), _d(width), _d(x), _d(jac), b.ctypes.data_as(_i32p)
    )
    return x, jac, b


def vegas_adapt(edges: np.ndarray, hist: np.ndarray, count: np.ndarray, eps: float) -> np.ndarray:
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    hist = np.ascontiguousarray(hist, dtype=np.float64)
    count = np.ascontiguousarray(count, dtype=np.float64)
    dim, bins = hist.shape
    out = np.empty((dim, bins + 1))
    lib().orc_vegas_adapt(dim, bins, _d(edges), _d(hist), _d(count), eps, _d(out))
    return out


def nlo_bounds(r_grid_values: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """S/bk/derivatives.py:424-431: (log z, log z') in [ln(r0/2), ln(2 r_max)], angles in [0, 2 pi]."""

    lo_r = math.log(0.5 * r_grid_values[0])
    hi_r = math.log(2.0 * r_grid_values[-1])
    lo = np.array([lo_r, lo_r, 0.0, 0.0])
    width = np.array([hi_r - lo_r, hi_r - lo_r, 2 * math.pi, 2 * math.pi])
    return lo, width


def bk_nlo_sums(
    parent_r: np.ndarray,
    spline: Spline,
    edges: np.ndarray,
    lo: np.ndarray,
    width: np.ndarray,
    table: SobolTable,
    first: int,
    n: int,
    p: BKParams | None = None,
    want_hist: bool = False,
    want_magnitude: bool = False,
):
    """Per-parent (sum w, sum w^2, sum K2 w, sum Kf w) and optional histograms.

    ``spline`` is the S spline for the reference integrand (p.stable = 0) and
    the N spline for the cancellation-free one (p.stable = 1).
    """

    p = p or params()
    parent_r = np.ascontiguousarray(parent_r, dtype=np.float64)
    P = parent_r.size
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    bins = edges.shape[-1] - 1
    sums = np.empty((P, 4))
    hist = np.empty((P, 4, bins)) if want_hist else None
    count = np.empty((P, 4, bins)) if want_hist else None
    mag = np.empty(P) if want_magnitude else None
    st = spline.struct()
    lib().orc_bk_nlo_sums(
        P,
        _d(parent_r),
        ctypes.byref(st),
        bins,
        _d(edges),
        _d(np.ascontiguousarray(lo, dtype=np.float64)),
        _d(np.ascontiguousarray(width, dtype=np.float64)),
        _u(table.direction),
        _u(table.shift),
        _u(table.first_q),
        first,
        n,
        ctypes.byref(p),
        _d(sums),
        _d(hist) if want_hist else None,
        _d(count) if want_hist else None,
        _d(mag) if want_magnitude else None,
    )
    if want_magnitude:
        return sums, hist, count, mag
    return sums, hist, count


# ------------------------------------------------------------ RK driver

# Butcher tableaux of the rapidity steps (S/bk/evolution.py:190-236 for
# Kutta's third-order step, :265-276 for the midpoint rule; Euler and
# classical RK4 as the device evolver offers them).
TABLEAUX = {
    "euler": ([[]], [1.0]),
    "rk2": ([[], [0.5]], [0.0, 1.0]),
    "kutta3": ([[], [0.5], [-1.0, 2.0]], [1 / 6, 4 / 6, 1 / 6]),
    "rk4": ([[], [0.5], [0.0, 0.5], [0.0, 0.0, 1.0]], [1 / 6, 2 / 6, 2 / 6, 1 / 6]),
}


def rk_step(N: np.ndarray, h: float, method: str, derivative) -> np.ndarray:
    """One explicit RK step N + h sum_s b_s K_s, K_s = f(N + h sum_j a_sj K_j)."""

    a, b = TABLEAUX[method]
    K = []
    for s in range(len(b)):
        stage = N + h * sum(a[s][j] * K[j] for j in range(len(a[s]))) if s else N
        K.append(derivative(stage))
    return N + h * sum(bs * k for bs, k in zip(b, K))


def rhs(r, N, edges, table, n, p, *, first=0, nr=8, nt=24, k1_rule="fixed"):
    """K1 and K2 + Kf per r node (active set of S/bk/evolution.py:501: index
    >= 1 and N <= 0.99999; others 0): K1 on the fixed rule (nr x nt nodes per
    log-r panel) or the nested adaptive Gauss-Kronrod rule of
    rapidity_derivative_lo (85 x 85 intervals, 1e-3), and the VEGAS sum over
    Sobol points [first, first + n) through each parent's frozen grid."""

    sN, sS = evolution_splines(r, N)
    active = np.nonzero((np.arange(r.size) >= 1) & (N <= 0.99999))[0]
    k1 = np.zeros(r.size)
    k2kf = np.zeros(r.size)
    if active.size:
        if k1_rule == "adaptive":
            k1[active] = k1_adaptive(r[active], sN, p=p)[0]
        else:
            k1[active] = k1_fixed(r[active], sN, nr=nr, nt=nt, p=p)
        lo, width = nlo_bounds(r)
        spline = sN if p.stable else sS
        sums, _, _ = bk_nlo_sums(r[active], spline, edges[active], lo, width, table, first, n, p)
        k2kf[active] = sums[:, 0] / n
    return k1, k2kf


def evolve(r, N0, edges, table, n, p, *, steps, h, method, force_positive=True, k1_rule="fixed"):
    """Fixed-step rapidity evolution (S/bk/evolution.py:85-159): dN/dy = K1 +
    K2 + Kf with non-finite values set to 0, clamp to [0, 1] after each step;
    returns N at every step, [steps + 1, R]."""

    N = np.array(N0, dtype=np.float64)
    out = [N.copy()]

    def derivative(stage):
        k1, k2kf = rhs(r, stage, edges, table, n, p, k1_rule=k1_rule)
        d = k1 + k2kf
        return np.where(np.isfinite(d), d, 0.0)

    for _ in range(steps):
        N = rk_step(N, h, method, derivative)
        if force_positive:
            N = np.clip(N, 0.0, 1.0)
        out.append(N.copy())
    return np.array(out)


def num_threads() -> int:
    return int(lib().orc_num_threads())
