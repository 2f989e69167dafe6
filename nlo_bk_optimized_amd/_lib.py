"""ctypes binding of ``libnlo_gfx950.so`` (the C ABI in include/nlo_gfx950.h).

PyTorch supplies device memory and the current HIP stream; every numeric
result comes from the library's HIP kernels.  A missing library is an error:
there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import torch

LIBRARY = Path(__file__).resolve().parent / "libnlo_gfx950.so"

NLO_OK = 0
NLO_API_VERSION = 3  # include/nlo_gfx950.h
NLO_SOBOL_U32, NLO_SOBOL_F32, NLO_SOBOL_F64 = 0, 1, 2
NLO_PREC_FP64, NLO_PREC_FP32, NLO_PREC_FP64_STABLE = 0, 1, 2
NLO_BK_SUM_ONLY = 0x100
NLO_BK_NOALLOC_SCRATCH_BYTES = 16 << 20
NLO_RK_EULER, NLO_RK_MIDPOINT, NLO_RK_KUTTA3, NLO_RK_CLASSIC4, NLO_RK_ADAPTIVE2 = 0, 1, 2, 3, 4
NLO_K1_FIXED, NLO_K1_ADAPTIVE = 0, 1
MAX_GK_INTERVALS = 100
# physics choice codes of nlo_bk_physics (zero = production)
RC_LO_CODES = {"BALITSKY_LO": 0, "FIXED_LO": 1, "PARENT_LO": 2, "SMALLEST_LO": 3, "FAC_LO": 4, "BEUF_LO": 5}
RC_NLO_CODES = {"SMALLEST_NLO": 0, "FIXED_NLO": 1, "PARENT_NLO": 2}
RESUM_RC_CODES = {"RESUM_RC_SMALLEST": 0, "RESUM_RC_PARENT": 1, "RESUM_RC_FIXED": 2}
ORDER_CODES = {"NLO_RESUM_DLOG_SLOG": 0, "NLO_RESUM_DLOG": 1, "NLO": 2, "LO_RESUM_DLOG_SLOG": 3,
               "LO_RESUM_DLOG": 4, "LO": 5}
SOBOL_MAXDIM = 8
SOBOL_BITS = 30
NLO_FN_K0, NLO_FN_K1, NLO_FN_K2, NLO_FN_LI2 = 0, 1, 2, 3
NLO_DIPOLE_GBW, NLO_DIPOLE_BK_TABLE = 0, 1
NLO_DIS_METHOD_VEGAS, NLO_DIS_METHOD_GAUSS_LEGENDRE = 0, 1
DIS_BATCH_MAX_ITERATIONS = 16
DIS_KINDS = {
    "LO": 0, "T0": 1, "T1": 2, "T2": 3, "Omega_L_const": 4, "ab": 5, "cd": 6,
    "I1": 7, "I2": 8, "I3": 9, "I2_nested": 10, "I3_nested": 11,
}
DIS_DIMENSIONS = {"LO": 2, "T0": 2, "T1": 3, "T2": 4, "Omega_L_const": 2, "ab": 3, "cd": 4,
                  "I1": 5, "I2": 6, "I3": 7, "I2_nested": 5, "I3_nested": 5}
DIS_MAX_QUARKS = 8
FIXED_SOBOL_MAX_ROUNDS = 8
FIXED_SOBOL_MAX_REPLICATES = 32


def spline_words(R: int) -> int:
    """NLO_SPLINE_WORDS(R): [r_min, r_max, ln r (R), N records, S records]."""

    return 2 + R + 8 * (R - 1)


class Physics(ctypes.Structure):
    _fields_ = [
        ("nc", ctypes.c_double),
        ("nf", ctypes.c_double),
        ("lambda_qcd", ctypes.c_double),
        ("c2", ctypes.c_double),
        ("ksub", ctypes.c_double),
        ("minr", ctypes.c_double),
        ("symmetrize", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("rc_lo", ctypes.c_int32),
        ("rc_nlo", ctypes.c_int32),
        ("resum_rc", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("fixed_as", ctypes.c_double),
    ]


class SobolTable(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("direction", (ctypes.c_uint32 * SOBOL_BITS) * SOBOL_MAXDIM),
        ("shift", ctypes.c_uint32 * SOBOL_MAXDIM),
        ("first", ctypes.c_uint32 * SOBOL_MAXDIM),
    ]


class EvolverDesc(ctypes.Structure):
    _fields_ = [
        ("R", ctypes.c_int32),
        ("bins", ctypes.c_int32),
        ("n_samples", ctypes.c_int64),
        ("first_index", ctypes.c_int64),
        ("precision", ctypes.c_int32),
        ("rk_method", ctypes.c_int32),
        ("step", ctypes.c_double),
        ("force_positive", ctypes.c_int32),
        ("nr", ctypes.c_int32),
        ("nt", ctypes.c_int32),
        ("nlo", ctypes.c_int32),
        ("physics", Physics),
        ("sobol", SobolTable),
        ("abserr", ctypes.c_double),
        ("relerr", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("kinematical", ctypes.c_int32),
        ("history_capacity", ctypes.c_int32),
        ("d_history_y", ctypes.c_void_p),
        ("d_history_N", ctypes.c_void_p),
        ("k1_rule", ctypes.c_int32),
        ("k1_rint", ctypes.c_int32),
        ("k1_tint", ctypes.c_int32),
        ("k1_reserved", ctypes.c_int32),
        ("k1_epsrel", ctypes.c_double),
    ]


class BatchDesc(ctypes.Structure):
    """nlo_bk_batch_desc."""

    _fields_ = [
        ("runs", ctypes.c_int32),
        ("bins", ctypes.c_int32),
        ("n_samples", ctypes.c_int64),
        ("first_index", ctypes.c_int64),
        ("precision", ctypes.c_int32),
        ("rk_method", ctypes.c_int32),
        ("step", ctypes.c_double),
        ("force_positive", ctypes.c_int32),
        ("nr", ctypes.c_int32),
        ("nt", ctypes.c_int32),
        ("nlo", ctypes.c_int32),
        ("spt", ctypes.c_int32),
        ("grid_blocks", ctypes.c_int32),
        ("physics", Physics),
        ("sobol", SobolTable),
        ("k1_rule", ctypes.c_int32),
        ("k1_rint", ctypes.c_int32),
        ("k1_tint", ctypes.c_int32),
        ("k1_reserved", ctypes.c_int32),
        ("k1_epsrel", ctypes.c_double),
    ]


class DisDipole(ctypes.Structure):
    """nlo_dis_dipole."""

    _fields_ = [
        ("kind", ctypes.c_int32),
        ("initialized", ctypes.c_int32),
        ("bilinear", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("qs0sqr", ctypes.c_double),
        ("lambda_", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("x0", ctypes.c_double),
        ("R", ctypes.c_int32),
        ("T", ctypes.c_int32),
        ("d_r", ctypes.c_void_p),
        ("d_Y", ctypes.c_void_p),
        ("d_N", ctypes.c_void_p),
        ("y_init", ctypes.c_double),
        ("maxr_interpolate", ctypes.c_double),
        ("d_init_coef", ctypes.c_void_p),
        ("d_local_coef", ctypes.c_void_p),
    ]


class DisDesc(ctypes.Structure):
    """nlo_dis_desc."""

    _fields_ = [
        ("Q2", ctypes.c_double),
        ("xbj", ctypes.c_double),
        ("transverse", ctypes.c_int32),
        ("n_quarks", ctypes.c_int32),
        ("mass", ctypes.c_double * DIS_MAX_QUARKS),
        ("charge", ctypes.c_double * DIS_MAX_QUARKS),
        ("maxr", ctypes.c_double),
        ("finite_nc", ctypes.c_int32),
        ("parent_coupling", ctypes.c_int32),
        ("smooth_coupling", ctypes.c_int32),
        ("mass_dependent_x", ctypes.c_int32),
        ("c2_alpha", ctypes.c_double),
        ("max_alpha", ctypes.c_double),
        ("nf_alphas", ctypes.c_int32),
        ("nested_points", ctypes.c_int32),
        ("dipole", DisDipole),
    ]


class DisItem(ctypes.Structure):
    """nlo_dis_item."""

    _fields_ = [
        ("desc", DisDesc),
        ("contribution", ctypes.c_int32),
        ("method", ctypes.c_int32),
        ("n", ctypes.c_int64),
        ("sobol_table", ctypes.c_int32),
        ("max_iterations", ctypes.c_int32),
        ("min_iterations", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("first_index", ctypes.c_int64),
    ]


class DisBatchDesc(ctypes.Structure):
    """nlo_dis_batch_desc."""

    _fields_ = [
        ("n_items", ctypes.c_int32),
        ("n_tables", ctypes.c_int32),
        ("items", ctypes.POINTER(DisItem)),
        ("tables", ctypes.POINTER(SobolTable)),
        ("d_initial_edges", ctypes.c_void_p),
        ("bins", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("epsrel", ctypes.c_double),
        ("epsabs", ctypes.c_double),
    ]


class FixedSobolItem(ctypes.Structure):
    """nlo_dis_fixed_sobol_item."""

    _fields_ = [
        ("desc", DisDesc),
        ("table_set", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class FixedSobolDesc(ctypes.Structure):
    """nlo_dis_fixed_sobol_desc."""

    _fields_ = [
        ("n_items", ctypes.c_int32),
        ("n_tables", ctypes.c_int32),
        ("items", ctypes.POINTER(FixedSobolItem)),
        ("tables", ctypes.POINTER(SobolTable)),
        ("training_samples", ctypes.c_int32),
        ("training_rounds", ctypes.c_int32),
        ("samples_per_replicate", ctypes.c_int32),
        ("replicates", ctypes.c_int32),
        ("bins", ctypes.c_int32),
        ("fold_angle", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("i1_weight", ctypes.c_double),
        ("epsrel", ctypes.c_double),
        ("d_grid", ctypes.c_void_p),
    ]


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_size = ctypes.c_size_t

# name -> (restype, argtypes)
SIGNATURES: dict[str, tuple] = {
    "nlo_api_version": (ctypes.c_int, []),
    "nlo_last_error": (ctypes.c_char_p, []),
    "nlo_gauss_legendre": (ctypes.c_int, [_i32, _vp, _vp]),
    "nlo_sobol_draw": (ctypes.c_int, [ctypes.POINTER(SobolTable), _i64, _i64, _i32, _vp, _vp]),
    "nlo_vegas_map": (ctypes.c_int, [_i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "nlo_vegas_histogram_workspace": (_size, [_i32, _i32, _i64]),
    "nlo_vegas_histogram": (ctypes.c_int, [_i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _size, _vp]),
    "nlo_vegas_adapt": (ctypes.c_int, [_i32, _i32, _i32, _vp, _vp, _vp, _f64, _vp, _vp]),
    "nlo_bk_spline_build": (ctypes.c_int, [_i32, _vp, _vp, _i32, _vp, _vp, _vp]),
    "nlo_bk_k1": (ctypes.c_int, [_i32, _vp, _vp, _vp, ctypes.POINTER(Physics), _i32, _i32, _vp, _vp]),
    "nlo_bk_k1_history": (
        ctypes.c_int,
        [_i32, _vp, _vp, _vp, ctypes.POINTER(Physics), _i32, _i32, _vp, _vp, _vp, _i32, _vp, _vp],
    ),
    "nlo_bk_k1_adaptive": (
        ctypes.c_int,
        [_i32, _vp, _vp, _vp, ctypes.POINTER(Physics), _i32, _i32, _f64, _vp, _vp, _vp, _i32, _vp, _vp,
         _vp, _vp],
    ),
    "nlo_bk_nlo_workspace": (_size, [_i32, _i32, _i64, _i32]),
    "nlo_bk_nlo": (
        ctypes.c_int,
        [
            _i32,
            _vp,
            _vp,
            _vp,
            _vp,
            _vp,
            _i32,
            ctypes.POINTER(SobolTable),
            _i64,
            _i64,
            _i32,
            ctypes.POINTER(Physics),
            _vp,
            _vp,
            _vp,
            _vp,
            _size,
            _vp,
        ],
    ),
    "nlo_bk_nlo_noalloc": (
        ctypes.c_int,
        [_i32, _vp, _vp, _vp, _vp, _vp, _i32, ctypes.POINTER(SobolTable), _i64, _i64, _i32,
         ctypes.POINTER(Physics), _vp, _vp],
    ),
    "nlo_bk_rhs_finalize": (ctypes.c_int, [_i32, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "nlo_bk_evolver_workspace": (_size, [ctypes.POINTER(EvolverDesc)]),
    "nlo_bk_evolver_create": (
        ctypes.c_int,
        [ctypes.POINTER(EvolverDesc), _vp, _vp, _vp, _size, _vp, ctypes.POINTER(_vp)],
    ),
    "nlo_bk_evolver_state": (_vp, [_vp]),
    "nlo_bk_evolver_nonfinite_flag": (_vp, [_vp]),
    "nlo_bk_evolver_k1": (_vp, [_vp]),
    "nlo_bk_evolver_active_evaluations": (_vp, [_vp]),
    "nlo_bk_evolver_k2kf": (_vp, [_vp]),
    "nlo_bk_evolver_step": (ctypes.c_int, [_vp, _vp]),
    "nlo_bk_evolver_adaptive_state": (
        ctypes.c_int,
        [_vp, ctypes.POINTER(_f64), ctypes.POINTER(_f64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)],
    ),
    "nlo_bk_evolver_destroy": (ctypes.c_int, [_vp]),
    "nlo_bk_batch_workspace": (_size, [ctypes.POINTER(BatchDesc), _vp]),
    "nlo_bk_batch_create": (
        ctypes.c_int,
        [ctypes.POINTER(BatchDesc), _vp, _vp, _vp, _vp, _size, _vp, ctypes.POINTER(_vp)],
    ),
    "nlo_bk_batch_state": (_vp, [_vp]),
    "nlo_bk_batch_nonfinite_flag": (_vp, [_vp]),
    "nlo_bk_batch_active_evaluations": (_vp, [_vp]),
    "nlo_bk_batch_step": (ctypes.c_int, [_vp, _vp]),
    "nlo_bk_batch_tuning": (
        ctypes.c_int,
        [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_f64)],
    ),
    "nlo_bk_batch_destroy": (ctypes.c_int, [_vp]),
    "nlo_dis_special": (ctypes.c_int, [_i32, _i64, _vp, _i32, _vp, _vp]),
    "nlo_dis_dipole_amplitude": (ctypes.c_int, [ctypes.POINTER(DisDipole), _i64, _vp, _vp, _vp, _vp]),
    "nlo_dis_bk_initialize": (ctypes.c_int, [ctypes.POINTER(DisDipole), _f64, _vp, _vp, _vp]),
    "nlo_dis_dimensions": (ctypes.c_int, [_i32]),
    "nlo_dis_integrand": (ctypes.c_int, [ctypes.POINTER(DisDesc), _i32, _i64, _vp, _i32, _vp, _vp]),
    "nlo_dis_vegas_workspace": (_size, [_i32, _i32, _i64]),
    "nlo_dis_vegas": (
        ctypes.c_int,
        [ctypes.POINTER(DisDesc), _i32, _i32, _vp, ctypes.POINTER(SobolTable), _i64, _i64, _i32, _vp, _vp,
         _vp, _vp, _size, _vp],
    ),
    "nlo_dis_gauss_legendre": (ctypes.c_int, [ctypes.POINTER(DisDesc), _i32, _i32, _i32, _vp, _vp]),
    "nlo_dis_bk_local_splines": (ctypes.c_int, [ctypes.POINTER(DisDipole), _vp, _vp]),
    "nlo_dis_batch_workspace": (_size, [ctypes.POINTER(DisBatchDesc)]),
    "nlo_dis_batch_run": (ctypes.c_int, [ctypes.POINTER(DisBatchDesc), _vp, _vp, _size, _vp]),
    "nlo_dis_qg_i23": (ctypes.c_int, [ctypes.POINTER(DisDesc), _i64, _vp, _i32, _vp, _vp, _vp]),
    "nlo_dis_endpoint_map": (ctypes.c_int, [_i32, _i64, _vp, _vp, _i32, _f64, _i32, _vp, _vp, _vp, _vp, _vp]),
    "nlo_dis_fixed_sobol_workspace": (_size, [ctypes.POINTER(FixedSobolDesc)]),
    "nlo_dis_fixed_sobol": (ctypes.c_int, [ctypes.POINTER(FixedSobolDesc), _vp, _vp, _vp, _size, _vp]),
}

_lib: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    """Load the HIP library once; raise if it has not been built."""

    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise RuntimeError(
                f"{LIBRARY} is missing: build it with `python -m nlo_bk_optimized_amd.build` "
                "(there is no CPU fallback)"
            )
        handle = ctypes.CDLL(str(LIBRARY))
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = restype
            fn.argtypes = argtypes
        version = handle.nlo_api_version()
        if version != NLO_API_VERSION:
            raise RuntimeError(f"{LIBRARY} has interface version {version}, these bindings expect "
                               f"{NLO_API_VERSION}: rebuild it with `python -m nlo_bk_optimized_amd.build`")
        _lib = handle
    return _lib


def check(status: int, what: str = "") -> None:
    """Raise RuntimeError for a non-zero status, with the library's message."""

    if status != NLO_OK:
        message = lib().nlo_last_error().decode(errors="replace")
        raise RuntimeError(f"{what or 'libnlo_gfx950'} failed (status {status}): {message}")


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(device: torch.device | None = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_cuda(t: torch.Tensor, dtype: torch.dtype, name: str) -> None:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
