"""The C ABI boundary without a GPU: the library loads, exports every entry
point include/nlo_gfx950.h declares, its host-only functions are right, and
argument errors come back as status codes with messages before any launch."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from nlo_bk_optimized_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "nlo_gfx950.h"


def declared_functions() -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(nlo_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    assert "nlo_bk_nlo" in names and "nlo_sobol_draw" in names and "nlo_bk_evolver_step" in names
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIBRARY))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    header_version = int(re.search(r"#define NLO_API_VERSION (\d+)", HEADER.read_text()).group(1))
    assert _lib.NLO_API_VERSION == header_version
    assert _lib.lib().nlo_api_version() == header_version


def _gauss_legendre_mp(n: int):
    """Nodes and weights by Newton iteration at 40 digits (the exact rule)."""

    import mpmath as mp

    mp.mp.dps = 40
    xs, ws = [], []
    for i in range(n):
        z = mp.cos(mp.pi * (i + mp.mpf(3) / 4) / (n + mp.mpf(1) / 2))
        for _ in range(60):
            p0, p1 = mp.mpf(1), z
            for k in range(2, n + 1):
                p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
            dp = n * (z * p1 - p0) / (z * z - 1)
            z -= p1 / dp
        xs.append(z)
        ws.append(2 / ((1 - z * z) * dp * dp))
    order = sorted(range(n), key=lambda k: xs[k])
    return np.array([float(xs[k]) for k in order]), np.array([float(ws[k]) for k in order])


@pytest.mark.parametrize("n", [1, 2, 8, 24, 48, 64])
def test_gauss_legendre_exact_to_double(n):
    x = np.empty(n)
    w = np.empty(n)
    assert _lib.lib().nlo_gauss_legendre(n, x.ctypes.data, w.ctypes.data) == 0
    if n == 1:
        np.testing.assert_array_equal(x, [0.0])
        np.testing.assert_array_equal(w, [2.0])
        return
    ex, ew = _gauss_legendre_mp(n)
    np.testing.assert_allclose(x, ex, rtol=0, atol=2.3e-16)
    np.testing.assert_allclose(w, ew, rtol=1e-15, atol=0)
    # numpy's leggauss (the reference's table source, K/quadrature.py:18) agrees to its accuracy
    rx, rw = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(x, rx, rtol=0, atol=1e-14)


def test_invalid_arguments_return_status_and_message():
    L = _lib.lib()
    t = _lib.SobolTable()
    t.dim = 0
    assert L.nlo_sobol_draw(ctypes.byref(t), 0, 10, 0, None, None) == 1
    assert b"dim" in L.nlo_last_error()
    assert L.nlo_vegas_map(4, 1, 10, None, None, None, None, None, None, None, None) == 1
    assert b"bins" in L.nlo_last_error()
    # physics is validated before any launch: NF must be 0, 3 or 5, fixed
    # couplings come in pairs, codes must be known
    for kwargs, word in ((dict(nf=4.0), b"NF"), (dict(rc_lo=1), b"fixed coupling"),
                         (dict(order=9), b"Order"), (dict(rc_lo=1, rc_nlo=1, fixed_as=0.0), b"FIXED_AS")):
        ph = _lib.Physics(nc=3.0, nf=3.0, lambda_qcd=0.241, c2=1.0, ksub=0.65, minr=1e-6, symmetrize=1,
                          fixed_as=0.2)
        for k, v in kwargs.items():
            setattr(ph, k, v)
        rc = L.nlo_bk_k1(8, 1, 1, 1, ctypes.byref(ph), 8, 24, 1, None)
        assert rc == 1 and word in L.nlo_last_error(), kwargs
    ph = _lib.Physics(nc=3.0, nf=3.0, lambda_qcd=0.241, c2=1.0, ksub=0.65, minr=1e-6, symmetrize=1)
    assert L.nlo_bk_k1_history(8, 1, 1, 1, ctypes.byref(ph), 8, 24, 1, 1, 1, 0, 1, None) == 1
    assert b"capacity" in L.nlo_last_error()
    # batched driver: every run's grid is checked before anything is allocated
    d = _lib.BatchDesc(runs=2, bins=50, n_samples=4096, precision=1, rk_method=1, step=0.1, nr=8, nt=24,
                       nlo=1)
    d.physics = ph
    d.sobol.dim = 4
    sizes = (ctypes.c_int32 * 2)(64, 2)
    assert L.nlo_bk_batch_workspace(ctypes.byref(d), sizes) == 0
    handle = ctypes.c_void_p()
    assert L.nlo_bk_batch_create(ctypes.byref(d), sizes, 1, 1, 1, 1 << 30, None, ctypes.byref(handle)) == 1
    assert b"run 1" in L.nlo_last_error()
    d.spt = 5
    sizes[1] = 32
    assert L.nlo_bk_batch_create(ctypes.byref(d), sizes, 1, 1, 1, 1 << 30, None, ctypes.byref(handle)) == 1
    assert b"samples per lane" in L.nlo_last_error()
    with pytest.raises(RuntimeError, match="status 1"):
        _lib.check(L.nlo_vegas_adapt(0, 4, 50, None, None, None, 1e-16, None, None), "adapt")


def test_workspace_queries_are_host_only_and_monotone():
    L = _lib.lib()
    a = L.nlo_bk_nlo_workspace(64, 50, 1 << 14, 0)
    b = L.nlo_bk_nlo_workspace(64, 50, 1 << 16, 0)
    c = L.nlo_bk_nlo_workspace(64, 50, 1 << 14, 1)
    assert 0 < a <= b and c > a
    assert L.nlo_bk_nlo_workspace(2, 50, 10, 0) == 0  # R < 3 is invalid
    # one block of partials plus one group of partials
    assert L.nlo_vegas_histogram_workspace(4, 50, 4096) == 2 * (4 * 50 * 2 * 8)
    d = _lib.EvolverDesc()
    d.R, d.bins, d.n_samples, d.nlo = 64, 50, 1 << 14, 1
    assert L.nlo_bk_evolver_workspace(ctypes.byref(d)) > a


STRUCTS = {
    "nlo_bk_physics": _lib.Physics,
    "nlo_sobol_table": _lib.SobolTable,
    "nlo_bk_evolver_desc": _lib.EvolverDesc,
    "nlo_bk_batch_desc": _lib.BatchDesc,
    "nlo_dis_dipole": _lib.DisDipole,
    "nlo_dis_desc": _lib.DisDesc,
    "nlo_dis_item": _lib.DisItem,
    "nlo_dis_batch_desc": _lib.DisBatchDesc,
}


C_FIELD = {"lambda_": "lambda"}  # Python keyword avoided in the ctypes mirror


def test_struct_layouts_match_the_header(tmp_path):
    # sizeof and every field's offsetof from the header itself (host gcc)
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "nlo_gfx950.h"', "int main(void) {"]
    for cname, ct in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in ct._fields_:
            cfield = C_FIELD.get(field, field)
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {cfield}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-D__HIP_PLATFORM_AMD__", f"-I{HEADER.parent}", "-I/opt/rocm/include", str(src),
                    "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, ct in STRUCTS.items():
        assert got[(cname, "size")] == ctypes.sizeof(ct), cname
        for field, _ in ct._fields_:
            assert got[(cname, field)] == getattr(ct, field).offset, (cname, field)


def test_dis_dimensions_and_argument_errors():
    L = _lib.lib()
    assert [L.nlo_dis_dimensions(k) for k in range(12)] == [2, 2, 3, 4, 2, 3, 4, 5, 6, 7, 5, 5]
    assert L.nlo_dis_dimensions(12) == -1
    for name, k in _lib.DIS_KINDS.items():
        assert _lib.DIS_DIMENSIONS[name] == L.nlo_dis_dimensions(k)
    d = _lib.DisDesc()
    d.Q2, d.xbj, d.n_quarks, d.maxr, d.c2_alpha, d.max_alpha, d.nested_points = 9.0, 1e-3, 1, 30.0, 1.0, 0.7, 48
    d.dipole.kind, d.dipole.x0, d.dipole.qs0sqr = 0, 1.0, 1.0
    # invalid contribution, precision and Sobol dimension fail before any launch
    assert L.nlo_dis_integrand(ctypes.byref(d), 12, 4, None, 1, None, None) == 1
    assert L.nlo_dis_integrand(ctypes.byref(d), 0, 4, None, 2, None, None) == 1
    t = _lib.SobolTable()
    t.dim = 4
    assert L.nlo_dis_vegas(ctypes.byref(d), 0, 32, None, ctypes.byref(t), 0, 64, 0, None, None, None, None, 0,
                           None) == 1
    assert b"dimension" in L.nlo_last_error()
    d.nested_points = 4
    t.dim = 2
    assert L.nlo_dis_gauss_legendre(ctypes.byref(d), 0, 24, 0, ctypes.c_void_p(8), None) == 1
    assert b"nested_points" in L.nlo_last_error()
    assert L.nlo_dis_vegas_workspace(9, 32, 1 << 20) > L.nlo_dis_vegas_workspace(9, 32, 1 << 10) > 0
