"""nlo_bk_nlo_noalloc (the K2 + Kf sums with the partials in the library's
static device scratch, no allocation) through the C ABI against nlo_bk_nlo
with a caller workspace: bit for bit, on consecutive calls of decreasing
size (the scratch keeps the previous call's contents), and the refusal of a
workload whose partials exceed the scratch."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from tests.bk_helpers import random_edges, workload


def _noalloc(dev, r, N, t64, t32, edges, table, n, config, precision, sum_only=False):
    from nlo_bk_optimized_amd import _lib

    R = r.numel()
    sums = torch.full((R, 4), float("nan"), dtype=torch.float64, device=r.device)
    ph = dev.physics(config)
    code = dev.PRECISION_CODES[precision] | (_lib.NLO_BK_SUM_ONLY if sum_only else 0)
    rc = _lib.lib().nlo_bk_nlo_noalloc(R, dev.ptr(r), dev.ptr(N), dev.ptr(t64), dev.ptr(t32), dev.ptr(edges),
                                       edges.shape[-1] - 1, ctypes.byref(table), 0, n, code, ctypes.byref(ph),
                                       dev.ptr(sums), dev.stream_handle(r.device))
    return rc, sums


@pytest.mark.gpu
def test_noalloc_equals_workspace_entry_point_on_shrinking_calls():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from nlo_bk_optimized_amd import device as dev
    from nlo_bk_optimized_amd.bk.config import BKConfig

    cfg = BKConfig(MINR=1e-6)
    # the second call is smaller: it must not see the first call's partials
    for R, n, seed, precision, sum_only in ((128, 1 << 18, 3, "fp32", True), (24, 4096 + 640, 5, "fp32", False),
                                            (40, 1 << 14, 7, "fp64_stable", False), (17, 2048, 9, "fp64", True)):
        print(f"noalloc R={R} n={n} {precision} sum_only={sum_only}", flush=True)
        r, N = workload(R=R)
        rt = torch.tensor(r, dtype=torch.float64, device="cuda")
        Nt = torch.tensor(N, dtype=torch.float64, device="cuda")
        edges = torch.tensor(random_edges(R, 50, seed=seed), dtype=torch.float64, device="cuda")
        t64, t32 = dev.spline_build(rt, Nt)
        table = dev.sobol_table(4, seed)
        want, _, _ = dev.bk_nlo(rt, Nt, t64, t32, edges, table, n, cfg, precision=precision, sum_only=sum_only)
        rc, got = _noalloc(dev, rt, Nt, t64, t32, edges, table, n, cfg, precision, sum_only)
        torch.cuda.synchronize()
        assert rc == 0
        assert torch.equal(got, want)
        assert torch.count_nonzero(got[:, 0]) > 0


def test_noalloc_refuses_more_than_its_scratch():
    """Checked on the host before anything is launched (no device needed)."""

    from nlo_bk_optimized_amd import _lib
    from nlo_bk_optimized_amd import device as dev
    from nlo_bk_optimized_amd.bk.config import BKConfig

    L = _lib.lib()
    R, bins, n = 1024, 50, 1 << 22
    assert L.nlo_bk_nlo_workspace(R, bins, n, 0) > _lib.NLO_BK_NOALLOC_SCRATCH_BYTES
    assert L.nlo_bk_nlo_workspace(R, bins, n // 2, 0) <= _lib.NLO_BK_NOALLOC_SCRATCH_BYTES
    table = _lib.SobolTable()
    table.dim = 4
    ph = dev.physics(BKConfig())
    fake = ctypes.c_void_p(256)  # never dereferenced: the size check comes first
    rc = L.nlo_bk_nlo_noalloc(R, fake, fake, fake, fake, fake, bins, ctypes.byref(table), 0, n,
                              _lib.NLO_PREC_FP32, ctypes.byref(ph), fake, None)
    assert rc == 3  # NLO_ERR_WORKSPACE
    msg = L.nlo_last_error().decode()
    assert "static scratch" in msg and "nlo_bk_nlo with a workspace" in msg
    # a histogram-free workload inside the limit passes the same checks; other
    # argument errors are reported as for nlo_bk_nlo
    rc = L.nlo_bk_nlo_noalloc(2, fake, fake, fake, fake, fake, bins, ctypes.byref(table), 0, n,
                              _lib.NLO_PREC_FP32, ctypes.byref(ph), fake, None)
    assert rc == 1 and "nlo_bk_nlo_noalloc: R=2" in L.nlo_last_error().decode()
