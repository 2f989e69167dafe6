"""Write golden vectors for the BK physics composites from the reference's own code.

Runs ONLY in the build container, where the reference tree is mounted
read-only; the vectors it writes (``bk_reference_vectors.npz``) are data and
travel with the tests.  It imports the reference's pure-Python CPU path in
FP64 and evaluates only functions that no closed form or library routine pins:

* ``Kernel_lo`` with the Balitsky coupling and the double- and single-log
  resummation (S/bk_kernels/k1.py:20-155);
* ``_lo_integrand`` (S/bk/derivatives.py:434-468);
* ``Kernel_nlo`` (S/bk_kernels/k2.py:8-39) and ``Kernel_nlo_fermion``
  (S/bk_kernels/kf.py:10-41);
* ``_nlo_integrand`` (S/bk/derivatives.py:490-612);
* ``_adapt_vegas_edges`` (S/numerics/integration.py:804-866, eager branch);
* ``MV.dipole_amplitude`` (S/bk/initial_conditions.py:29-35) and the
  evolution splines (S/bk/evolution.py:536-560), as the inputs the composites
  above consume.

Usage: ``python tests/golden/make_bk_golden.py`` (rewrites the .npz).
"""

from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np
import torch

REFERENCE_SRC = Path("/root/reference/torch_nlo_clean/src")
OUT = Path(__file__).with_name("bk_reference_vectors.npz")


def main() -> None:
    sys.path.insert(0, str(REFERENCE_SRC))
    from nlo_torch.bk.config import BKConfig
    from nlo_torch.bk.derivatives import _lo_integrand, _nlo_integrand
    from nlo_torch.bk.evolution import _evolution_interpolators, _r_grid
    from nlo_torch.bk.initial_conditions import MV
    from nlo_torch.bk_kernels.k1 import Kernel_lo
    from nlo_torch.bk_kernels.k2 import Kernel_nlo
    from nlo_torch.bk_kernels.kf import Kernel_nlo_fermion
    from nlo_torch.numerics.integration import _adapt_vegas_edges

    torch.set_default_dtype(torch.float64)
    dt = torch.float64
    rng = np.random.default_rng(20261202)
    out: dict[str, np.ndarray] = {}

    # Configuration of the BK workloads: 64 log-spaced r in [1e-6, 1e2].
    config = BKConfig(RPOINTS=63, MINR=1e-6, MAXR=1e2, CUDA_FUSION=False)
    r = _r_grid(config, device="cpu", dtype=dt)
    ic = MV(qs0sqr=0.2, x0=0.01, ec=1.0, lambdaqcd=0.241, gamma=1.0)
    N = ic.dipole_amplitude(r)
    out["cfg_rpoints"] = np.array(63)
    out["cfg_minr"] = np.array(1e-6)
    out["cfg_maxr"] = np.array(1e2)
    out["r_grid"] = r.numpy()
    out["mv_N"] = N.numpy()
    # Second MV parameter set (gamma != 1, e_c != 1) on extra points,
    # including the small-exponent branch.
    r_extra = torch.tensor(np.exp(rng.uniform(math.log(1e-8), math.log(1e2), 64)), dtype=dt)
    ic2 = MV(qs0sqr=0.37, x0=0.01, ec=7.5, lambdaqcd=0.241, gamma=1.13)
    out["mv2_params"] = np.array([0.37, 1.13, 7.5, 0.241])
    out["mv2_r"] = r_extra.numpy()
    out["mv2_N"] = ic2.dipole_amplitude(r_extra).numpy()

    # Evolution splines (CPU path): interpolator_N (underflow 0 / overflow 1)
    # and interpolator_S.
    interp_N, interp_S = _evolution_interpolators(N, r, config)
    for name, interp in (("splN", interp_N), ("splS", interp_S)):
        s = interp._spline
        out[f"{name}_x"] = s.x.numpy()
        out[f"{name}_coef"] = torch.cat((s.a, s.b, s.c, s.d)).numpy()

    # Kernel_lo at generic points plus edge cases (theta = 0 and z = r gives X = 0;
    # r below 1.01 MINR switches the double-log resummation off).
    n = 400
    rr = np.exp(rng.uniform(math.log(2e-6), math.log(80.0), n))
    zz = np.exp(rng.uniform(math.log(1e-6), math.log(100.0), n))
    th = rng.uniform(0.0, math.pi, n)
    rr[:3] = [0.5, 1e-6, 1.005e-6]
    zz[:3] = [0.5, 0.3, 2e-3]
    th[:3] = [0.0, 1.0, 2.0]
    out["klo_r"], out["klo_z"], out["klo_theta"] = rr, zz, th
    out["klo_value"] = Kernel_lo(
        torch.tensor(rr), torch.tensor(zz), torch.tensor(th), config
    ).numpy()
    out["lo_integrand"] = np.array(
        [
            _lo_integrand(
                torch.tensor(a), torch.tensor([b]), torch.tensor([c]), interp_N, r, config,
                None, None, None,
            ).item()
            for a, b, c in zip(rr, zz, th)
        ]
    )

    # The same two functions with torch.special.bessel_j1 replaced by
    # scipy.special.j1: torch's J1 deviates from scipy's by up to 4.7e-7
    # absolute on [5, 8], and this variant pins the rest of the algorithm at
    # full FP64 precision.
    import scipy.special

    import nlo_torch.bk_kernels.k1 as k1_module

    raw_j1 = k1_module.bessel_J1
    k1_module.bessel_J1 = lambda x: torch.from_numpy(scipy.special.j1(x.numpy()))
    out["klo_value_accj1"] = Kernel_lo(
        torch.tensor(rr), torch.tensor(zz), torch.tensor(th), config
    ).numpy()
    out["lo_integrand_accj1"] = np.array(
        [
            _lo_integrand(
                torch.tensor(a), torch.tensor([b]), torch.tensor([c]), interp_N, r, config,
                None, None, None,
            ).item()
            for a, b, c in zip(rr, zz, th)
        ]
    )
    k1_module.bessel_J1 = raw_j1

    # Kernel_nlo / Kernel_nlo_fermion on explicit distances (including the
    # near-degenerate X Y' ~ X' Y band, where both kernels stay finite).
    m = 400
    Rr = np.exp(rng.uniform(math.log(1e-5), math.log(50.0), m))
    X = np.exp(rng.uniform(math.log(1e-6), math.log(100.0), m))
    Y = np.exp(rng.uniform(math.log(1e-6), math.log(100.0), m))
    X2 = np.exp(rng.uniform(math.log(1e-6), math.log(100.0), m))
    Y2 = np.exp(rng.uniform(math.log(1e-6), math.log(100.0), m))
    ZZ = np.exp(rng.uniform(math.log(1e-6), math.log(100.0), m))
    Y2[:40] = X2[:40] * Y[:40] / X[:40] * (1.0 + rng.uniform(-1e-3, 1e-3, 40))
    out["knlo_in"] = np.stack((Rr, X, Y, X2, Y2, ZZ))
    T = [torch.tensor(a) for a in (Rr, X, Y, X2, Y2, ZZ)]
    out["knlo_value"] = Kernel_nlo(*T).numpy()
    out["knlof_value"] = Kernel_nlo_fermion(*T, config).numpy()

    # _nlo_integrand (eager CPU branch) on the NLO Vegas box.
    k = 600
    lo_r = math.log(0.5 * r[0].item())
    hi_r = math.log(2.0 * r[-1].item())
    lz = rng.uniform(lo_r, hi_r, k)
    lz2 = rng.uniform(lo_r, hi_r, k)
    tz = rng.uniform(0.0, 2 * math.pi, k)
    tz2 = rng.uniform(0.0, 2 * math.pi, k)
    parent = r.numpy()[rng.integers(1, r.numel(), k)]
    out["nlo_in"] = np.stack((parent, lz, lz2, tz, tz2))
    vals = []
    for a, b, c, d, e in zip(parent, lz, lz2, tz, tz2):
        v = _nlo_integrand(
            torch.tensor(a),
            torch.tensor([math.exp(b)]),
            torch.tensor([d]),
            torch.tensor([math.exp(c)]),
            torch.tensor([e]),
            interp_S,
            r,
            config,
        )
        vals.append(v.item() * math.exp(2 * b + 2 * c))
    out["nlo_value"] = np.array(vals)

    # _adapt_vegas_edges (eager branch), 50 and 32 bins.
    for bins in (50, 32):
        edges = np.sort(rng.uniform(0, 1, (4, bins + 1)), axis=1)
        edges[:, 0] = 0.0
        edges[:, -1] = 1.0
        hist = rng.exponential(1.0, (4, bins)) * rng.uniform(0, 10, (4, 1))
        hist[0, 3] = 0.0
        count = rng.integers(0, 200, (4, bins)).astype(np.float64)
        count[1, 5] = 0.0
        new = _adapt_vegas_edges(
            torch.tensor(edges), torch.tensor(hist), torch.tensor(count), cuda_fusion=False
        )
        out[f"adapt{bins}_edges"] = edges
        out[f"adapt{bins}_hist"] = hist
        out[f"adapt{bins}_count"] = count
        out[f"adapt{bins}_new"] = new.numpy()

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
