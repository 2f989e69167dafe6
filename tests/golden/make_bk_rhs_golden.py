"""Write golden values of the reference's adaptive K1 right-hand side.

Runs ONLY in the build container, where the reference tree is mounted
read-only; the values it writes (``bk_rhs_reference_vectors.npz``) are data
and travel with the tests.  It imports the reference's pure-Python CPU path
in FP64 and evaluates ``rapidity_derivative_lo`` (S/bk/derivatives.py:43-110:
the nested adaptive Gauss-Kronrod 21 rule in ln z over [ln r_0, ln r_{R-1}]
around the same rule in theta over [0, pi]) for a few parents of the 64-node
MV workload (R = 64 on [1e-6, 1e2], Qs0^2 = 0.2, gamma = 1), with the
production physics and INTACCURACY = 1e-3 and 1e-6.

torch.special.bessel_j1, which the reference's double-log resummation calls,
deviates from the true J1 by up to 4.7e-7 absolute on [5, 8]; a second set
replaces it with scipy.special.j1, the accurate J1 the oracle and the device
use, so that the adaptive rule takes the same decisions and the values agree
to rounding.

Usage: ``python tests/golden/make_bk_rhs_golden.py`` (rewrites the .npz).
"""

from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np
import torch

REFERENCE_SRC = Path("/root/reference/torch_nlo_clean/src")
OUT = Path(__file__).with_name("bk_rhs_reference_vectors.npz")
PARENTS = (1, 12, 25, 34, 40, 45)


def main() -> None:
    sys.path.insert(0, str(REFERENCE_SRC))
    import scipy.special
    from nlo_torch.bk.config import BKConfig
    from nlo_torch.bk.derivatives import rapidity_derivative_lo
    from nlo_torch.bk.evolution import _evolution_interpolators, _r_grid
    from nlo_torch.bk.initial_conditions import MV
    from nlo_torch.bk_kernels import k1 as k1_module

    torch.set_default_dtype(torch.float64)
    dt = torch.float64
    out: dict[str, np.ndarray] = {}
    ic = MV(qs0sqr=0.2, x0=0.01, ec=1.0, lambdaqcd=0.241, gamma=1.0)
    raw_j1 = k1_module.bessel_J1
    for tag, accuracy in (("e3", 1e-3), ("e6", 1e-6)):
        config = BKConfig(RPOINTS=63, MINR=1e-6, MAXR=1e2, CUDA_FUSION=False, INTACCURACY=accuracy)
        r = _r_grid(config, device="cpu", dtype=dt)
        N = ic.dipole_amplitude(r)
        interp_N, _ = _evolution_interpolators(N, r, config)
        out["r_grid"] = r.numpy()
        out["N0"] = N.numpy()
        out["parents"] = np.array(PARENTS)
        for j1_tag, j1 in (("torchj1", raw_j1),
                           ("accj1", lambda x: torch.from_numpy(scipy.special.j1(x.numpy())))):
            k1_module.bessel_J1 = j1
            vals, errs, evals = [], [], []
            for p in PARENTS:
                t0 = time.time()
                res = rapidity_derivative_lo(r[p:p + 1], interp_N, r, config)
                vals.append(float(res.value))
                errs.append(float(res.error))
                evals.append(int(res.n_eval))
                print(f"{tag} {j1_tag} parent {p} r={float(r[p]):.3e}: K1 = {vals[-1]:.15e} +- {errs[-1]:.2e} "
                      f"({evals[-1]} evaluations, {time.time() - t0:.1f} s)", flush=True)
            out[f"k1_{tag}_{j1_tag}"] = np.array(vals)
            out[f"k1err_{tag}_{j1_tag}"] = np.array(errs)
            out[f"k1evals_{tag}_{j1_tag}"] = np.array(evals)
        k1_module.bessel_J1 = raw_j1
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
