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
  above consume;
* ``Kernel_lo``, ``_lo_integrand`` and ``_nlo_integrand`` for the other
  couplings, orders and NF (S/bk_kernels/k1.py:37-150,
  S/bk/derivatives.py:596-632) and ``bk_alpha_s`` with the NF = 5 flavour
  thresholds (S/coupling.py:11-46);
* ``_interpolate_history_N`` and the kinematically constrained
  ``_lo_integrand`` (S/bk/derivatives.py:445-468, :712-758);
* ``_adaptive_rk2_to`` (S/bk/evolution.py:239-326) driven by a closed-form
  right-hand side.

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

    physics_options(out, rng, r, N, interp_N, interp_S)
    kinematical(out, rng, r)
    adaptive_rk2(out)

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(out)} arrays)")


# (name, RC_LO, RC_NLO, RESUM_RC, Order, NF, FIXED_AS, C2, KSUB) of the
# non-production physics variants; the row order fixes the vector names.
VARIANTS = (
    ("parent_dls", "PARENT_LO", "PARENT_NLO", "RESUM_RC_PARENT", "NLO_RESUM_DLOG_SLOG", 3, 0.2, 1.0, 0.65),
    ("smallest_lod", "SMALLEST_LO", "SMALLEST_NLO", "RESUM_RC_SMALLEST", "LO_RESUM_DLOG", 3, 0.2, 2.0, 0.65),
    ("fac_nlod", "FAC_LO", "SMALLEST_NLO", "RESUM_RC_FIXED", "NLO_RESUM_DLOG", 3, 0.25, 1.0, 0.65),
    ("beuf_dls", "BEUF_LO", "SMALLEST_NLO", "RESUM_RC_SMALLEST", "NLO_RESUM_DLOG_SLOG", 3, 0.2, 1.0, 0.4),
    ("fixed_dls", "FIXED_LO", "FIXED_NLO", "RESUM_RC_PARENT", "NLO_RESUM_DLOG_SLOG", 3, 0.21, 1.0, 0.65),
    ("nf5_dls", "BALITSKY_LO", "SMALLEST_NLO", "RESUM_RC_SMALLEST", "NLO_RESUM_DLOG_SLOG", 5, 0.2, 1.0, 0.65),
    ("nf0_lo", "BALITSKY_LO", "PARENT_NLO", "RESUM_RC_SMALLEST", "LO", 0, 0.2, 1.0, 0.65),
    ("bal_nlo", "BALITSKY_LO", "SMALLEST_NLO", "RESUM_RC_PARENT", "NLO", 3, 0.2, 1.0, 0.65),
    ("bal_lodls_nf5", "BALITSKY_LO", "PARENT_NLO", "RESUM_RC_PARENT", "LO_RESUM_DLOG_SLOG", 5, 0.2, 1.5, 0.65),
)


def variant_config(v, **extra):
    from nlo_torch.bk.config import (
        BKConfig,
        BKOrder,
        ResummationCoupling,
        RunningCouplingLO,
        RunningCouplingNLO,
    )

    _, lo, nlo, resum, order, nf, fixed_as, c2, ksub = v
    return BKConfig(RPOINTS=63, MINR=1e-6, MAXR=1e2, CUDA_FUSION=False, RC_LO=RunningCouplingLO[lo],
                    RC_NLO=RunningCouplingNLO[nlo], RESUM_RC=ResummationCoupling[resum],
                    Order=BKOrder[order], NF=nf, FIXED_AS=fixed_as, C2=c2, KSUB=ksub, **extra)


def physics_options(out, rng, r, N, interp_N, interp_S):
    """Kernel_lo, _lo_integrand, _nlo_integrand and bk_alpha_s for every
    coupling / order / NF choice other than the production one, with the
    accurate J1 (scipy) in the double-log resummation."""

    import scipy.special

    import nlo_torch.bk_kernels.k1 as k1_module
    from nlo_torch.bk.derivatives import _lo_integrand, _nlo_integrand
    from nlo_torch.coupling import bk_alpha_s

    # columns: RC_LO, RC_NLO, RESUM_RC, Order (as the NLO_RC_*, NLO_RESUM_RC_*,
    # NLO_ORDER_* codes of the C interface), NF, FIXED_AS, C2, KSUB
    out["var_table"] = np.array([[
        ["BALITSKY_LO", "FIXED_LO", "PARENT_LO", "SMALLEST_LO", "FAC_LO", "BEUF_LO"].index(v[1]),
        ["SMALLEST_NLO", "FIXED_NLO", "PARENT_NLO"].index(v[2]),
        ["RESUM_RC_SMALLEST", "RESUM_RC_PARENT", "RESUM_RC_FIXED"].index(v[3]),
        ["NLO_RESUM_DLOG_SLOG", "NLO_RESUM_DLOG", "NLO", "LO_RESUM_DLOG_SLOG", "LO_RESUM_DLOG", "LO"].index(v[4]),
        v[5], v[6], v[7], v[8]] for v in VARIANTS], dtype=np.float64)

    # couplings: NF = 5 thresholds (both sides of 1.3 and 4.5 GeV, the IR
    # cap) and NF = 0, C2 = 2
    ra = np.concatenate([np.exp(rng.uniform(math.log(1e-4), math.log(40.0), 200)),
                         2.0 / np.array([1.3, 4.5]) * (1 + np.array([[-1e-9], [1e-9]])).ravel()])
    out["alpha_r"] = ra
    out["alpha_nf5"] = bk_alpha_s(torch.tensor(ra), NF=5).numpy()
    out["alpha_nf5_c2"] = bk_alpha_s(torch.tensor(ra), NF=5, C2=2.0).numpy()
    out["alpha_nf0_c2"] = bk_alpha_s(torch.tensor(ra), NF=0, C2=2.0).numpy()

    n = 200
    rr = np.exp(rng.uniform(math.log(2e-6), math.log(80.0), n))
    zz = np.exp(rng.uniform(math.log(1e-6), math.log(100.0), n))
    th = rng.uniform(0.0, math.pi, n)
    rr[:2] = [0.5, 1.005e-6]
    zz[:2] = [0.3, 2e-3]
    out["var_klo_in"] = np.stack((rr, zz, th))
    k = 200
    lo_r = math.log(0.5 * r[0].item())
    hi_r = math.log(2.0 * r[-1].item())
    nlo_in = np.stack((r.numpy()[rng.integers(1, r.numel(), k)], rng.uniform(lo_r, hi_r, k),
                       rng.uniform(lo_r, hi_r, k), rng.uniform(0.0, 2 * math.pi, k),
                       rng.uniform(0.0, 2 * math.pi, k)))
    out["var_nlo_in"] = nlo_in

    raw_j1 = k1_module.bessel_J1
    k1_module.bessel_J1 = lambda x: torch.from_numpy(scipy.special.j1(x.numpy()))
    try:
        for v in VARIANTS:
            cfg = variant_config(v)
            name = v[0]
            out[f"var_{name}_klo"] = k1_module.Kernel_lo(
                torch.tensor(rr), torch.tensor(zz), torch.tensor(th), cfg).numpy()
            out[f"var_{name}_lo_integrand"] = np.array([
                _lo_integrand(torch.tensor(a), torch.tensor([b]), torch.tensor([c]), interp_N, r, cfg,
                              None, None, None).item() for a, b, c in zip(rr, zz, th)])
            vals = []
            for a, b, c, d, e in nlo_in.T:
                vv = _nlo_integrand(torch.tensor(a), torch.tensor([math.exp(b)]), torch.tensor([d]),
                                    torch.tensor([math.exp(c)]), torch.tensor([e]), interp_S, r, cfg)
                vals.append(vv.item() * math.exp(2 * b + 2 * c))
            out[f"var_{name}_nlo"] = np.array(vals)
    finally:
        k1_module.bessel_J1 = raw_j1


def kinematical(out, rng, r):
    """_interpolate_history_N and the kinematically constrained
    _lo_integrand (S/bk/derivatives.py:445-468, :712-758) on a synthetic
    history: MV amplitudes with growing saturation scale at y = 0, 0.2, ..."""

    from nlo_torch.bk.config import BKConfig
    from nlo_torch.bk.derivatives import _interpolate_history_N, _lo_integrand
    from nlo_torch.bk.evolution import _evolution_interpolators
    from nlo_torch.bk.initial_conditions import MV

    rows = 6
    hy = torch.tensor([0.2 * k for k in range(rows)])
    hN = torch.stack([MV(qs0sqr=0.2 * math.exp(0.35 * k * 0.2), x0=0.01, ec=1.0, lambdaqcd=0.241,
                         gamma=1.0).dipole_amplitude(r) for k in range(rows)])
    out["kin_hist_y"] = hy.numpy()
    out["kin_hist_N"] = hN.numpy()
    cfg = BKConfig(RPOINTS=63, MINR=1e-6, MAXR=1e2, CUDA_FUSION=False, KINEMATICAL_CONSTRAINT=True,
                   EULER_METHOD=True)
    interp_N, _ = _evolution_interpolators(hN[-1], r, cfg)
    # history lookups: rapidities on and between rows, r on and off the grid
    m = 300
    qr = np.exp(rng.uniform(math.log(3e-7), math.log(300.0), m))
    qy = rng.uniform(0.0, float(hy[-1]), m)
    qy[:20] = hy.numpy()[rng.integers(0, rows, 20)]
    qr[20:30] = r.numpy()[rng.integers(0, r.numel(), 10)]
    qy[-5:] = float(hy[-1])
    out["kin_q_r"], out["kin_q_y"] = qr, qy
    out["kin_q_N"] = _interpolate_history_N(torch.tensor(qr), torch.tensor(qy), interp_N, r, hy,
                                            hN).numpy()
    n = 300
    rr = np.exp(rng.uniform(math.log(2e-6), math.log(80.0), n))
    zz = np.exp(rng.uniform(math.log(1e-6), math.log(100.0), n))
    th = rng.uniform(0.0, math.pi, n)
    out["kin_lo_in"] = np.stack((rr, zz, th))
    y_now = hy[-1]
    out["kin_lo_value"] = np.array([
        _lo_integrand(torch.tensor(a), torch.tensor([b]), torch.tensor([c]), interp_N, r, cfg,
                      y_now, hy, hN).item() for a, b, c in zip(rr, zz, th)])


def adaptive_rk2(out):
    """_adaptive_rk2_to (S/bk/evolution.py:239-326) with the BK derivative
    replaced by a closed-form logistic right-hand side, so that the vectors
    pin the step-size controller alone: accepted steps, rejections, the
    returned step size and the final state over several output steps."""

    import nlo_torch.bk.evolution as evolution
    from nlo_torch.bk.config import BKConfig

    rate = torch.tensor([0.0, 0.3, 1.1, 2.7, 6.0, 15.0, 40.0])

    def derivative(N, r, rapidity, history_y, history_N, config, *, seed, vegas_states=None):
        return rate * N * (1 - N) * (1 + 0.5 * torch.sin(3 * rapidity))

    saved = evolution._evolve_derivative
    evolution._evolve_derivative = derivative
    try:
        for case, (step, h0, abserr, relerr) in enumerate(((0.2, 0.2, 1e-6, 1e-4), (0.5, 2.0, 1e-8, 1e-6),
                                                            (0.1, 0.01, 1e-3, 1e-2))):
            cfg = BKConfig(DE_SOLVER_STEP=step, DE_SOLVER_ABSERR=abserr, DE_SOLVER_RELERR=relerr)
            N = torch.tensor([0.5, 0.01, 0.02, 0.05, 1e-3, 1e-4, 1e-5])
            y, h = 0.0, h0
            traj = []
            for _ in range(4):
                N, y, h = evolution._adaptive_rk2_to(N, y, y + step, h, rate, None, None, cfg, None)
                traj.append(np.concatenate([[y, h], N.numpy()]))
            out[f"ark2_case{case}"] = np.array([step, h0, abserr, relerr])
            out[f"ark2_traj{case}"] = np.array(traj)
        out["ark2_rate"] = rate.numpy()
        out["ark2_N0"] = np.array([0.5, 0.01, 0.02, 0.05, 1e-3, 1e-4, 1e-5])
    finally:
        evolution._evolve_derivative = saved


if __name__ == "__main__":
    main()
