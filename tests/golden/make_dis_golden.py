"""Write golden vectors for the NLO DIS integrands from the reference's own code.

Runs ONLY in the build container, where the reference tree is mounted
read-only; the vectors it writes (``dis_reference_vectors.npz``) are data and
travel with the tests.  It imports the reference's pure-Python CPU path and
evaluates, in FP64 at seeded random points:

* ``Li2`` (S/numerics/special.py:35-97) and the Bessel K0 / K1 / K2 wrappers;
* ``dis_alpha_s_freeze`` / ``dis_alpha_s_smooth`` (S/coupling.py:49-78);
* ``GBW.dipole_amplitude`` (S/dipole/amplitude.py:38-43);
* ``BKDipole.dipole_amplitude`` uninitialised and initialised, SPLINE_LINEAR
  and LINEAR_LINEAR (S/dipole/amplitude.py:49-270) on the table in
  ``xb_001_mc50w.dat``;
* ``integrand_photon_target_LO`` (S/dis/lo.py:12-51);
* ``integrand_dip_massive`` for all six contributions (S/dis/observables.py:107-172);
* ``integrand_qgunsub_massive`` I1 / I2 / I3 (S/dis/observables.py:175-250);
* the nested I2 / I3 values as the 48-node mapped Gauss-Legendre sums of the
  reference's own un-nested impact factors (its nested form is CUDA-only);
* ``photon_proton_cross_section_LO_d2b`` (deterministic tensor Gauss-Legendre).

Usage: ``python tests/golden/make_dis_golden.py`` (rewrites the .npz).
"""

from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np
import torch

REFERENCE_SRC = Path("/root/reference/torch_nlo_clean/src")
OUT = Path(__file__).with_name("dis_reference_vectors.npz")
TABLE = Path(__file__).with_name("xb_001_mc50w.dat")


def main() -> None:
    sys.path.insert(0, str(REFERENCE_SRC))
    from nlo_torch.coupling import dis_alpha_s_freeze, dis_alpha_s_smooth
    from nlo_torch.dipole.amplitude import GBW, AmplitudeInterpolation, BKDipole
    from nlo_torch.dipole.table import load_bk_table
    from nlo_torch.dis import observables as obs
    from nlo_torch.dis.config import (
        DISConfig,
        NcScheme,
        Polarization,
        Quark,
        QuarkType,
        RunningCouplingIRScheme,
        RunningCouplingScheme,
    )
    from nlo_torch.dis.lo import integrand_photon_target_LO
    from nlo_torch.dis.longitudinal import (
        ILNLOqg_massive_tripole_part_I2,
        ILNLOqg_massive_tripole_part_I3,
    )
    from nlo_torch.dis.transverse import (
        ITNLOqg_massive_tripole_part_I2,
        ITNLOqg_massive_tripole_part_I3,
    )
    from nlo_torch.numerics.special import Li2, bessel_K0, bessel_K1, bessel_K2

    torch.set_default_dtype(torch.float64)
    dt = torch.float64
    rng = np.random.default_rng(20261203)
    out: dict[str, np.ndarray] = {}

    def T(a):
        return torch.tensor(np.asarray(a, dtype=np.float64), dtype=dt)

    # special functions
    xl = np.concatenate([rng.uniform(-40, -1, 40), rng.uniform(-1, 0, 40), rng.uniform(0, 0.5, 40),
                         rng.uniform(0.5, 1, 40), [-1.0, 0.0, 0.5, 1.0, -1e-8, 1 - 1e-9]])
    out["li2_x"] = xl
    out["li2"] = Li2(T(xl)).numpy()
    xb = np.concatenate([np.exp(rng.uniform(math.log(1e-7), math.log(60), 200)), [1e-7, 0.5, 2.0, 8.0, 40.0]])
    out["bessel_x"] = xb
    out["bessel_k0"] = bessel_K0(T(xb)).numpy()
    out["bessel_k1"] = bessel_K1(T(xb)).numpy()
    out["bessel_k2"] = bessel_K2(T(xb)).numpy()

    # couplings
    ra = np.exp(rng.uniform(math.log(1e-4), math.log(60), 100))
    out["alpha_r"] = ra
    out["alpha_freeze_nf4"] = dis_alpha_s_freeze(T(ra), Nf=4, C2=1.0, max_alpha_s=0.7).numpy()
    out["alpha_freeze_nf1_c2"] = dis_alpha_s_freeze(T(ra), Nf=1, C2=2.5, max_alpha_s=0.5).numpy()
    out["alpha_smooth_nf4"] = dis_alpha_s_smooth(T(ra), Nf=4, C2=1.0).numpy()

    # GBW
    gr = np.exp(rng.uniform(math.log(1e-6), math.log(50), 100))
    gy = rng.uniform(-1, 12, 100)
    gb = GBW(Qs0sqr=0.24, lambda_=0.29, gamma=1.1, x0=0.01)
    out["gbw_r"], out["gbw_Y"] = gr, gy
    out["gbw_N"] = gb.dipole_amplitude(T(gr), T(gy)).numpy()

    # BK table dipole
    table = load_bk_table(TABLE)
    bkd = BKDipole(table)
    br = np.concatenate([np.exp(rng.uniform(math.log(1e-6), math.log(40), 120)), [table.r[0].item(), table.r[-1].item()]])
    by = rng.uniform(float(bkd.Y[0]), float(bkd.Y[-1]), br.size)
    by[:10] = rng.uniform(0, float(bkd.Y[0]) + 1e-9, 10)
    out["bk_r"], out["bk_Y"] = br, by
    out["bk_N_local"] = bkd.dipole_amplitude(T(br), T(by)).numpy()
    y_init = 0.5 * (float(bkd.Y[0]) + float(bkd.Y[-1]))
    bkd.initialize_interpolation(y_init)
    out["bk_init_Y"] = np.array(y_init)
    out["bk_N_init"] = bkd.dipole_amplitude(T(br), T(np.full_like(br, y_init))).numpy()
    out["bk_maxr_interpolate"] = np.array(bkd.maxr_interpolate)
    bkl = BKDipole(table, interpolation=AmplitudeInterpolation.LINEAR_LINEAR)
    out["bk_N_bilinear"] = bkl.dipole_amplitude(T(br), T(by)).numpy()
    bkl.initialize_interpolation(y_init)
    out["bk_N_bilinear_init"] = bkl.dipole_amplitude(T(br), T(np.full_like(br, y_init))).numpy()
    out["bk_bilinear_maxr_interpolate"] = np.array(bkl.maxr_interpolate)

    light = Quark(QuarkType.LIGHT)
    charm = Quark(QuarkType.C, 1.4)
    Q2s = (9.0, 2.5)
    xbjs = (1e-3, 3e-4)
    gbw_dis = GBW(Qs0sqr=1.0, lambda_=0.3, gamma=1.0, x0=1.0)

    # LO integrand (one quark set, both polarisations)
    n = 200
    r = np.exp(rng.uniform(math.log(1e-5), math.log(30), n))
    z = rng.uniform(0, 1, n)
    out["lo_r"], out["lo_z"] = r, z
    for pol in ("T", "L"):
        v = integrand_photon_target_LO(T(r), T(z), T(1e-3), T(9.0), getattr(Polarization, pol), (light, charm), gbw_dis)
        out[f"lo_{pol}"] = v.numpy()

    configs = {
        "default": DISConfig(cuda_fusion=False, cuda_nested=False),
        "variant": DISConfig(cuda_fusion=False, cuda_nested=False, maxr=5.0, nc_scheme=NcScheme.LargeNC,
                             rc_scheme=RunningCouplingScheme.PARENT,
                             rc_ir_scheme=RunningCouplingIRScheme.SMOOTH, C2_alpha=2.0, nf_alphas=1),
    }
    out["cfg_variant"] = np.array([5.0, 0, 1, 1, 2.0, 1])  # maxr, finite_nc, parent, smooth, C2, nf

    # dipole-sector integrands
    for dims, contribs in ((2, (("T", "T0"), ("L", "Omega_L_const"))), (3, (("T", "T1"), ("L", "ab"))),
                           (4, (("T", "T2"), ("L", "cd")))):
        x = rng.uniform(0, 1, (n, dims))
        x[:5, 1] = rng.uniform(0, 1e-4, 5)
        out[f"dip_x{dims}"] = x
        for pol, c in contribs:
            for qname, q in (("light", light), ("charm", charm)):
                for cname, cfg in configs.items():
                    v = obs.integrand_dip_massive(T(x), T(Q2s[0]), T(xbjs[0]), getattr(Polarization, pol), q, c,
                                                  gbw_dis, cfg)
                    out[f"dip_{c}_{qname}_{cname}"] = v.numpy()

    # qqbar-g integrands
    for dims, c in ((5, "I1"), (6, "I2"), (7, "I3")):
        x = rng.uniform(0, 1, (n, dims))
        out[f"qg_x{dims}"] = x
        for pol in ("T", "L"):
            for qname, q in (("light", light), ("charm", charm)):
                for cname, cfg in configs.items():
                    v = obs.integrand_qgunsub_massive(T(x), T(Q2s[0]), T(xbjs[0]), getattr(Polarization, pol), q, c,
                                                      gbw_dis, cfg)
                    out[f"qg_{c}_{pol}_{qname}_{cname}"] = v.numpy()

    # nested I2/I3: 48-node mapped GL sums of the un-nested impact factors
    nn = 24
    g5 = rng.uniform(0, 1, (nn, 5))
    out["nest_x5"] = g5
    xg, wg = np.polynomial.legendre.leggauss(48)
    tg = (xg + 1) / 2
    yn, wn = tg**3, wg / 2 * 3 * tg**2
    cfg = configs["default"]
    for pol in ("T", "L"):
        for qname, q in (("light", light), ("charm", charm)):
            Q2 = T(Q2s[0])
            xbj = T(xbjs[0])
            z2min = obs.z2_lower_bound(xbj, Q2)
            xt = T(g5)
            z1 = (1 - z2min) * xt[:, 0]
            z2 = (1 - z1 - z2min) * xt[:, 1] + z2min
            x01 = cfg.maxr * xt[:, 2]
            x02 = cfg.maxr * xt[:, 3]
            phi = 2 * math.pi * xt[:, 4]
            x01sq, x02sq = x01.square(), x02.square()
            x21sq = (x01sq + x02sq - 2 * x01 * x02 * torch.cos(phi)).clamp_min(0)
            f2 = ITNLOqg_massive_tripole_part_I2 if pol == "T" else ILNLOqg_massive_tripole_part_I2
            f3 = ITNLOqg_massive_tripole_part_I3 if pol == "T" else ILNLOqg_massive_tripole_part_I3
            i2 = torch.zeros(nn)
            i3 = torch.zeros(nn)
            for a in range(48):
                ya = torch.full((nn,), yn[a])
                i2 += wn[a] * f2(Q2, q.mass, z1, z2, x01sq, x02sq, x21sq, ya)
                for b in range(48):
                    yb = torch.full((nn,), yn[b])
                    i3 += wn[a] * wn[b] * f3(Q2, q.mass, z1, z2, x01sq, x02sq, x21sq, ya, yb)
            out[f"nest_I2_{pol}_{qname}"] = i2.numpy()
            out[f"nest_I3_{pol}_{qname}"] = i3.numpy()

    # LO cross sections (tensor Gauss-Legendre, deterministic)
    lo_vals = []
    for Q2, xbj in ((9.0, 1e-3), (2.5, 3e-4), (50.0, 1e-2)):
        for pol in ("T", "L"):
            res = obs.photon_proton_cross_section_LO_d2b(T(Q2), T(xbj), getattr(Polarization, pol), gbw_dis,
                                                         configs["default"], quadrature_points=24)
            lo_vals.append([Q2, xbj, 1.0 if pol == "T" else 0.0, res.value.item(), res.error.item()])
    out["lo_sigma"] = np.array(lo_vals)

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({len(out)} arrays)")


if __name__ == "__main__":
    main()
