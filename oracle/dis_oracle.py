"""FP64 numpy restatement of the reference's NLO DIS integrands.

TEST INFRASTRUCTURE ONLY: the reference the HIP DIS path is tested against.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py`` may import this
module; the product package never imports it.

Every function restates the reference formula it cites (S/ =
torch_nlo_clean/src/nlo_torch/).  Arithmetic is IEEE double in numpy; the
special functions are scipy's K0 and K1, which pin the reference's
torch.special calls, and a Bernoulli-series dilogarithm pinned against mpmath.

Parity status: every public function here is checked against vectors that the
reference's own FP64 code produced (tests/golden/dis_reference_vectors.npz,
written by tests/golden/make_dis_golden.py), and the pieces with closed forms
(couplings, GBW, LO wave function, Li2 identities, Gauss-Legendre nesting) also
against those forms in tests/test_dis_oracle.py.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from scipy import special as sp

LAMBDA_QCD = 0.241
ALPHA_EM = 1 / 137.035999
NC = 3.0
CF = (NC * NC - 1) / (2 * NC)
Q0SQR = 1.0
ZMIN = 1e-6
EULER_GAMMA = 0.5772156649015329

# ----------------------------------------------------------------- specials


def k0(x):
    return sp.k0(x)


def k1(x):
    return sp.k1(x)


def k2(x):
    """K2 = K0 + 2 K1 / x, +inf at 0 (S/numerics/special.py:26-32)."""

    x = np.asarray(x, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        v = k0(x) + 2 * k1(x) / np.where(x == 0, 1.0, x)
    return np.where(x == 0, np.inf, v)


def _bernoulli_over_factorial(count):
    """B_n / (n + 1)! for n < count (B_1 = -1/2), exact rationals rounded once."""

    from fractions import Fraction
    from math import comb

    B = [Fraction(1)]
    for m in range(1, count):
        B.append(-sum(comb(m + 1, k) * B[k] for k in range(m)) / Fraction(m + 1))
    return [float(B[n] / math.factorial(n + 1)) for n in range(count)]


_LI2_COEF = _bernoulli_over_factorial(40)


def li2(x):
    """Real dilogarithm for x <= 1 (S/numerics/special.py:35-97).

    Reduced to [-1, 1/2] by Li2(x) = pi^2/6 - ln x ln(1-x) - Li2(1-x) (x > 1/2)
    and Li2(x) = -pi^2/6 - ln^2(-x)/2 - Li2(1/x) (x < -1); there the Bernoulli
    series Li2 = sum_n B_n u^(n+1)/(n+1)!, u = -ln(1-x), |u| <= ln 2.
    """

    x = np.asarray(x, dtype=np.float64)

    def core(v):
        u = -np.log1p(-v)
        acc = np.zeros_like(u)
        for c in reversed(_LI2_COEF):
            acc = acc * u + c
        return acc * u

    out = np.empty_like(x)
    hi = x > 0.5
    lo = x < -1
    mid = ~(hi | lo)
    out[mid] = core(x[mid])
    xh = x[hi]
    with np.errstate(divide="ignore", invalid="ignore"):
        out[hi] = np.where(xh == 1, math.pi**2 / 6, math.pi**2 / 6 - np.log(xh) * np.log1p(-xh) - core(1 - xh))
    xl = x[lo]
    out[lo] = -(math.pi**2) / 6 - 0.5 * np.log(-xl) ** 2 - core(1 / xl)
    return out


# ---------------------------------------------------------------- couplings


def alpha_freeze(r, nf, c2=1.0, max_alpha=0.7, lambda_qcd=LAMBDA_QCD, nc=NC):
    """S/coupling.py:49-66: 1/(b0 ln(4 C2/(r^2 Lambda^2))), capped at max_alpha (and
    max_alpha where the log argument is below 1)."""

    r = np.asarray(r, dtype=np.float64)
    b0 = (11 * nc - 2 * nf) / (12 * math.pi)
    arg = 4 * c2 / (r * r * lambda_qcd**2)
    with np.errstate(divide="ignore", invalid="ignore"):
        a = 1 / (b0 * np.log(arg))
    return np.where((a > max_alpha) | (arg < 1), max_alpha, a)


def alpha_smooth(r, nf, c2=1.0, lambda_qcd=LAMBDA_QCD, nc=NC):
    """S/coupling.py:69-97: 1/(b0 0.2 logaddexp(10 ln 2.5, 5 ln(4 C2/(r^2 Lambda^2))))."""

    r = np.asarray(r, dtype=np.float64)
    b0 = (11 * nc - 2 * nf) / (12 * math.pi)
    mu0 = 2 / 0.2 * math.log(2.5)
    scale = np.log(4 * c2 / (r * r * lambda_qcd**2)) / 0.2
    return 1 / (b0 * 0.2 * np.logaddexp(mu0, scale))


# ------------------------------------------------------------------- dipoles


@dataclass(frozen=True)
class GBWParams:
    qs0sqr: float = 0.1
    lambda_: float = 0.3
    gamma: float = 1.0
    x0: float = 1.0


def gbw(r, Y, p: GBWParams):
    """S/dipole/amplitude.py:38-43."""

    r = np.asarray(r, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    y_eff = np.maximum(Y, math.log(1 / p.x0))
    qs2 = p.qs0sqr * np.exp(p.lambda_ * y_eff)
    arg = np.power(r * r * qs2, p.gamma) / 4
    return np.where(np.abs(arg) < 1e-7, arg, 1 - np.exp(-arg))


def natural_spline_eval(x_nodes, y_nodes, x):
    """Natural cubic spline through (x_nodes, y_nodes) evaluated at x (one spline).

    Used for BKDipole's r-splines (S/numerics/interpolation.py NaturalCubicSpline):
    pinned against scipy.interpolate.CubicSpline(bc_type='natural') in the tests.
    """

    from scipy.interpolate import CubicSpline

    return CubicSpline(x_nodes, y_nodes, bc_type="natural")(x)


class BKDipoleOracle:
    """N(r, Y) from a BK table (S/dipole/amplitude.py:49-270)."""

    def __init__(self, r, y, N, x0, bilinear=False):
        self.r = np.asarray(r, dtype=np.float64)
        self.N = np.asarray(N, dtype=np.float64)
        self.Y = np.asarray(y, dtype=np.float64) + max(0.0, math.log(1 / x0))
        self.bilinear = bilinear

    def _bilinear(self, r, Y):
        """Bilinear interpolation in r and Y (:176-202)."""

        R, T = self.r.size, self.Y.size
        ri = np.searchsorted(self.r, r, side="right") - 1
        yi = np.searchsorted(self.Y, Y, side="right") - 1
        ri2 = np.minimum(ri + 1, R - 1)
        yi2 = np.minimum(yi + 1, T - 1)
        r1, r2, Y1, Y2 = self.r[ri], self.r[ri2], self.Y[yi], self.Y[yi2]
        dr = np.where(ri2 != ri, r2 - r1, 1.0)
        dY = np.where(yi2 != yi, Y2 - Y1, 1.0)
        N11, N12 = self.N[yi, ri], self.N[yi, ri2]
        N21, N22 = self.N[yi2, ri], self.N[yi2, ri2]
        v = (N11 * (r2 - r) * (Y2 - Y) + N12 * (r - r1) * (Y2 - Y) + N21 * (r2 - r) * (Y - Y1)
             + N22 * (r - r1) * (Y - Y1)) / (dr * dY)
        last = N11 + (r - r1) / dr * (N12 - N11)
        v = np.where(yi2 == yi, last, v)
        return np.where(ri2 == ri, 1.0, v)

    def _local(self, r, Y):
        """Local natural spline over 7 r nodes (9 at the top) of the Y-linear
        interpolated table rows (:184-240)."""

        R = self.r.size
        rind = np.searchsorted(self.r, r, side="right") - 1
        yind = np.searchsorted(self.Y, Y, side="right") - 1
        out = np.empty_like(r)
        for i in range(r.size):
            if rind[i] + 3 > R - 1:
                start, pts = R - 9, 9
            else:
                start, pts = max(rind[i] - 3, 0), 7
            idx = np.arange(start, start + pts)
            y1 = yind[i]
            y2 = min(y1 + 1, self.Y.size - 1)
            lower, upper = self.N[y1, idx], self.N[y2, idx]
            frac = (Y[i] - self.Y[y1]) / (self.Y[y2] - self.Y[y1]) if y2 != y1 else 0.0
            n = lower + frac * (upper - lower)
            out[i] = natural_spline_eval(self.r[idx], n, r[i])
        return out

    def amplitude(self, r, Y):
        """Uninitialised path: clamp r into the table, freeze Y below the first row,
        local spline, clamp to [0, 1] (:133-168)."""

        r, Y = np.broadcast_arrays(np.asarray(r, dtype=np.float64), np.asarray(Y, dtype=np.float64))
        r = r.reshape(-1).copy()
        Y = Y.reshape(-1).copy()
        r = np.where(r < self.r[0], self.r[0] * 1.000001, r)
        r = np.where(r > self.r[-1], self.r[-1] * 0.999999, r)
        Y = np.where((Y >= 0) & (Y < self.Y[0]), self.Y[0], Y)
        if np.any((Y < self.Y[0]) | (Y > self.Y[-1])):
            raise ValueError("Y outside the table")
        return np.clip(self._bilinear(r, Y) if self.bilinear else self._local(r, Y), 0.0, 1.0)

    def initialized(self, Y):
        """Initialised path at one Y (:111-131, :170-175, :242-260): a natural spline
        in r through the amplitude at the (slightly moved) table nodes, 1 beyond the
        first r (scanning up from 0.01) where N >= 0.99999."""

        Y = float(Y)
        if 0 <= Y <= self.Y[0]:
            Y = float(self.Y[0])
        er = self.r.copy()
        er[0] *= 1.0001
        er[-1] *= 0.9999
        n_nodes = self.amplitude(er, np.full_like(er, Y))
        from scipy.interpolate import CubicSpline

        spline = CubicSpline(self.r, n_nodes, bc_type="natural")
        # maxr_interpolate: step search from r = 0.01 (:242-260)
        step, prev, rr, maxr = 2.0, 0.1, 0.01, -1.0
        for _ in range(41):
            if rr > self.r[-1]:
                break
            v = float(np.clip(spline(np.clip(rr, self.r[0] * 1.000001, self.r[-1] * 0.999999)), 0, 1))
            if v >= 0.99999:
                if step < 1e-2:
                    maxr = rr
                    break
                step /= 1.5
                rr = prev
            prev = rr
            rr += step

        def fn(r):
            r = np.asarray(r, dtype=np.float64)
            r = np.where(r < self.r[0], self.r[0] * 1.000001, r)
            r = np.where(r > self.r[-1], self.r[-1] * 0.999999, r)
            v = spline(r)
            if maxr > 0:
                v = np.where(r >= maxr, 1.0, v)
            return np.clip(v, 0.0, 1.0)

        fn.maxr = maxr
        return fn


# ---------------------------------------------------------------------- LO


@dataclass(frozen=True)
class QuarkSpec:
    mass: float
    charge: float


LIGHT = QuarkSpec(0.002, math.sqrt(2 / 3))  # QuarkType.LIGHT default mass via Quark(LIGHT)
CHARM = QuarkSpec(1.4, 2 / 3)


def lo_wavefunction(r, z, Q2, quark: QuarkSpec, transverse: bool):
    """|Psi|^2 without the dipole (S/dis/lo.py:25-47)."""

    r = np.asarray(r, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    mf = quark.mass
    eps = np.sqrt(Q2 * z * (1 - z) + mf * mf)
    a = r * eps
    active = (a >= 1e-7) & (a <= 5e2)
    s = np.where(active, a, 1.0)
    if transverse:
        c = quark.charge**2 * ((1 - 2 * z + 2 * z * z) * eps * eps * k1(s) ** 2 + (mf * k0(s)) ** 2)
    else:
        c = quark.charge**2 * 4 * Q2 * z * z * (1 - z) ** 2 * k0(s) ** 2
    return np.where(active, c, 0.0)


# ------------------------------------------------------ dipole-sector terms


def _omega_t_v_u(Q, z, mf):
    g = math.sqrt(1 + 4 * (mf / Q) ** 2)
    return (1 + 1 / (2 * z)) * (np.log(1 - z) + g * np.log((1 + g) / (1 + g - 2 * z))) - 1 / (2 * z) * (
        (z + 0.5) * (1 - g) + mf * mf / (Q * Q)
    ) * np.log((z * (1 - z) * Q * Q + mf * mf) / (mf * mf))


def omega_t_v(Q, z, mf):
    """S/dis/transverse.py:49-62."""

    return _omega_t_v_u(Q, z, mf) + _omega_t_v_u(Q, 1 - z, mf)


def _omega_t_n_u(Q, z, mf):
    g = math.sqrt(1 + 4 * (mf / Q) ** 2)
    return (1 + z - 2 * z * z) / z * (np.log(1 - z) + g * np.log((1 + g) / (1 + g - 2 * z))) + (1 - z) / z * (
        (0.5 + z) * (g - 1) - mf * mf / (Q * Q)
    ) * np.log((z * (1 - z) * Q * Q + mf * mf) / (mf * mf))


def omega_t_n(Q, z, mf):
    """S/dis/transverse.py:65-76."""

    return _omega_t_n_u(Q, z, mf) - _omega_t_n_u(Q, 1 - z, mf)


def l_dip(Q2, z, mf):
    """S/dis/longitudinal.py:133-143."""

    g = math.sqrt(1 + 4 * mf * mf / Q2)
    s = (
        li2(1 / (1 - (1 - g) / (2 * z)))
        + li2(1 / (1 - (1 + g) / (2 * z)))
        + li2(1 / (1 - (1 - g) / (2 * (1 - z))))
        + li2(1 / (1 - (1 + g) / (2 * (1 - z))))
    )
    return s - (math.pi**2 / 6 - 0.5 * np.log(z / (1 - z)) ** 2)


def omega_l_v(Q2, z, mf):
    """S/dis/longitudinal.py:118-130."""

    g = math.sqrt(1 + 4 * mf * mf / Q2)
    return (
        1 / (2 * z) * (np.log(1 - z) + g * np.log((1 + g) / (1 + g - 2 * z)))
        + 1 / (2 * (1 - z)) * (np.log(z) + g * np.log((1 + g) / (1 + g - 2 * (1 - z))))
        + 1 / (4 * z * (1 - z)) * (g - 1 + 2 * mf * mf / Q2) * np.log((z * (1 - z) * Q2 + mf * mf) / (mf * mf))
    )


def it_dip_0(Q2, z1, x01sq, mf):
    """S/dis/transverse.py:27-46."""

    x01 = np.sqrt(x01sq)
    Q = math.sqrt(Q2)
    kz = np.sqrt(z1 * (1 - z1) * Q2 + mf * mf)
    common = -(math.pi**2) / 3 + math.pi**2 / 6 + 0.5 * np.log(z1 / (1 - z1)) ** 2 + omega_t_v(Q, z1, mf) + l_dip(Q2, z1, mf)
    t1 = (kz * k1(x01 * kz)) ** 2 * ((z1 * z1 + (1 - z1) ** 2) * (2.5 + common) + (2 * z1 - 1) / 2 * omega_t_n(Q, z1, mf))
    t2 = (mf * k0(x01 * kz)) ** 2 * (3 + common)
    return t1 + t2


def _it_v1_u(Q, z, mf, r, xi):
    kz = np.sqrt(z * (1 - z) * Q * Q + mf * mf)
    dk = xi / (1 - xi) * (1 - z) * mf * mf
    sh = np.sqrt(kz * kz + dk)
    diff = sh * k1(r * sh) - kz * k1(r * kz)
    t1 = 1 / xi * (2 * np.log(xi) / (1 - xi) - (1 + xi) / 2) * diff
    t2 = -(np.log(xi) / (1 - xi) ** 2 + z / (1 - xi) + z / 2) * (1 - z) * mf * mf / sh * k1(r * sh)
    return t1 + t2


def _it_vms1_u(Q, z, mf, r, xi):
    kz = np.sqrt(z * (1 - z) * Q * Q + mf * mf)
    sh = np.sqrt(kz * kz + xi / (1 - xi) * (1 - z) * mf * mf)
    t1 = 1 / xi * (2 * np.log(xi) / (1 - xi) - (1 + xi) / 2) * (k0(r * sh) - k0(r * kz))
    t2 = (-1.5 * (1 - z) / (1 - xi) + (1 - z) / 2) * k0(r * sh)
    return t1 + t2


def it_dip_1(Q2, z1, x01sq, mf, xi):
    """S/dis/transverse.py:79-148."""

    x01 = np.sqrt(x01sq)
    Q = math.sqrt(Q2)
    kz = np.sqrt(z1 * (1 - z1) * Q2 + mf * mf)
    v1 = _it_v1_u(Q, z1, mf, x01, xi) + _it_v1_u(Q, 1 - z1, mf, x01, xi)
    vms1 = _it_vms1_u(Q, z1, mf, x01, xi) + _it_vms1_u(Q, 1 - z1, mf, x01, xi)
    return kz * k1(x01 * kz) * (z1 * z1 + (1 - z1) ** 2) * v1 + mf * mf * k0(x01 * kz) * vms1


def _k1_difference_stable(r, kb, dk2):
    """S/dis/transverse.py:13-24."""

    b2 = kb * kb
    s2 = b2 + dk2
    valid = (s2 > 0) & (kb > 0)
    rel = np.abs(dk2) / (b2 + np.abs(dk2) + 1e-300)
    lin = -0.5 * r * dk2 * k0(r * kb)
    sh = np.sqrt(np.maximum(s2, 0))
    direct = sh * k1(r * sh) - kb * k1(r * kb)
    v = np.where(rel < 1e-5, lin, direct)
    return np.where(valid, v, 0.0)


def _it_v2_u(Q, z, mf, r, ych, yu):
    chi = z * ych
    u = (1 - yu) / yu
    kz = np.sqrt(z * (1 - z) * Q * Q + mf * mf)
    kc = np.sqrt(chi * (1 - chi) * Q * Q + mf * mf)
    dk = u * (1 - z) / (1 - chi) * kc * kc
    kd = _k1_difference_stable(r, kz, dk)
    sh = np.sqrt(kz * kz + dk)
    t1 = -1 / (1 - chi) / (u * (u + 1)) * mf * mf / (kc * kc) * (2 * chi + (u / (u + 1)) ** 2 / z * (z - chi) * (1 - 2 * chi)) * kd
    t2 = (
        -1 / (1 - chi) ** 2 / (u + 1) * (z - chi)
        * (1 - 2 * u / (1 + u) * (z - chi) + (u / (u + 1)) ** 2 / z * (z - chi) ** 2)
        * mf * mf / sh * k1(r * sh)
    )
    return z / (yu * yu) * (t1 + t2)


def _it_vms2_u(Q, z, mf, r, ych, yu):
    chi = z * ych
    u = (1 - yu) / yu
    kz = np.sqrt(z * (1 - z) * Q * Q + mf * mf)
    kc = np.sqrt(chi * (1 - chi) * Q * Q + mf * mf)
    sh = np.sqrt(kz * kz + u * (1 - z) / (1 - chi) * kc * kc)
    t1 = 1 / (1 - chi) / (u + 1) ** 2 * (-z - u / (1 + u) * (z + u * chi) / z * (chi - (1 - z))) * k0(r * sh)
    t2 = (
        1 / (u + 1) ** 3
        * (kz * kz / (kc * kc) * (1 + u * chi * (1 - chi) / (z * (1 - z)))
           - mf * mf / (kc * kc) * chi / (1 - chi) * (2 * (1 + u) ** 2 / u + u / (z * (1 - z)) * (z - chi) ** 2))
        * (k0(r * sh) - k0(r * kz))
    )
    return z / (yu * yu) * (t1 + t2)


def _it_n_u(Q, z, mf, r, ych, yu):
    chi = z * ych
    u = (1 - yu) / yu
    kz = np.sqrt(z * (1 - z) * Q * Q + mf * mf)
    kc = np.sqrt(chi * (1 - chi) * Q * Q + mf * mf)
    dk = u * (1 - z) / (1 - chi) * kc * kc
    sh = np.sqrt(kz * kz + dk)
    t1 = 2 * (1 - z) / z / (u + 1) ** 3 * ((2 + u) * u * z + u * u * chi) * sh * k1(r * sh)
    t2 = (
        2 * (1 - z) / z / (u + 1) ** 3 * mf * mf / (kc * kc)
        * (z / (1 - z) + chi / (1 - chi) * (u - 2 * z - 2 * u * chi))
        * (sh * k1(r * sh) - kz * k1(r * kz))
    )
    return z / (yu * yu) * (t1 + t2)


def it_dip_2(Q2, z1, x01sq, mf, ych, yu):
    """S/dis/transverse.py:151-306."""

    x01 = np.sqrt(x01sq)
    Q = math.sqrt(Q2)
    kz = np.sqrt(z1 * (1 - z1) * Q2 + mf * mf)
    v2 = _it_v2_u(Q, z1, mf, x01, ych, yu) + _it_v2_u(Q, 1 - z1, mf, x01, ych, yu)
    n = _it_n_u(Q, z1, mf, x01, ych, yu) - _it_n_u(Q, 1 - z1, mf, x01, ych, yu)
    vms2 = _it_vms2_u(Q, z1, mf, x01, ych, yu) + _it_vms2_u(Q, 1 - z1, mf, x01, ych, yu)
    return kz * k1(x01 * kz) * ((z1 * z1 + (1 - z1) ** 2) * v2 + (2 * z1 - 1) / 2 * n) + mf * mf * k0(x01 * kz) * vms2


def il_dip_const(Q2, z, r, mf):
    """S/dis/longitudinal.py:95-115."""

    ff = 4 * Q2 * (z * (1 - z)) ** 2
    a = np.sqrt(Q2 * z * (1 - z) + mf * mf) * r
    active = a >= 1e-7
    s = np.where(active, a, 1.0)
    v = ff * k0(s) ** 2 * (2.5 + (-(math.pi**2) / 3 + math.pi**2 / 6) + 0.5 * np.log(z / (1 - z)) ** 2 + omega_l_v(Q2, z, mf) + l_dip(Q2, z, mf))
    return np.where(active, v, 0.0)


def il_dip_ab(Q2, z1, r, mf, xi):
    """S/dis/longitudinal.py:74-92."""

    ff = 4 * Q2 * (z1 * (1 - z1)) ** 2
    kz = np.sqrt(z1 * (1 - z1) * Q2 + mf * mf)
    b1 = k0(kz * r)
    b2 = k0(np.sqrt(kz * kz + (1 - z1) * xi / (1 - xi) * mf * mf) * r)
    b3 = k0(np.sqrt(kz * kz + z1 * xi / (1 - xi) * mf * mf) * r)
    return ff * (b1 / xi * (-2 * np.log(xi) / (1 - xi) + (1 + xi) / 2) * (2 * b1 - b2 - b3))


def il_dip_cd(Q2, z1, r, mf, xi, x):
    """S/dis/longitudinal.py:12-71."""

    ff = 4 * Q2 * (z1 * (1 - z1)) ** 2
    kz = np.sqrt(z1 * (1 - z1) * Q2 + mf * mf)
    a = kz * r
    active = a >= 1e-7
    s = np.where(active, a, 1.0)
    clm1 = z1 * z1 * (1 - xi) / (1 - z1) * (-xi * xi + x * (1 - xi) * (1 + (1 - xi) * (1 + z1 * xi / (1 - z1))) / (x * (1 - xi) + xi / (1 - z1)))
    clm2 = (1 - z1) ** 2 * (1 - xi) / z1 * (-xi * xi + x * (1 - xi) * (1 + (1 - xi) * (1 + (1 - z1) * xi / z1)) / (x * (1 - xi) + xi / z1))
    kap1 = xi * mf * mf / ((1 - xi) * (1 - x) * (x * (1 - xi) + xi / (1 - z1))) * (xi * (1 - x) + x * (1 - z1 * (1 - xi) / (1 - z1)))
    kap2 = xi * mf * mf / ((1 - xi) * (1 - x) * (x * (1 - xi) + xi / z1)) * (xi * (1 - x) + x * (1 - (1 - z1) * (1 - xi) / z1))
    b1 = k0(s)
    v = b1 * mf * mf * (
        (b1 - k0(r * np.sqrt(kz * kz / (1 - x) + kap1))) * clm1
        / ((1 - xi) * (1 - x) * (x * (1 - xi) + xi / (1 - z1)) * (x / (1 - x) * kz * kz + kap1))
        + (b1 - k0(r * np.sqrt(kz * kz / (1 - x) + kap2))) * clm2
        / ((1 - xi) * (1 - x) * (x * (1 - xi) + xi / z1) * (x / (1 - x) * kz * kz + kap2))
    )
    return ff * np.where(active, v, 0.0)


# ------------------------------------------------------------------- qqbar g


def tripole_variables(Q, z1, z2, x01sq, x02sq, x21sq):
    """S/dis/transverse.py:854-909 (the longitudinal set is its first twelve)."""

    z0 = 1 - z1 - z2
    x20x21 = -0.5 * (x01sq - x21sq - x02sq)
    v = dict(z0=z0, x20x21=x20x21)
    v["Qbar_j"] = Q * np.sqrt(z1 * (1 - z1))
    v["Qbar_k"] = Q * np.sqrt(z0 * (1 - z0))
    v["omega_j"] = z0 * z2 / (z1 * (z0 + z2) ** 2)
    v["omega_k"] = z1 * z2 / (z0 * (z1 + z2) ** 2)
    v["lambda_j"] = z1 * z2 / z0
    v["lambda_k"] = z0 * z2 / z1
    v["x2_j"] = np.sqrt(x02sq)
    v["x2_k"] = np.sqrt(x21sq)
    v["x3_j"] = np.sqrt(z0 * z0 / (z0 + z2) ** 2 * x02sq + x21sq - 2 * z0 / (z0 + z2) * x20x21)
    v["x3_k"] = np.sqrt(z1 * z1 / (z1 + z2) ** 2 * x21sq + x02sq - 2 * z1 / (z1 + z2) * x20x21)
    v["x2j_x3j"] = x20x21 - z0 / (z0 + z2) * x02sq
    v["x2k_x3k"] = -x20x21 + z1 / (z1 + z2) * x21sq
    v["x2j_x3k"] = -x02sq + z1 / (z1 + z2) * x20x21
    v["x2k_x3j"] = x21sq - z0 / (z0 + z2) * x20x21
    v["x3j_x3k"] = z0 / (z0 + z2) * x02sq + z1 / (z1 + z2) * x21sq - (1 + z0 * z1 / ((z0 + z2) * (z1 + z2))) * x20x21
    return v


def g_simplified(a, b, qbar, mf, x2, x3, omega, lam, y):
    """S/dis/longitudinal.py:303-336."""

    order = a + b - 2
    num = y * lam * mf * mf + qbar * qbar + mf * mf
    den = y * x3 * x3 + omega * x2 * x2
    arg = np.sqrt(num * den / y)
    bes = {0: k0, 1: k1, 2: k2}[order](arg)
    return y ** (-0.5 * (2 - a + b)) * 2.0 ** (a + b - 1) * omega ** (b - 1) * (num / den) ** (0.5 * (a + b - 2)) * bes


def _g_bar(a, b, qbar, mf, x2, x3, omega, lam, y):
    return g_simplified(a, b, qbar, mf, x2, x3, omega, lam, y) - g_simplified(a, b, qbar, mf, x2, x3, omega, 0.0 * lam, y)


def _g12_sing(qbar, mf, x2, x3, omega):
    return k0(np.sqrt((qbar * qbar + mf * mf) * (x3 * x3 + omega * x2 * x2))) / (x2 * x2)


def _g22_sing(qbar, mf, x2, x3, omega):
    s = x3 * x3 + omega * x2 * x2
    return np.sqrt((qbar * qbar + mf * mf) / s) * k1(np.sqrt((qbar * qbar + mf * mf) * s)) / (x2 * x2)


def _h(qbar, mf, x2, x3, omega, lam):
    s = x3 * x3 + omega * x2 * x2
    m = qbar * qbar + mf * mf * (1 + lam)
    return 4 * np.sqrt(m / s) * k1(np.sqrt(m * s))


def it_qg_uvsub(Q2, mf, z1, z2, x01sq, x02sq, x21sq):
    """ITNLOqg_massive_dipole_uvsub = IT_dipole_jk_I1 + IT_dipole_jkm_I1
    (S/dis/transverse.py:308-321, 371-399, 437-463)."""

    v = tripole_variables(math.sqrt(Q2), z1, z2, x01sq, x02sq, x21sq)
    z0 = v["z0"]
    eg = math.exp(EULER_GAMMA)
    mj = v["Qbar_j"] ** 2 + mf * mf
    mk = v["Qbar_k"] ** 2 + mf * mf
    pj = (2 * z0 * (z0 + z2) + z2 * z2) / (z0 + z2) ** 2
    pk = (2 * z1 * (z1 + z2) + z2 * z2) / (z1 + z2) ** 2
    ej = -np.exp(-v["x2_j"] ** 2 / (x01sq * eg)) / v["x2_j"] ** 2
    ek = -np.exp(-v["x2_k"] ** 2 / (x01sq * eg)) / v["x2_k"] ** 2
    jk = pj * (1 - 2 * z1 * (1 - z1)) * mj * ej * k1(np.sqrt(x01sq * mj)) ** 2 + pk * (1 - 2 * z0 * (1 - z0)) * mk * ek * k1(np.sqrt(x01sq * mk)) ** 2
    jkm = mf * mf * (pj * ej * k0(np.sqrt(x01sq * mj)) ** 2 + pk * ek * k0(np.sqrt(x01sq * mk)) ** 2)
    return jk + jkm


def it_qg_i1(Q2, mf, z1, z2, x01sq, x02sq, x21sq):
    """ITNLOqg_massive_tripole_part_I1 (S/dis/transverse.py:324-337, 402-434, 466-585)."""

    v = tripole_variables(math.sqrt(Q2), z1, z2, x01sq, x02sq, x21sq)
    z0 = v["z0"]
    pj = (2 * z0 * (z0 + z2) + z2 * z2) / (z0 + z2) ** 2
    pk = (2 * z1 * (z1 + z2) + z2 * z2) / (z1 + z2) ** 2
    mj = v["Qbar_j"] ** 2 + mf * mf
    mk = v["Qbar_k"] ** 2 + mf * mf
    sj = v["x3_j"] ** 2 + v["omega_j"] * v["x2_j"] ** 2
    sk = v["x3_k"] ** 2 + v["omega_k"] * v["x2_k"] ** 2
    jk = pj * (1 - 2 * z1 * (1 - z1)) * mj / v["x2_j"] ** 2 * v["x3_j"] ** 2 / sj * k1(np.sqrt(sj) * np.sqrt(mj)) ** 2 \
        + pk * (1 - 2 * z0 * (1 - z0)) * mk / v["x2_k"] ** 2 * v["x3_k"] ** 2 / sk * k1(np.sqrt(sk) * np.sqrt(mk)) ** 2
    jkm = mf * mf * (pj / v["x2_j"] ** 2 * k0(np.sqrt(sj) * np.sqrt(mj)) ** 2 + pk / v["x2_k"] ** 2 * k0(np.sqrt(sk) * np.sqrt(mk)) ** 2)
    g22j = _g22_sing(v["Qbar_j"], mf, v["x2_j"], v["x3_j"], v["omega_j"])
    g22k = _g22_sing(v["Qbar_k"], mf, v["x2_k"], v["x3_k"], v["omega_k"])
    hj = _h(v["Qbar_j"], mf, v["x2_j"], v["x3_j"], v["omega_j"], v["lambda_j"])
    hk = _h(v["Qbar_k"], mf, v["x2_k"], v["x3_k"], v["omega_k"], v["lambda_k"])
    cross = _cross(v, z1, z2)
    F = 0.5 * (
        4 / ((z0 + z2) * (z1 + z2)) * cross * g22j * g22k
        - (z0 + z2) * z1 * z2 / (z1 + z2) ** 2 * v["x2j_x3j"] * hk * g22j
        + (z1 + z2) * z0 * z2 / (z0 + z2) ** 2 * v["x2k_x3k"] * hj * g22k
        - z0 * z0 * z1 * z2 / (z0 + z2) ** 3 * v["x2j_x3j"] * hj * g22j
        + z1 * z1 * z0 * z2 / (z1 + z2) ** 3 * v["x2k_x3k"] * hk * g22k
        + (z0 * z2) ** 2 / (8 * (z0 + z2) ** 4) * hj * hj
        + (z1 * z2) ** 2 / (8 * (z1 + z2) ** 4) * hk * hk
    )
    g12j = _g12_sing(v["Qbar_j"], mf, v["x2_j"], v["x3_j"], v["omega_j"])
    g12k = _g12_sing(v["Qbar_k"], mf, v["x2_k"], v["x3_k"], v["omega_k"])
    Fm = 0.5 * mf * mf * (-((2 * z0 + z2) * (2 * z1 + z2) + z2 * z2) * v["x20x21"] * 8 * g12j * 8 * g12k / (32 * (z0 + z2) * (z1 + z2)))
    return jk + jkm + F + Fm


def _cross(v, z1, z2):
    z0 = v["z0"]
    return z2 * (z0 - z1) ** 2 * (v["x2j_x3j"] * v["x2k_x3k"] - v["x2k_x3j"] * v["x2j_x3k"]) - (z1 * (z0 + z2) + z0 * (z1 + z2)) * (
        z0 * (z0 + z2) + z1 * (z1 + z2)
    ) * v["x20x21"] * v["x3j_x3k"]


def _t_moments(v, mf, y):
    """The eight transverse G moments at inner variable y."""

    j = (v["Qbar_j"], mf, v["x2_j"], v["x3_j"], v["omega_j"])
    k = (v["Qbar_k"], mf, v["x2_k"], v["x3_k"], v["omega_k"])
    return dict(
        b22j=_g_bar(2, 2, *j, v["lambda_j"], y), b22k=_g_bar(2, 2, *k, v["lambda_k"], y),
        b12j=_g_bar(1, 2, *j, v["lambda_j"], y), b12k=_g_bar(1, 2, *k, v["lambda_k"], y),
        g21j=g_simplified(2, 1, *j, v["lambda_j"], y), g21k=g_simplified(2, 1, *k, v["lambda_k"], y),
        g11j=g_simplified(1, 1, *j, v["lambda_j"], y), g11k=g_simplified(1, 1, *k, v["lambda_k"], y),
    )


def _it_i2_from_moments(v, mf, z1, z2, m):
    """_IT_tripole_I2 with the y-dependent G values given (S/dis/transverse.py:587-734)."""

    z0 = v["z0"]
    pj = (2 * z0 * (z0 + z2) + z2 * z2) / (z0 + z2) ** 2
    pk = (2 * z1 * (z1 + z2) + z2 * z2) / (z1 + z2) ** 2
    mj = v["Qbar_j"] ** 2 + mf * mf
    mk = v["Qbar_k"] ** 2 + mf * mf
    sj = v["x3_j"] ** 2 + v["omega_j"] * v["x2_j"] ** 2
    sk = v["x3_k"] ** 2 + v["omega_k"] * v["x2_k"] ** 2
    g12j = _g12_sing(v["Qbar_j"], mf, v["x2_j"], v["x3_j"], v["omega_j"])
    g12k = _g12_sing(v["Qbar_k"], mf, v["x2_k"], v["x3_k"], v["omega_k"])
    g22j = _g22_sing(v["Qbar_j"], mf, v["x2_j"], v["x3_j"], v["omega_j"])
    g22k = _g22_sing(v["Qbar_k"], mf, v["x2_k"], v["x3_k"], v["omega_k"])
    hj = _h(v["Qbar_j"], mf, v["x2_j"], v["x3_j"], v["omega_j"], v["lambda_j"])
    hk = _h(v["Qbar_k"], mf, v["x2_k"], v["x3_k"], v["omega_k"], v["lambda_k"])
    jk = pj * (1 - 2 * z1 * (1 - z1)) * m["b22j"] * v["x3_j"] ** 2 / 8 * np.sqrt(mj / sj) * k1(np.sqrt(sj) * np.sqrt(mj)) \
        + pk * (1 - 2 * z0 * (1 - z0)) * m["b22k"] * v["x3_k"] ** 2 / 8 * np.sqrt(mk / sk) * k1(np.sqrt(sk) * np.sqrt(mk))
    jkm = mf * mf * (pj * m["b12j"] / 4 * k0(np.sqrt(sj) * np.sqrt(mj)) + pk * m["b12k"] / 4 * k0(np.sqrt(sk) * np.sqrt(mk)))
    cross = _cross(v, z1, z2)
    F = 0.5 * (
        cross / (4 * (z0 + z2) * (z1 + z2)) * (m["b22k"] * g22j + m["b22j"] * g22k)
        - (z0 + z2) * z1 * z2 / (16 * (z1 + z2) ** 2) * v["x2j_x3j"] * hk * m["b22j"]
        + (z1 + z2) * z0 * z2 / (16 * (z0 + z2) ** 2) * v["x2k_x3k"] * hj * m["b22k"]
        - z0 * z0 * z1 * z2 / (16 * (z0 + z2) ** 3) * v["x2j_x3j"] * hj * m["b22j"]
        + z1 * z1 * z0 * z2 / (16 * (z1 + z2) ** 3) * v["x2k_x3k"] * hk * m["b22k"]
    )
    Fm = 0.5 * mf * mf * (
        -z0 * z1 * z2 * z2 / (16 * (z0 + z2) ** 3) * v["x2j_x3j"] * m["g21j"] * 8 * g12j
        + z0 * z1 * z2 * z2 / (16 * (z1 + z2) ** 3) * v["x2k_x3k"] * m["g21k"] * 8 * g12k
        - ((2 * z0 + z2) * (2 * z1 + z2) + z2 * z2) * v["x20x21"] / (32 * (z0 + z2) * (z1 + z2)) * (m["b12k"] * 8 * g12j + m["b12j"] * 8 * g12k)
        - (z0 * z2) ** 2 / (16 * (z0 + z2) * (z1 + z2) ** 2) * v["x2j_x3k"] * m["g21k"] * 8 * g12j
        + (z1 * z2) ** 2 / (16 * (z0 + z2) ** 2 * (z1 + z2)) * v["x2k_x3j"] * m["g21j"] * 8 * g12k
        - z0 * z1 * z2 * z2 / (16 * (z0 + z2) ** 3) * v["x2j_x3j"] * m["g11j"] * 16 * g22j
        + z0 * z1 * z2 * z2 / (16 * (z1 + z2) ** 3) * v["x2k_x3k"] * m["g11k"] * 16 * g22k
        - (z0 + z2) * z2 * z2 / (16 * (z1 + z2) ** 2) * v["x2j_x3j"] * m["g11k"] * 16 * g22j
        + (z1 + z2) * z2 * z2 / (16 * (z0 + z2) ** 2) * v["x2k_x3k"] * m["g11j"] * 16 * g22k
        + z0 * z2**3 / (4 * (z0 + z2) ** 4) * hj * m["g11j"]
        + z1 * z2**3 / (4 * (z1 + z2) ** 4) * hk * m["g11k"]
    )
    return jk + jkm + F + Fm


def _it_i3_from_moments(v, mf, z1, z2, m1, m2):
    """_IT_tripole_I3 with G values at y_t1 (m1) and y_t2 (m2) (S/dis/transverse.py:737-851)."""

    z0 = v["z0"]
    pj = (2 * z0 * (z0 + z2) + z2 * z2) / (z0 + z2) ** 2
    pk = (2 * z1 * (z1 + z2) + z2 * z2) / (z1 + z2) ** 2
    x2j2, x2k2, x3j2, x3k2 = v["x2_j"] ** 2, v["x2_k"] ** 2, v["x3_j"] ** 2, v["x3_k"] ** 2
    jkm = mf * mf * (pj * x2j2 / 64 * m1["b12j"] * m2["b12j"] + pk * x2k2 / 64 * m1["b12k"] * m2["b12k"])
    jk = pj * (1 - 2 * z1 * (1 - z1)) * x3j2 * x2j2 / 256 * m1["b22j"] * m2["b22j"] \
        + pk * (1 - 2 * z0 * (1 - z0)) * x3k2 * x2k2 / 256 * m1["b22k"] * m2["b22k"]
    F = 0.5 * m1["b22j"] * m2["b22k"] / (64 * (z0 + z2) * (z1 + z2)) * _cross(v, z1, z2)
    Fm = 0.5 * mf * mf * (
        z2**4 / (64 * (z0 + z2) ** 4) * (4 * z1 * (z1 - 1) + 2) * x3j2 * m1["g21j"] * m2["g21j"]
        - z0 * z1 * z2 * z2 / (16 * (z0 + z2) ** 3) * v["x2j_x3j"] * m1["b12j"] * m2["g21j"]
        + z2**4 / (64 * (z1 + z2) ** 4) * (4 * z0 * (z0 - 1) + 2) * x3k2 * m1["g21k"] * m2["g21k"]
        + z0 * z1 * z2 * z2 / (16 * (z1 + z2) ** 3) * v["x2k_x3k"] * m1["b12k"] * m2["g21k"]
        - ((2 * z0 + z2) * (2 * z1 + z2) + z2 * z2) * v["x20x21"] / (32 * (z0 + z2) * (z1 + z2)) * m1["b12j"] * m2["b12k"]
        + z2**4 / (32 * (z0 + z2) ** 2 * (z1 + z2) ** 2) * ((2 * z0 + z2) * (2 * z1 + z2) + z2 * z2) * v["x3j_x3k"] * m1["g21j"] * m2["g21k"]
        + mf * mf / 8 * (z2 / (z0 + z2)) ** 4 * m1["g11j"] * m2["g11j"]
        + mf * mf / 8 * (z2 / (z1 + z2)) ** 4 * m1["g11k"] * m2["g11k"]
        - (z0 * z2) ** 2 / (16 * (z0 + z2) * (z1 + z2) ** 2) * v["x2j_x3k"] * m1["b12j"] * m2["g21k"]
        + (z1 * z2) ** 2 / (16 * (z1 + z2) * (z0 + z2) ** 2) * v["x2k_x3j"] * m1["b12k"] * m2["g21j"]
        - z0 * z1 * z2 * z2 / (16 * (z0 + z2) ** 3) * v["x2j_x3j"] * m1["g11j"] * m2["b22j"]
        + z0 * z1 * z2 * z2 / (16 * (z1 + z2) ** 3) * v["x2k_x3k"] * m1["g11k"] * m2["b22k"]
        - (z0 + z2) * z2 * z2 / (16 * (z1 + z2) ** 2) * v["x2j_x3j"] * m1["g11k"] * m2["b22j"]
        + (z1 + z2) * z2 * z2 / (16 * (z0 + z2) ** 2) * v["x2k_x3k"] * m1["g11j"] * m2["b22k"]
    )
    return jk + jkm + F + Fm


def it_qg_i2(Q2, mf, z1, z2, x01sq, x02sq, x21sq, yt):
    """ITNLOqg_massive_tripole_part_I2 (S/dis/transverse.py:340-353)."""

    v = tripole_variables(math.sqrt(Q2), z1, z2, x01sq, x02sq, x21sq)
    return _it_i2_from_moments(v, mf, z1, z2, _t_moments(v, mf, yt))


def it_qg_i3(Q2, mf, z1, z2, x01sq, x02sq, x21sq, yt1, yt2):
    """ITNLOqg_massive_tripole_part_I3 (S/dis/transverse.py:356-368)."""

    v = tripole_variables(math.sqrt(Q2), z1, z2, x01sq, x02sq, x21sq)
    return _it_i3_from_moments(v, mf, z1, z2, _t_moments(v, mf, yt1), _t_moments(v, mf, yt2))


def il_qg_i1(Q2, mf, z1, z2, x01sq, x02sq, x21sq):
    """ILNLOqg_massive_tripole_part_I1 (S/dis/longitudinal.py:146-183)."""

    v = tripole_variables(math.sqrt(Q2), z1, z2, x01sq, x02sq, x21sq)
    z0 = v["z0"]
    bk = k0(np.sqrt(v["Qbar_j"] ** 2 + mf * mf) * np.sqrt(v["x3_j"] ** 2 + v["omega_j"] * x02sq))
    bl = k0(np.sqrt(v["Qbar_k"] ** 2 + mf * mf) * np.sqrt(v["x3_k"] ** 2 + v["omega_k"] * x21sq))
    tk = z1 * z1 * (2 * z0 * (z0 + z2) + z2 * z2) / x02sq * bk * bk
    tl = z0 * z0 * (2 * z1 * (z1 + z2) + z2 * z2) / x21sq * bl * bl
    tkl = -2 * z0 * z1 * (z0 * (1 - z0) + z1 * (1 - z1)) * v["x20x21"] / (x02sq * x21sq) * bk * bl
    return 4 * Q2 * (tk + tl + tkl)


def il_qg_uvsub(Q2, mf, z1, z2, x01sq, x02sq, x21sq):
    """ILNLOqg_massive_dipole_uvsub (S/dis/longitudinal.py:186-214)."""

    z0 = 1 - z1 - z2
    Q = math.sqrt(Q2)
    qk = Q * np.sqrt(z1 * (1 - z1))
    ql = Q * np.sqrt(z0 * (1 - z0))
    eg = math.exp(EULER_GAMMA)
    tk = -z1 * z1 * (2 * z0 * (z0 + z2) + z2 * z2) / x02sq * np.exp(-x02sq / x01sq / eg) * k0(np.sqrt(qk * qk + mf * mf) * np.sqrt(x01sq)) ** 2
    tl = -z0 * z0 * (2 * z1 * (z1 + z2) + z2 * z2) / x21sq * np.exp(-x21sq / x01sq / eg) * k0(np.sqrt(ql * ql + mf * mf) * np.sqrt(x01sq)) ** 2
    return 4 * Q2 * (tk + tl)


def _l_moments(v, mf, y):
    j = (v["Qbar_j"], mf, v["x2_j"], v["x3_j"], v["omega_j"])
    k = (v["Qbar_k"], mf, v["x2_k"], v["x3_k"], v["omega_k"])
    return dict(
        b12k=_g_bar(1, 2, *j, v["lambda_j"], y), b12l=_g_bar(1, 2, *k, v["lambda_k"], y),
        g11k=g_simplified(1, 1, *j, v["lambda_j"], y), g11l=g_simplified(1, 1, *k, v["lambda_k"], y),
    )


def _il_i2_from_moments(Q2, v, mf, z1, z2, m):
    """S/dis/longitudinal.py:217-255 with the y-dependent G values given."""

    z0 = v["z0"]
    x02sq, x21sq = v["x2_j"] ** 2, v["x2_k"] ** 2
    bk = k0(np.sqrt(v["Qbar_j"] ** 2 + mf * mf) * np.sqrt(v["x3_j"] ** 2 + v["omega_j"] * x02sq))
    bl = k0(np.sqrt(v["Qbar_k"] ** 2 + mf * mf) * np.sqrt(v["x3_k"] ** 2 + v["omega_k"] * x21sq))
    tk = z1 * z1 * (2 * z0 * (z0 + z2) + z2 * z2) / 4 * m["b12k"] * bk
    tl = z0 * z0 * (2 * z1 * (z1 + z2) + z2 * z2) / 4 * m["b12l"] * bl
    tkl = -z0 * z1 * (z0 * (1 - z0) + z1 * (1 - z1)) * v["x20x21"] / 4 * (m["b12k"] * bl / x21sq + m["b12l"] * bk / x02sq)
    return 4 * Q2 * (tk + tl + tkl)


def _il_i3_from_moments(Q2, v, mf, z1, z2, m1, m2):
    """S/dis/longitudinal.py:258-300 with the G values at y_t1 (m1), y_t2 (m2)."""

    z0 = v["z0"]
    x02sq, x21sq = v["x2_j"] ** 2, v["x2_k"] ** 2
    tk = z1 * z1 * (2 * z0 * (z0 + z2) + z2 * z2) * x02sq / 64 * m1["b12k"] * m2["b12k"]
    tl = z0 * z0 * (2 * z1 * (z1 + z2) + z2 * z2) * x21sq / 64 * m1["b12l"] * m2["b12l"]
    tkl = -z1 * z0 * (z1 * (1 - z1) + z0 * (1 - z0)) * v["x20x21"] / 32 * m1["b12k"] * m2["b12l"]
    tmf = mf * mf / 16 * z2**4 * (
        (z1 / (z0 + z2)) ** 2 * m1["g11k"] * m2["g11k"]
        + (z0 / (z1 + z2)) ** 2 * m1["g11l"] * m2["g11l"]
        - 2 * z0 / (z1 + z2) * z1 / (z0 + z2) * m1["g11k"] * m2["g11l"]
    )
    return 4 * Q2 * (tk + tl + tkl + tmf)


def il_qg_i2(Q2, mf, z1, z2, x01sq, x02sq, x21sq, yt):
    v = tripole_variables(math.sqrt(Q2), z1, z2, x01sq, x02sq, x21sq)
    return _il_i2_from_moments(Q2, v, mf, z1, z2, _l_moments(v, mf, yt))


def il_qg_i3(Q2, mf, z1, z2, x01sq, x02sq, x21sq, yt1, yt2):
    v = tripole_variables(math.sqrt(Q2), z1, z2, x01sq, x02sq, x21sq)
    return _il_i3_from_moments(Q2, v, mf, z1, z2, _l_moments(v, mf, yt1), _l_moments(v, mf, yt2))


def mapped_gl_rule(points):
    """Inner rule of the nested I2/I3 integrals: Gauss-Legendre in t on [0, 1],
    y = t^3, weight w/2 * 3 t^2 (K/dis/i2_i3/extension.py:21-36)."""

    x, w = np.polynomial.legendre.leggauss(points)
    t = (x + 1) / 2
    return t**3, w / 2 * 3 * t * t


def qg_nested(Q2, mf, z1, z2, x01sq, x02sq, x21sq, transverse, points=48):
    """(I2, I3) with the y_t integrals done on the mapped rule: I2 is linear and I3
    bilinear in the G values, so both follow from the integrated moments
    (K/dis/i2_i3/kernel.cu:768-912)."""

    v = tripole_variables(math.sqrt(Q2), z1, z2, x01sq, x02sq, x21sq)
    ys, ws = mapped_gl_rule(points)
    moments = _t_moments if transverse else _l_moments
    acc = None
    for y, w in zip(ys, ws):
        m = moments(v, mf, np.full_like(z1, y))
        if acc is None:
            acc = {k: w * val for k, val in m.items()}
        else:
            for k, val in m.items():
                acc[k] = acc[k] + w * val
    if transverse:
        return _it_i2_from_moments(v, mf, z1, z2, acc), _it_i3_from_moments(v, mf, z1, z2, acc, acc)
    return _il_i2_from_moments(Q2, v, mf, z1, z2, acc), _il_i3_from_moments(Q2, v, mf, z1, z2, acc, acc)


# ----------------------------------------------------------- full integrands


@dataclass(frozen=True)
class DisSetup:
    """The subset of DISConfig the integrands read (S/dis/config.py:101-145)."""

    maxr: float = 30.0
    finite_nc: bool = True
    parent_coupling: bool = False  # RunningCouplingScheme.PARENT
    smooth_coupling: bool = False  # RunningCouplingIRScheme.SMOOTH
    c2_alpha: float = 1.0
    nf_alphas: int = 4  # active flavours (light = 3, charm = 1 with the default quarks)
    max_alpha: float = 0.7
    mass_dependent_x: bool = False


def _alpha(r, s: DisSetup):
    if s.smooth_coupling:
        return alpha_smooth(r, s.nf_alphas, s.c2_alpha)
    return alpha_freeze(r, s.nf_alphas, s.c2_alpha, s.max_alpha)


def rapidity_dipole(xbj, Q2, mf, mass_dependent):
    """S/dis/observables.py:55-68."""

    W2 = (1 - xbj) * Q2 / xbj
    X = (Q2 + 4 * mf * mf) / (Q2 + W2) if mass_dependent else Q2 / (Q2 + W2)
    return math.log(1 / X)


def dipole_integrand(x, Q2, xbj, transverse, quark: QuarkSpec, contribution, amp, x0, s: DisSetup,
                     min_r=1e-30, max_r=1e3):
    """integrand_dip_massive (S/dis/observables.py:107-172) for unit-cube points x;
    amp(r, Y) is the dipole amplitude."""

    z1 = ZMIN + x[:, 0] * (1 - 2 * ZMIN)
    x01 = s.maxr * x[:, 1]
    inside = (x01 >= min_r) & (x01 <= max_r)
    x01e = np.clip(x01, min_r, max_r)
    mf = quark.mass
    Y = np.full_like(x01e, rapidity_dipole(xbj, Q2, mf, s.mass_dependent_x))
    N01 = amp(x01e, Y)
    with np.errstate(all="ignore"):
        if contribution == "Omega_L_const":
            res = N01 * il_dip_const(Q2, z1, x01e, mf)
        elif contribution == "ab":
            res = N01 * il_dip_ab(Q2, z1, x01e, mf, x[:, 2])
        elif contribution == "cd":
            res = N01 * il_dip_cd(Q2, z1, x01e, mf, x[:, 2], x[:, 3])
        elif contribution == "T0":
            res = N01 * it_dip_0(Q2, z1, x01e**2, mf)
        elif contribution == "T1":
            res = N01 * it_dip_1(Q2, z1, x01e**2, mf, x[:, 2])
        else:
            res = N01 * it_dip_2(Q2, z1, x01e**2, mf, x[:, 2], x[:, 3])
        res = res * (x01 * s.maxr * (1 - 2 * ZMIN)) * (_alpha(x01e, s) * CF / math.pi)
    return np.where(inside & np.isfinite(res), res, 0.0)


def qg_geometry(x, Q2, xbj, s: DisSetup):
    """Phase-space map of integrand_qgunsub_massive (S/dis/observables.py:186-200)."""

    z2min = Q0SQR / (Q2 / xbj)
    z1 = (1 - z2min) * x[:, 0]
    z2 = (1 - z1 - z2min) * x[:, 1] + z2min
    x01 = s.maxr * x[:, 2]
    x02 = s.maxr * x[:, 3]
    phi = 2 * math.pi * x[:, 4]
    x01sq, x02sq = x01 * x01, x02 * x02
    x21sq = np.maximum(x01sq + x02sq - 2 * x01 * x02 * np.cos(phi), 0)
    jac = (1 - z2min) * (1 - z1 - z2min) * x01 * x02 * s.maxr**2 * 2 * math.pi
    Y = np.log((Q2 / xbj) * z2 / Q0SQR)
    return z2min, z1, z2, x01, x02, x01sq, x02sq, x21sq, jac, Y


def tripole(amp, x01, x02, x21, Y, finite_nc):
    """S/dis/observables.py:86-104."""

    S01 = 1 - amp(x01, Y)
    S02 = 1 - amp(x02, Y)
    S12 = 1 - amp(x21, Y)
    if finite_nc:
        return 1 - NC / (2 * CF) * (S02 * S12 - S01 / NC**2)
    return 1 - S02 * S12


def qg_integrand(x, Q2, xbj, transverse, quark: QuarkSpec, contribution, amp, s: DisSetup, nested_points=None):
    """integrand_qgunsub_massive (S/dis/observables.py:175-250) or, with
    nested_points, integrand_qg_nested_massive (:253-331)."""

    z2min, z1, z2, x01, x02, x01sq, x02sq, x21sq, jac, Y = qg_geometry(x, Q2, xbj, s)
    if z2min > 1:
        return np.zeros(x.shape[0])
    x21 = np.sqrt(x21sq)
    Ye = np.maximum(Y, 0)
    mf = quark.mass
    with np.errstate(all="ignore"):
        N012 = tripole(amp, x01, x02, x21, Ye, s.finite_nc)
        r_a = x01 if s.parent_coupling else np.minimum(x01, np.minimum(x02, x21))
        af = _alpha(r_a, s) * CF / math.pi
        if contribution == "I1":
            N01 = amp(x01, Ye)
            if transverse:
                res = N01 * it_qg_uvsub(Q2, mf, z1, z2, x01sq, x02sq, x21sq) + N012 * it_qg_i1(Q2, mf, z1, z2, x01sq, x02sq, x21sq)
            else:
                res = N01 * il_qg_uvsub(Q2, mf, z1, z2, x01sq, x02sq, x21sq) + N012 * il_qg_i1(Q2, mf, z1, z2, x01sq, x02sq, x21sq)
        elif nested_points is not None:
            i2, i3 = qg_nested(Q2, mf, z1, z2, x01sq, x02sq, x21sq, transverse, nested_points)
            res = N012 * (i2 if contribution == "I2" else i3)
        elif contribution == "I2":
            f = it_qg_i2 if transverse else il_qg_i2
            res = N012 * f(Q2, mf, z1, z2, x01sq, x02sq, x21sq, x[:, 5])
        else:
            f = it_qg_i3 if transverse else il_qg_i3
            res = N012 * f(Q2, mf, z1, z2, x01sq, x02sq, x21sq, x[:, 5], x[:, 6])
        res = res * jac * af / z2
    return np.where((Y >= 0) & np.isfinite(res), res, 0.0)


def lo_integrand(x, Q2, xbj, transverse, quarks, amp, x0, s: DisSetup):
    """The LO cross-section integrand on the unit square: 2 pi r maxr |Psi|^2 N
    (S/dis/observables.py:345-352, S/dis/lo.py:12-51)."""

    z = x[:, 0]
    r = s.maxr * x[:, 1]
    tot = np.zeros_like(r)
    for q in quarks:
        tot = tot + lo_wavefunction(r, z, Q2, q, transverse)
    Y = np.full_like(r, max(math.log(x0 / xbj), 0.0))
    return 2 * math.pi * r * s.maxr * tot * amp(r, Y)


def gl_tensor(fn, dims, points):
    """Tensor Gauss-Legendre on the unit cube: (value at `points`, value at
    points // 2) (S/numerics/integration.py:95-117)."""

    def once(p):
        x, w = np.polynomial.legendre.leggauss(p)
        x = (x + 1) / 2
        w = w / 2
        grids = np.meshgrid(*([x] * dims), indexing="ij")
        wg = np.meshgrid(*([w] * dims), indexing="ij")
        pts = np.stack([g.reshape(-1) for g in grids], 1)
        ww = np.prod(np.stack([g.reshape(-1) for g in wg], 1), 1)
        return float(np.sum(ww * fn(pts)))

    return once(points), once(max(2, points // 2))
