"""Algorithmic operation count of one K2 + Kf sample of the FP32 integrator.

TEST AND MEASUREMENT INFRASTRUCTURE ONLY (like the rest of ``oracle/``): it is
the source of the flops-per-sample constant ``bench.py`` prices the dominant
kernel with, and ``tests/test_flop_count.py`` holds the two together.

The sample is written here as the mathematical formula the FP32 kernel
evaluates (``nlo_body`` / ``nlo_integrand_fp32`` in
``nlo_bk_optimized_amd/csrc/``), over a number type that counts operations:

* the VEGAS map of four Sobol coordinates through the per-bin absolute left
  end X_b and width W_b (bin and fraction from the integer product; the
  fraction's scaling is one multiplication), the Jacobian product;
* the cancellation-free K2 + Kf integrand (the algebra of
  ``orc_nlo_integrand_stable`` in ``nlo_oracle.c``, S/bk/derivatives.py:490-612,
  S/bk_kernels/k2.py:8-39, kf.py:10-41): distances from sin^2(theta/2),
  ln(X/Y) = log1p(rho (rho - 2 cos theta)) / 2, five log-log spline lookups
  N = 2^(cubic in the local coordinate), S = 1 - N, the dipole combinations,
  g(L) = L coth L - 1, h(L) = 2L / expm1(2L), h(-L) = h(L) e^{2L}, the ratios
  A, B, C, D, both kernels and their swapped copies, the |K2| cut, the smooth
  coupling at the smallest distance and the normalisation;
* the FP64 accumulation (w = f J, sum += w).

Convention: add, sub, mul, div, reciprocal, sqrt, reciprocal sqrt, exp, exp2,
log, log1p, expm1, sin, cos and coth count 1; a fused multiply-add counts 2;
compares, selects, abs, min, max, median-of-three, conversions, the table
index and all integer (Sobol, bin) work count 0.  A named function counts 1
however the kernel evaluates it (the FP32 kernel takes sin, log1p, expm1 and
g(L) from short series, ln and e^x from the hardware log2 / exp2): the count is
the formula's, and the executed instruction count beside it measures the
implementation.

Evaluated in float64 the restatement must equal ``orc_nlo_integrand_stable``
(the C oracle) to rounding: that pins every operand of the counted formula
(a swapped ratio or a dropped term fails the comparison).
"""

from __future__ import annotations

import math

import numpy as np

_COUNT = [0]


class Num:
    """A float64 value whose arithmetic is counted in ``_COUNT``."""

    __slots__ = ("v",)

    def __init__(self, v: float):
        self.v = float(v)

    @staticmethod
    def _val(x) -> float:
        return x.v if isinstance(x, Num) else float(x)

    def _op(self, other, f):
        _COUNT[0] += 1
        return Num(f(self.v, Num._val(other)))

    def __add__(self, o):
        return self._op(o, lambda a, b: a + b)

    def __radd__(self, o):
        return self._op(o, lambda a, b: b + a)

    def __sub__(self, o):
        return self._op(o, lambda a, b: a - b)

    def __rsub__(self, o):
        return self._op(o, lambda a, b: b - a)

    def __mul__(self, o):
        return self._op(o, lambda a, b: a * b)

    def __rmul__(self, o):
        return self._op(o, lambda a, b: b * a)

    def __truediv__(self, o):
        return self._op(o, lambda a, b: a / b)

    def __rtruediv__(self, o):
        return self._op(o, lambda a, b: b / a)

    def __neg__(self):  # sign flips are free
        return Num(-self.v)

    def __abs__(self):
        return Num(abs(self.v))

    def __lt__(self, o):
        return self.v < Num._val(o)

    def __gt__(self, o):
        return self.v > Num._val(o)

    def __le__(self, o):
        return self.v <= Num._val(o)

    def __ge__(self, o):
        return self.v >= Num._val(o)


def _fn(f):
    def g(x):
        _COUNT[0] += 1
        return Num(f(Num._val(x)))

    return g


exp = _fn(math.exp)
exp2 = _fn(lambda x: 2.0**x)
log = _fn(math.log)
log1p = _fn(math.log1p)
expm1 = _fn(math.expm1)
sin = _fn(math.sin)
rsqrt = _fn(lambda x: 1.0 / math.sqrt(x))
rcp = _fn(lambda x: 1.0 / x)


def fma(a, b, c):
    _COUNT[0] += 2
    return Num(Num._val(a) * Num._val(b) + Num._val(c))


def fract(u):  # u - floor(u)
    _COUNT[0] += 1
    return Num(Num._val(u) - math.floor(Num._val(u)))


def lcoth_minus_one(L):
    """g(L) = L coth L - 1 (coth, mul, sub: 3), its value without cancellation."""

    _COUNT[0] += 3
    a = abs(Num._val(L))
    if a < 1e-3:
        y = a * a
        v = y * (1 / 3 - y * (1 / 45 - y * 2 / 945))
    else:
        v = a / math.tanh(a) - 1.0
    return Num(v)


def _minimum(*xs):
    return min(xs, key=Num._val)


def _maximum(*xs):
    return max(xs, key=Num._val)


class LocalSpline:
    """The FP32 kernel's N table: per interval the cubic of log2 N in the local
    coordinate t in [0, 1), u = ln d / h + off clamped to [0, R] (records 0 and
    R are the constants N = 2^-126 below the grid and N = 1 above)."""

    def __init__(self, x: np.ndarray, coef: np.ndarray):
        n = x.size
        m = n - 1
        a, b, c, d = coef[0:m], coef[m:2 * m], coef[2 * m:3 * m], coef[3 * m:4 * m]
        h = (x[-1] - x[0]) / m
        l2e = 1.0 / math.log(2.0)
        rec = np.zeros((n + 1, 4))
        rec[0] = (-126.0, 0.0, 0.0, 0.0)
        rec[1:n] = np.stack([a * l2e, b * h * l2e, c * h * h * l2e, d * h ** 3 * l2e], axis=1)
        rec[n] = (0.0, 0.0, 0.0, 0.0)
        self.rec = rec
        self.inv_h = 1.0 / h
        self.off = 1.0 - x[0] / h
        self.umax = float(n)

    def log2n(self, log_d):
        u = fma(log_d, self.inv_h, self.off)
        u = Num(min(max(Num._val(u), 0.0), self.umax))
        t = fract(u)
        c = self.rec[min(int(u.v), self.rec.shape[0] - 1)]
        return fma(t, fma(t, fma(t, c[3], c[2]), c[1]), c[0])

    def n(self, log_d):
        return exp2(self.log2n(log_d))


def integrand(r: float, log_r: float, lz, lz2, tz, tz2, sp: LocalSpline, *, nf_over_nc: float = 1.0,
              nc: float = 3.0, norm: float = 1.0 / (8.0 * math.pi**4), b0: float = 0.0,
              log_mu0: float = 0.0, log_scale0: float = 0.0):
    """One sample of the cancellation-free K2 + Kf integrand (symmetrised,
    N-spline, smooth coupling at the smallest distance): (total, k2, kf)."""

    z, z2 = exp(lz), exp(lz2)
    s1, s2 = sin(0.5 * tz), sin(0.5 * tz2)
    s3 = sin(0.5 * (tz - tz2))
    s1q, s2q, s3q = s1 * s1, s2 * s2, s3 * s3
    xsq = fma(4.0 * r * z, s1q, (r - z) * (r - z))
    x2sq = fma(4.0 * r * z2, s2q, (r - z2) * (r - z2))
    zsq = fma(4.0 * z * z2, s3q, (z - z2) * (z - z2))
    if not (xsq.v >= 1e-37 and x2sq.v >= 1e-37 and zsq.v >= 1e-37):
        return Num(0.0), Num(0.0), Num(0.0)
    iX, iX2, iZ = rsqrt(xsq), rsqrt(x2sq), rsqrt(zsq)
    rho1, rho2 = r * rcp(z), r * rcp(z2)
    # ln(X / Y) = log1p(rho (rho - 2 cos t)) / 2 with cos t = 1 - 2 sin^2(t / 2)
    c1, c2 = 1.0 - 2.0 * s1q, 1.0 - 2.0 * s2q
    lxy = 0.5 * log1p(rho1 * (rho1 - 2.0 * c1))
    lx2y2 = 0.5 * log1p(rho2 * (rho2 - 2.0 * c2))
    log_x, log_x2, log_zz = lz + lxy, lz2 + lx2y2, 0.5 * log(zsq)
    n_x, n_y, n_x2, n_y2, n_z = (sp.n(v) for v in (log_x, lz, log_x2, lz2, log_zz))
    s_x, s_y, s_x2, s_y2 = (1.0 - v for v in (n_x, n_y, n_x2, n_y2))
    dip = s_x * (n_z + n_y2 - n_y - n_z * n_y2)
    dips = s_x2 * (n_z + n_y - n_y2 - n_z * n_y)
    L = lxy - lx2y2
    two_l = 2.0 * L
    h = two_l / expm1(two_l) if two_l.v != 0.0 else Num(1.0)
    hm = h * exp(two_l)
    g = lcoth_minus_one(L)
    A = z * z2 * (iZ * iZ)
    rZ = r * iZ
    B, C, D = rZ * z2 * iX2, r * r * (iX * iX2), rZ * z * iX
    B2, C2, D2 = B * B, C * C, D * D
    A2g = 2.0 * A * A * g
    k = A2g - 4.0 * h * B2 + h * C2 + 2.0 * L * D2
    ks = A2g - 4.0 * hm * D2 + hm * C2 - 2.0 * L * B2
    zz2 = z * z2
    cut = 1e10 * (zz2 * zz2)  # |K2| > 1e10 with the z^2 z'^2 Jacobian carried
    if abs(k.v) > cut.v and abs(dip.v) < 1e-10:
        k, dip = Num(0.0), Num(0.0)
    if abs(ks.v) > cut.v and abs(dips.v) < 1e-10:
        ks, dips = Num(0.0), Num(0.0)
    k2r = (k * dip + ks * dips) * 0.5
    kf = (h * B2 - A2g) * nf_over_nc
    kfs = (hm * D2 - A2g) * nf_over_nc
    df, dfs = s_y * (n_x - n_x2), s_y2 * (n_x2 - n_x)
    kfr = -((kf * df + kfs * dfs) * 0.5)
    log_min = _minimum(Num(log_r), log_x, lz, log_x2, lz2, log_zz)
    right = 5.0 * (log_scale0 - 2.0 * log_min)
    hi = _maximum(Num(log_mu0), right)
    t = exp(-abs(log_mu0 - right))
    as_ = rcp(b0 * (0.2 * (hi + log1p(t))))
    asn = as_ * nc
    scale = asn * asn * norm
    k2r, kfr = k2r * scale, kfr * scale
    return k2r + kfr, k2r, kfr


def vegas_map(q: np.ndarray, bins: int, X: np.ndarray, W: np.ndarray):
    """x_d = X_b + W_b frac and J = prod W_b (the widths carry B and the volume
    except for the constant B^4 factor, one more multiplication)."""

    x, jac = [], None
    for d in range(4):
        q4 = (int(q[d]) << 2) & 0xFFFFFFFF
        b = (q4 * bins) >> 32
        frac = Num((q4 * bins) & 0xFFFFFFFF) * 2.0**-32
        x.append(fma(W[d, b], frac, X[d, b]))
        jac = Num(W[d, b]) if jac is None else jac * W[d, b]
    return x, float(bins) ** 4 * jac


def sample_ops(r: float, log_r: float, q: np.ndarray, bins: int, X: np.ndarray, W: np.ndarray,
               sp: LocalSpline, **physics) -> tuple[int, float]:
    """Counted operations of one sample (map, integrand, accumulation) and w."""

    _COUNT[0] = 0
    x, jac = vegas_map(q, bins, X, W)
    total, _, _ = integrand(r, log_r, *x, sp, **physics)
    w = total * jac
    acc = Num(0.0) + w  # sum += w
    return _COUNT[0], acc.v


def flops_per_sample(r: np.ndarray, N: np.ndarray, parents, q: np.ndarray, bins: int = 50, **physics) -> float:
    """Mean counted operations per sample over the given parents and Sobol
    states q [n, 4] (uniform grids on the kernel's box)."""

    from oracle import oracle as O

    sN, _ = O.evolution_splines(r, N)
    sp = LocalSpline(sN.x, sN.coef)
    lo_r = math.log(0.5 * r[0])
    w_r = math.log(2.0 * r[-1]) - lo_r
    e = np.linspace(0.0, 1.0, bins + 1)
    X = np.empty((4, bins))
    W = np.empty((4, bins))
    for d in range(4):
        lo_d, w_d = (lo_r, w_r) if d < 2 else (0.0, 2.0 * math.pi)
        X[d] = lo_d + w_d * e[:-1]
        W[d] = w_d * (e[1:] - e[:-1])
    total, count = 0, 0
    for p in parents:
        for row in q:
            ops, _ = sample_ops(float(r[p]), math.log(r[p]), row, bins, X, W, sp, **physics)
            total += ops
            count += 1
    return total / count
