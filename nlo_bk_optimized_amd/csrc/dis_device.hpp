// Device-side NLO DIS physics: modified Bessel K0/K1/K2, the dilogarithm, the
// DIS couplings, the GBW and BK-table dipoles, and every impact factor of the
// LO, dipole-sector and qqbar-g contributions, templated on the arithmetic
// type (double for the FP64 path, float for the FP32 path).  Reference
// prefixes: S/ = torch_nlo_clean/src/nlo_torch/, K/ = S/custom_kernels/.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "nlo_gfx950.h"

namespace nlo {
namespace dis {

#include "bessel_tables.inc"

constexpr double kPiD = 3.14159265358979323846;
constexpr double kEulerGamma = 0.57721566490153286061;

// ------------------------------------------------------------ special functions

template <typename T>
struct K01 {
    T k0, k1;
};

// K0(x) and K1(x) for x >= 0 (S/numerics/special.py:18-23 calls ATen's).
// x <= 2: the ascending series (A&S 9.6.11/9.6.13) with t = x^2/4,
//   K0 = -(ln(x/2) + gamma) I0 + sum H_k t^k / k!^2,
//   K1 = 1/x + (x/2)(ln(x/2) + gamma) sum t^k/(k!(k+1)!) - (x/4) sum (H_k + H_{k+1}) t^k/(k!(k+1)!),
// the four sums as polynomials in t (15 terms FP64, 8 FP32) by Horner's rule;
// x > 2: K = e^-x / sqrt(x) * sum c_m T_m(4/x - 1) (Chebyshev fit of the
// scaled function, coefficients from gen_bessel_tables.py; Clenshaw in FP64,
// the power-basis form by Horner in FP32).  FP32 takes 1/x, 1/sqrt(x) and ln
// from the hardware approximations; FP64 keeps IEEE divisions.
// Reciprocal, reciprocal square root and natural log: IEEE operations in
// FP64; in FP32 the hardware approximations (about 1 ulp; the log with an
// absolute error ~1e-7), where no FP32 result is a small difference of them.
template <typename T>
__device__ __forceinline__ T dis_rcp(T x) {
    if constexpr (sizeof(T) == 4)
        return __builtin_amdgcn_rcpf(x);
    else
        return T(1) / x;
}
template <typename T>
__device__ __forceinline__ T dis_rsqrt(T x) {
    if constexpr (sizeof(T) == 4)
        return __builtin_amdgcn_rsqf(x);
    else
        return T(1) / sqrt(x);
}
// Square root: IEEE in FP64; in FP32 the hardware one (about 1 ulp, without
// the correctly rounded expansion's denormal scaling: every FP32 argument here
// is a product of squared physical scales, zero or far above 1e-38).
template <typename T>
__device__ __forceinline__ T dis_sqrt(T x) {
    if constexpr (sizeof(T) == 4)
        return __builtin_amdgcn_sqrtf(x);
    else
        return sqrt(x);
}

template <typename T>
__device__ __forceinline__ K01<T> bessel_k01(T x) {
    constexpr bool f64 = sizeof(T) == 8;
    if (!(x > T(0))) {
        const T v = x == T(0) ? T(INFINITY) : T(NAN);
        return {v, v};
    }
    if (x <= T(2)) {
        const T t = T(0.25) * x * x;
        T i0, s0, i1, s1;
        if constexpr (f64) {
            // the four series as polynomials in t by Horner's rule, the loop
            // kept rolled: its 60 double coefficients are read per use (scalar
            // loads) instead of occupying registers across the integrand
            constexpr int n = kBesselSeriesTerms64;
            i0 = kBesselSeriesI0_64[n - 1], s0 = kBesselSeriesS0_64[n - 1];
            i1 = kBesselSeriesI1_64[n - 1], s1 = kBesselSeriesS1_64[n - 1];
#pragma unroll 1
            for (int k = n - 2; k >= 0; --k) {
                i0 = fma(i0, t, kBesselSeriesI0_64[k]);
                s0 = fma(s0, t, kBesselSeriesS0_64[k]);
                i1 = fma(i1, t, kBesselSeriesI1_64[k]);
                s1 = fma(s1, t, kBesselSeriesS1_64[k]);
            }
            const T lg = log(T(0.5) * x) + T(kEulerGamma);
            return {s0 - lg * i0, T(1) / x + T(0.5) * x * lg * i1 - T(0.25) * x * s1};
        } else {
            // the four series as polynomials in t (Horner, exact rational
            // coefficients rounded once; every term is positive)
            constexpr int n = kBesselSeriesTerms32;
            i0 = kBesselSeriesI0_32[n - 1], s0 = kBesselSeriesS0_32[n - 1];
            i1 = kBesselSeriesI1_32[n - 1], s1 = kBesselSeriesS1_32[n - 1];
#pragma unroll
            for (int k = n - 2; k >= 0; --k) {
                i0 = fmaf(i0, t, kBesselSeriesI0_32[k]);
                s0 = fmaf(s0, t, kBesselSeriesS0_32[k]);
                i1 = fmaf(i1, t, kBesselSeriesI1_32[k]);
                s1 = fmaf(s1, t, kBesselSeriesS1_32[k]);
            }
            const T lg = __builtin_amdgcn_logf(T(0.5) * x) * T(0.69314718055994531) + T(kEulerGamma);
            return {s0 - lg * i0, dis_rcp(x) + T(0.5) * x * lg * i1 - T(0.25) * x * s1};
        }
    }
    const T s = (f64 ? T(4) / x : T(4) * dis_rcp(x)) - T(1);
    if constexpr (f64) {
        const T s2 = s + s;
        T b0k0 = T(0), b1k0 = T(0), b0k1 = T(0), b1k1 = T(0);
#pragma unroll
        for (int m = kBesselK0Deg64; m >= 1; --m) {
            const T t0 = s2 * b0k0 - b1k0 + T(kBesselK0Cheb64[m]);
            b1k0 = b0k0;
            b0k0 = t0;
            const T t1 = s2 * b0k1 - b1k1 + T(kBesselK1Cheb64[m]);
            b1k1 = b0k1;
            b0k1 = t1;
        }
        const T g0 = s * b0k0 - b1k0 + T(kBesselK0Cheb64[0]);
        const T g1 = s * b0k1 - b1k1 + T(kBesselK1Cheb64[0]);
        const T sc = exp(-x) / dis_sqrt(x);
        return {g0 * sc, g1 * sc};
    } else {
        // the degree-8 interpolant in the power basis of s (Horner)
        T g0 = kBesselK0Mono32[kBesselK0Deg32], g1 = kBesselK1Mono32[kBesselK1Deg32];
#pragma unroll
        for (int m = kBesselK0Deg32 - 1; m >= 0; --m) {
            g0 = fmaf(g0, s, kBesselK0Mono32[m]);
            g1 = fmaf(g1, s, kBesselK1Mono32[m]);
        }
        const T sc = exp(-x) * dis_rsqrt(x);
        return {g0 * sc, g1 * sc};
    }
}

template <typename T>
__device__ __forceinline__ T bessel_k0(T x) {
    return bessel_k01(x).k0;
}
template <typename T>
__device__ __forceinline__ T bessel_k1(T x) {
    return bessel_k01(x).k1;
}
// K2 = K0 + 2 K1 / x, +inf at 0 (S/numerics/special.py:26-32).
template <typename T>
__device__ __forceinline__ T bessel_k2(T x) {
    if (x == T(0)) return T(INFINITY);
    const K01<T> k = bessel_k01(x);
    return k.k0 + T(2) * k.k1 / x;
}

// Li2 on [-1, 1/2]: u = -ln(1 - x), u - u^2/4 + sum B_2k u^(2k+1)/(2k+1)!.
template <typename T>
__device__ __forceinline__ T dilog_core(T x) {
    const T u = -log1p(-x);
    const T u2 = u * u;
    T acc = T(0);
#pragma unroll
    for (int k = kDilogTerms - 1; k >= 0; --k) acc = acc * u2 + T(kDilogBernoulli[k]);
    return u - T(0.25) * u2 + u * u2 * acc;
}

// Real dilogarithm for x <= 1 (S/numerics/special.py:35-97): reflection
// Li2(x) = pi^2/6 - ln x ln(1-x) - Li2(1-x) above 1/2 and inversion
// Li2(x) = -pi^2/6 - ln^2(-x)/2 - Li2(1/x) below -1.
template <typename T>
__device__ __forceinline__ T dilog(T x) {
    const T pi2_6 = T(kPiD * kPiD / 6.0);
    if (x == T(1)) return pi2_6;
    if (x > T(0.5)) return pi2_6 - log(x) * log1p(-x) - dilog_core(T(1) - x);
    if (x < T(-1)) {
        const T l = log(-x);
        return -pi2_6 - T(0.5) * l * l - dilog_core(T(1) / x);
    }
    return dilog_core(x);
}

// --------------------------------------------------------------- couplings

struct DisCoupling {
    double b0;         // (11 Nc - 2 Nf) / (12 pi)
    double scale_num;  // 4 C2 / Lambda^2
    double max_alpha;
    int smooth;
};

// S/coupling.py:49-66 (freeze) and :69-97 (smooth).
template <typename T>
__device__ __forceinline__ T dis_alpha(T r, const DisCoupling& c) {
    const T arg = T(c.scale_num) / (r * r);
    if (c.smooth) {
        const T mu0 = T(2.0 / 0.2 * 0.91629073187415506518);  // 10 ln 2.5
        const T ls = log(arg) / T(0.2);
        const T hi = fmax(mu0, ls);
        return T(1) / (T(c.b0) * T(0.2) * (hi + log1p(exp(-fabs(mu0 - ls)))));
    }
    const T a = T(1) / (T(c.b0) * log(arg));
    return (a > T(c.max_alpha) || arg < T(1)) ? T(c.max_alpha) : a;
}

// ------------------------------------------------------------------ dipoles

// Device view of the dipole amplitude N(r, Y).
struct DisDipole {
    int kind;  // NLO_DIPOLE_GBW or NLO_DIPOLE_BK_TABLE
    // GBW (S/dipole/amplitude.py:22-46)
    double qs0sqr, lambda, gamma, y0_gbw;  // y0_gbw = ln(1 / x0)
    // BK table (S/dipole/amplitude.py:49-270): r[R] log-uniform, Y[T] (shifted
    // by max(0, ln(1/x0))), N[T x R]
    int R, T;
    const double* r;
    const double* Y;
    const double* N;
    // the spline initialised at one rapidity (natural spline in r over all
    // table nodes): y = a + d (b + d (c + d e)), 4 (R - 1) doubles; used where
    // Y matches y_init as BKDipole.interpolator_initialized decides
    int initialized;
    double y_init, maxr_interpolate;
    const double* init_coef;
    int bilinear;  // AmplitudeInterpolation.LINEAR_LINEAR
    // SPLINE_LINEAR: per row y and interval i the cubic in t = r - r_i of the
    // local natural spline on interval i (T (R - 1) x 4), or null (solve per query)
    const double* local_coef;
};

// GBW: 1 - exp(-(r^2 Qs^2(Y))^gamma / 4) (as arg below 1e-7), Qs^2 =
// Qs0^2 e^(lambda max(Y, ln(1/x0))).  1 - e^-a is formed as -expm1(-a).
template <typename T>
__device__ __forceinline__ T gbw_amplitude(T r, T Y, const DisDipole& d) {
    const T ye = fmax(Y, T(d.y0_gbw));
    const T qs2 = T(d.qs0sqr) * exp(T(d.lambda) * ye);
    const T a = (d.gamma == 1.0 ? r * r * qs2 : pow(r * r * qs2, T(d.gamma))) * T(0.25);
    return fabs(a) < T(1e-7) ? a : -expm1(-a);
}

// Natural cubic spline through n <= 9 points (x_i, y_i), evaluated at xq
// (S/numerics/interpolation.py NaturalCubicSpline, one spline per query).
template <typename T>
__device__ __forceinline__ T local_natural_spline(const T* x, const T* y, int n, T xq) {
    T m[9], cp[9], dp[9];
    // second derivatives: m_0 = m_{n-1} = 0, tridiagonal system for the interior
    const int ni = n - 2;
    for (int i = 0; i < ni; ++i) {
        const T h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
        const T diag = T(2) * (h0 + h1);
        const T rhs = T(6) * ((y[i + 2] - y[i + 1]) / h1 - (y[i + 1] - y[i]) / h0);
        const T lower = i > 0 ? h0 : T(0);
        const T den = diag - lower * (i > 0 ? cp[i - 1] : T(0));
        cp[i] = h1 / den;
        dp[i] = (rhs - lower * (i > 0 ? dp[i - 1] : T(0))) / den;
    }
    m[0] = T(0);
    m[n - 1] = T(0);
    for (int i = ni - 1; i >= 0; --i) m[i + 1] = dp[i] - (i < ni - 1 ? cp[i] * m[i + 2] : T(0));
    int k = 0;
    while (k < n - 2 && xq >= x[k + 1]) ++k;
    const T h = x[k + 1] - x[k];
    const T a = x[k + 1] - xq, b = xq - x[k];
    return (m[k] * a * a * a + m[k + 1] * b * b * b) / (T(6) * h) + (y[k] / h - m[k] * h / T(6)) * a +
           (y[k + 1] / h - m[k + 1] * h / T(6)) * b;
}

// The same spline's piece on interval k as a cubic in t = x - x[k]:
// c0 + t (c1 + t (c2 + t c3)) with c0 = y_k, c1 = (y_{k+1} - y_k) / h -
// h (2 m_k + m_{k+1}) / 6, c2 = m_k / 2, c3 = (m_{k+1} - m_k) / (6 h).
__device__ __forceinline__ void local_natural_piece(const double* x, const double* y, int n, int k, double* c) {
    double m[9], cp[9], dp[9];
    const int ni = n - 2;
    for (int i = 0; i < ni; ++i) {
        const double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
        const double diag = 2.0 * (h0 + h1);
        const double rhs = 6.0 * ((y[i + 2] - y[i + 1]) / h1 - (y[i + 1] - y[i]) / h0);
        const double lower = i > 0 ? h0 : 0.0;
        const double den = diag - lower * (i > 0 ? cp[i - 1] : 0.0);
        cp[i] = h1 / den;
        dp[i] = (rhs - lower * (i > 0 ? dp[i - 1] : 0.0)) / den;
    }
    m[0] = 0.0;
    m[n - 1] = 0.0;
    for (int i = ni - 1; i >= 0; --i) m[i + 1] = dp[i] - (i < ni - 1 ? cp[i] * m[i + 2] : 0.0);
    const double h = x[k + 1] - x[k];
    c[0] = y[k];
    c[1] = (y[k + 1] - y[k]) / h - h * (2.0 * m[k] + m[k + 1]) / 6.0;
    c[2] = m[k] / 2.0;
    c[3] = (m[k + 1] - m[k]) / (6.0 * h);
}

// The window of BKDipole._local_spline_amplitude for r in interval rind:
// 7 nodes from rind - 3 (clamped at 0), or the last 9 at the top of the grid.
__device__ __forceinline__ void local_window(int rind, int R, int& start, int& pts) {
    if (rind + 3 > R - 1) {
        start = R - 9;
        pts = 9;
    } else {
        start = rind - 3 < 0 ? 0 : rind - 3;
        pts = 7;
    }
}

// Largest i with v[i] <= q (searchsorted(..., right=True) - 1), -1 below v[0].
__device__ __forceinline__ int search_right(const double* v, int n, double q) {
    int lo = 0, hi = n;  // first index with v > q lies in [lo, hi]
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] <= q) lo = mid + 1;
        else hi = mid;
    }
    return lo - 1;
}

// BKDipole.interpolator_initialized (S/dipole/amplitude.py:95-103).
__device__ __forceinline__ bool bk_initialized(double Y, const DisDipole& d) {
    if (!d.initialized) return false;
    const bool near_zero = fabs(Y) < 0.001 && fabs(d.y_init) < 0.001;
    double den = fmin(Y, d.y_init);
    if (den == 0.0) den = 1.0;
    const bool match = d.y_init >= 0.0 && fabs(Y - d.y_init) / den < 0.001;
    return near_zero || match;
}

// BKDipole.dipole_amplitude (S/dipole/amplitude.py:133-168): r clamped into
// the table, Y frozen at the first row from below; the initialised spline where
// Y matches, else the local 7-point (9 at the top) natural spline of the
// Y-linear table rows (:184-240); clamped to [0, 1].  Evaluated in FP64 (the
// table and its interpolation are shared by both arithmetic paths).
// SOLVE = false (the integrands): the local-spline table is required (the
// host checks it), so no per-query solve is compiled in.
template <bool SOLVE = true>
__device__ __forceinline__ double bk_amplitude(double r, double Y, const DisDipole& d) {
    const double rmin = d.r[0], rmax = d.r[d.R - 1];
    if (r < rmin) r = rmin * 1.000001;
    if (r > rmax) r = rmax * 0.999999;
    if (Y >= 0.0 && Y < d.Y[0]) Y = d.Y[0];
    double v;
    if (bk_initialized(Y, d)) {
        int k = search_right(d.r, d.R, r);
        k = k < 0 ? 0 : (k > d.R - 2 ? d.R - 2 : k);
        const double t = r - d.r[k];
        const double* c = d.init_coef + 4 * k;
        v = c[0] + t * (c[1] + t * (c[2] + t * c[3]));
        if (d.maxr_interpolate > 0.0 && r >= d.maxr_interpolate) v = 1.0;
    } else if (d.bilinear) {
        // _bilinear_amplitude (S/dipole/amplitude.py:176-202)
        int rind = search_right(d.r, d.R, r);
        int yind = search_right(d.Y, d.T, Y);
        rind = rind < 0 ? 0 : rind;
        yind = yind < 0 ? 0 : yind;
        const int rind2 = rind + 1 < d.R ? rind + 1 : d.R - 1;
        const int yind2 = yind + 1 < d.T ? yind + 1 : d.T - 1;
        const double r1 = d.r[rind], r2 = d.r[rind2], Y1 = d.Y[yind], Y2 = d.Y[yind2];
        const double dr = rind2 != rind ? r2 - r1 : 1.0, dY = yind2 != yind ? Y2 - Y1 : 1.0;
        const double N11 = d.N[(size_t)yind * d.R + rind], N12 = d.N[(size_t)yind * d.R + rind2];
        const double N21 = d.N[(size_t)yind2 * d.R + rind], N22 = d.N[(size_t)yind2 * d.R + rind2];
        if (rind2 == rind) {
            v = 1.0;
        } else if (yind2 == yind) {
            v = N11 + (r - r1) / dr * (N12 - N11);
        } else {
            v = (N11 * (r2 - r) * (Y2 - Y) + N12 * (r - r1) * (Y2 - Y) + N21 * (r2 - r) * (Y - Y1) +
                 N22 * (r - r1) * (Y - Y1)) / (dr * dY);
        }
    } else if (!SOLVE || d.local_coef) {
        // the local splines are linear in their data: the Y-linear
        // combination of the two rows' precomputed cubics on interval rind
        int rind = search_right(d.r, d.R, r);
        rind = rind < 0 ? 0 : (rind > d.R - 2 ? d.R - 2 : rind);
        const int yind = search_right(d.Y, d.T, Y);
        const int y1 = yind < 0 ? 0 : yind;
        const int y2 = y1 + 1 < d.T ? y1 + 1 : d.T - 1;
        const double frac = y2 != y1 ? (Y - d.Y[y1]) / (d.Y[y2] - d.Y[y1]) : 0.0;
        const double t = r - d.r[rind];
        const double* c1 = d.local_coef + 4 * ((size_t)y1 * (d.R - 1) + rind);
        const double* c2 = d.local_coef + 4 * ((size_t)y2 * (d.R - 1) + rind);
        const double lo = c1[0] + t * (c1[1] + t * (c1[2] + t * c1[3]));
        const double up = c2[0] + t * (c2[1] + t * (c2[2] + t * c2[3]));
        v = lo + frac * (up - lo);
    } else if constexpr (SOLVE) {
        const int rind = search_right(d.r, d.R, r);
        const int yind = search_right(d.Y, d.T, Y);
        const int y1 = yind < 0 ? 0 : yind;
        const int y2 = y1 + 1 < d.T ? y1 + 1 : d.T - 1;
        int start, pts;
        local_window(rind, d.R, start, pts);
        const double frac = y2 != y1 ? (Y - d.Y[y1]) / (d.Y[y2] - d.Y[y1]) : 0.0;
        double xs[9], ys[9];
        for (int i = 0; i < pts; ++i) {
            const double lo = d.N[(size_t)y1 * d.R + start + i], up = d.N[(size_t)y2 * d.R + start + i];
            xs[i] = d.r[start + i];
            ys[i] = lo + frac * (up - lo);
        }
        v = local_natural_spline<double>(xs, ys, pts, r);
    }
    return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
}

template <typename T, bool SOLVE = true>
__device__ __forceinline__ T dipole_amplitude(T r, T Y, const DisDipole& d) {
    if (d.kind == NLO_DIPOLE_GBW) return gbw_amplitude<T>(r, Y, d);
    return (T)bk_amplitude<SOLVE>((double)r, (double)Y, d);
}

// ----------------------------------------------------------- LO (a19)

// |Psi_{T,L}|^2 of one quark without the dipole (S/dis/lo.py:25-47).
template <typename T>
__device__ __forceinline__ T lo_wavefunction(T r, T z, T Q2, T mf, T charge, bool transverse) {
    const T eps = dis_sqrt(Q2 * z * (T(1) - z) + mf * mf);
    const T a = r * eps;
    if (!(a >= T(1e-7) && a <= T(5e2))) return T(0);
    const K01<T> k = bessel_k01(a);
    if (transverse)
        return charge * charge *
               ((T(1) - T(2) * z + T(2) * z * z) * eps * eps * k.k1 * k.k1 + mf * mf * k.k0 * k.k0);
    return charge * charge * T(4) * Q2 * z * z * (T(1) - z) * (T(1) - z) * k.k0 * k.k0;
}

// ---------------------------------------------- dipole-sector terms (a20)

template <typename T>
__device__ __forceinline__ T omega_t_v_u(T Q, T z, T mf) {
    const T g = dis_sqrt(T(1) + T(4) * (mf / Q) * (mf / Q));
    return (T(1) + T(1) / (T(2) * z)) * (log(T(1) - z) + g * log((T(1) + g) / (T(1) + g - T(2) * z))) -
           T(1) / (T(2) * z) * ((z + T(0.5)) * (T(1) - g) + mf * mf / (Q * Q)) *
               log((z * (T(1) - z) * Q * Q + mf * mf) / (mf * mf));
}

template <typename T>
__device__ __forceinline__ T omega_t_n_u(T Q, T z, T mf) {
    const T g = dis_sqrt(T(1) + T(4) * (mf / Q) * (mf / Q));
    return (T(1) + z - T(2) * z * z) / z * (log(T(1) - z) + g * log((T(1) + g) / (T(1) + g - T(2) * z))) +
           (T(1) - z) / z * ((T(0.5) + z) * (g - T(1)) - mf * mf / (Q * Q)) *
               log((z * (T(1) - z) * Q * Q + mf * mf) / (mf * mf));
}

// L_dip (S/dis/longitudinal.py:133-143).
template <typename T>
__device__ __forceinline__ T l_dip(T Q2, T z, T mf) {
    const T g = dis_sqrt(T(1) + T(4) * mf * mf / Q2);
    const T s = dilog(T(1) / (T(1) - (T(1) - g) / (T(2) * z))) + dilog(T(1) / (T(1) - (T(1) + g) / (T(2) * z))) +
                dilog(T(1) / (T(1) - (T(1) - g) / (T(2) * (T(1) - z)))) +
                dilog(T(1) / (T(1) - (T(1) + g) / (T(2) * (T(1) - z))));
    const T lz = log(z / (T(1) - z));
    return s - (T(kPiD * kPiD / 6.0) - T(0.5) * lz * lz);
}

// OmegaL_V (S/dis/longitudinal.py:118-130).
template <typename T>
__device__ __forceinline__ T omega_l_v(T Q2, T z, T mf) {
    const T g = dis_sqrt(T(1) + T(4) * mf * mf / Q2);
    return T(1) / (T(2) * z) * (log(T(1) - z) + g * log((T(1) + g) / (T(1) + g - T(2) * z))) +
           T(1) / (T(2) * (T(1) - z)) * (log(z) + g * log((T(1) + g) / (T(1) + g - T(2) * (T(1) - z)))) +
           T(1) / (T(4) * z * (T(1) - z)) * (g - T(1) + T(2) * mf * mf / Q2) *
               log((z * (T(1) - z) * Q2 + mf * mf) / (mf * mf));
}

// ITdip_massive_0 (S/dis/transverse.py:27-46).
template <typename T>
__device__ __forceinline__ T it_dip_0(T Q2, T z1, T x01, T mf) {
    const T Q = dis_sqrt(Q2);
    const T kz = dis_sqrt(z1 * (T(1) - z1) * Q2 + mf * mf);
    const T lz = log(z1 / (T(1) - z1));
    const T common = T(-kPiD * kPiD / 3.0 + kPiD * kPiD / 6.0) + T(0.5) * lz * lz + omega_t_v_u(Q, z1, mf) +
                     omega_t_v_u(Q, T(1) - z1, mf) + l_dip(Q2, z1, mf);
    const K01<T> k = bessel_k01(x01 * kz);
    const T on = omega_t_n_u(Q, z1, mf) - omega_t_n_u(Q, T(1) - z1, mf);
    const T t1 = (kz * k.k1) * (kz * k.k1) *
                 ((z1 * z1 + (T(1) - z1) * (T(1) - z1)) * (T(2.5) + common) + (T(2) * z1 - T(1)) / T(2) * on);
    const T t2 = (mf * k.k0) * (mf * k.k0) * (T(3) + common);
    return t1 + t2;
}

template <typename T>
__device__ __forceinline__ void it_v1_vms1_u(T Q, T z, T mf, T r, T xi, T& v1, T& vms1) {
    const T kz = dis_sqrt(z * (T(1) - z) * Q * Q + mf * mf);
    const T dk = xi / (T(1) - xi) * (T(1) - z) * mf * mf;
    const T sh = dis_sqrt(kz * kz + dk);
    const K01<T> ks = bessel_k01(r * sh), kb = bessel_k01(r * kz);
    const T pre = T(1) / xi * (T(2) * log(xi) / (T(1) - xi) - (T(1) + xi) / T(2));
    v1 = pre * (sh * ks.k1 - kz * kb.k1) -
         (log(xi) / ((T(1) - xi) * (T(1) - xi)) + z / (T(1) - xi) + z / T(2)) * (T(1) - z) * mf * mf / sh * ks.k1;
    vms1 = pre * (ks.k0 - kb.k0) + (T(-1.5) * (T(1) - z) / (T(1) - xi) + (T(1) - z) / T(2)) * ks.k0;
}

// ITdip_massive_1 (S/dis/transverse.py:79-148).
template <typename T>
__device__ __forceinline__ T it_dip_1(T Q2, T z1, T x01, T mf, T xi) {
    const T Q = dis_sqrt(Q2);
    const T kz = dis_sqrt(z1 * (T(1) - z1) * Q2 + mf * mf);
    T a1, a2, b1, b2;
    it_v1_vms1_u(Q, z1, mf, x01, xi, a1, b1);
    it_v1_vms1_u(Q, T(1) - z1, mf, x01, xi, a2, b2);
    const K01<T> k = bessel_k01(x01 * kz);
    return kz * k.k1 * (z1 * z1 + (T(1) - z1) * (T(1) - z1)) * (a1 + a2) + mf * mf * k.k0 * (b1 + b2);
}

// _weighted_k1_difference_stable (S/dis/transverse.py:13-24).
template <typename T>
__device__ __forceinline__ T k1_difference_stable(T r, T kb, T dk2, const K01<T>& kbase) {
    const T b2 = kb * kb;
    const T s2 = b2 + dk2;
    if (!(s2 > T(0) && kb > T(0))) return T(0);
    const T rel = fabs(dk2) / (b2 + fabs(dk2) + T(1e-300));
    if (rel < T(1e-5)) return T(-0.5) * r * dk2 * kbase.k0;
    const T sh = dis_sqrt(fmax(s2, T(0)));
    return sh * bessel_k1(r * sh) - kb * kbase.k1;
}

// IT_V2 / IT_N / IT_VMS2 unsymmetric parts (S/dis/transverse.py:170-306).
template <typename T>
__device__ __forceinline__ void it_v2_n_vms2_u(T Q, T z, T mf, T r, T ych, T yu, T& v2, T& n, T& vms2) {
    const T chi = z * ych;
    const T u = (T(1) - yu) / yu;
    const T kz = dis_sqrt(z * (T(1) - z) * Q * Q + mf * mf);
    const T kc = dis_sqrt(chi * (T(1) - chi) * Q * Q + mf * mf);
    const T dk = u * (T(1) - z) / (T(1) - chi) * kc * kc;
    const T sh = dis_sqrt(kz * kz + dk);
    const K01<T> kb = bessel_k01(r * kz), ks = bessel_k01(r * sh);
    const T m2 = mf * mf, kc2 = kc * kc;
    const T up1 = u + T(1), uu = u / up1;
    const T pre = z / (yu * yu);
    // V2
    const T kd = k1_difference_stable(r, kz, dk, kb);
    const T t1 = T(-1) / (T(1) - chi) / (u * up1) * m2 / kc2 * (T(2) * chi + uu * uu / z * (z - chi) * (T(1) - T(2) * chi)) * kd;
    const T t2 = T(-1) / ((T(1) - chi) * (T(1) - chi)) / up1 * (z - chi) *
                 (T(1) - T(2) * uu * (z - chi) + uu * uu / z * (z - chi) * (z - chi)) * m2 / sh * ks.k1;
    v2 = pre * (t1 + t2);
    // N
    const T up3 = up1 * up1 * up1;
    const T n1 = T(2) * (T(1) - z) / z / up3 * ((T(2) + u) * u * z + u * u * chi) * sh * ks.k1;
    const T n2 = T(2) * (T(1) - z) / z / up3 * m2 / kc2 * (z / (T(1) - z) + chi / (T(1) - chi) * (u - T(2) * z - T(2) * u * chi)) *
                 (sh * ks.k1 - kz * kb.k1);
    n = pre * (n1 + n2);
    // VMS2
    const T s1 = T(1) / (T(1) - chi) / (up1 * up1) * (-z - uu * (z + u * chi) / z * (chi - (T(1) - z))) * ks.k0;
    const T s2 = T(1) / up3 *
                 (kz * kz / kc2 * (T(1) + u * chi * (T(1) - chi) / (z * (T(1) - z))) -
                  m2 / kc2 * chi / (T(1) - chi) * (T(2) * up1 * up1 / u + u / (z * (T(1) - z)) * (z - chi) * (z - chi))) *
                 (ks.k0 - kb.k0);
    vms2 = pre * (s1 + s2);
}

// ITdip_massive_2 (S/dis/transverse.py:151-168).
template <typename T>
__device__ __forceinline__ T it_dip_2(T Q2, T z1, T x01, T mf, T ych, T yu) {
    const T Q = dis_sqrt(Q2);
    const T kz = dis_sqrt(z1 * (T(1) - z1) * Q * Q + mf * mf);
    T v2a, na, sa, v2b, nb, sb;
    it_v2_n_vms2_u(Q, z1, mf, x01, ych, yu, v2a, na, sa);
    it_v2_n_vms2_u(Q, T(1) - z1, mf, x01, ych, yu, v2b, nb, sb);
    const K01<T> k = bessel_k01(x01 * kz);
    return kz * k.k1 * ((z1 * z1 + (T(1) - z1) * (T(1) - z1)) * (v2a + v2b) + (T(2) * z1 - T(1)) / T(2) * (na - nb)) +
           mf * mf * k.k0 * (sa + sb);
}

// ILdip_massive_Omega_L_Const (S/dis/longitudinal.py:95-115).
template <typename T>
__device__ __forceinline__ T il_dip_const(T Q2, T z, T r, T mf) {
    const T a = dis_sqrt(Q2 * z * (T(1) - z) + mf * mf) * r;
    if (!(a >= T(1e-7))) return T(0);
    const T lz = log(z / (T(1) - z));
    const T k0 = bessel_k0(a);
    return T(4) * Q2 * (z * (T(1) - z)) * (z * (T(1) - z)) * k0 * k0 *
           (T(2.5) + T(-kPiD * kPiD / 3.0 + kPiD * kPiD / 6.0) + T(0.5) * lz * lz + omega_l_v(Q2, z, mf) + l_dip(Q2, z, mf));
}

// ILdip_massive_Iab (S/dis/longitudinal.py:74-92).
template <typename T>
__device__ __forceinline__ T il_dip_ab(T Q2, T z1, T r, T mf, T xi) {
    const T kz = dis_sqrt(z1 * (T(1) - z1) * Q2 + mf * mf);
    const T b1 = bessel_k0(kz * r);
    const T b2 = bessel_k0(dis_sqrt(kz * kz + (T(1) - z1) * xi / (T(1) - xi) * mf * mf) * r);
    const T b3 = bessel_k0(dis_sqrt(kz * kz + z1 * xi / (T(1) - xi) * mf * mf) * r);
    return T(4) * Q2 * (z1 * (T(1) - z1)) * (z1 * (T(1) - z1)) *
           (b1 / xi * (T(-2) * log(xi) / (T(1) - xi) + (T(1) + xi) / T(2)) * (T(2) * b1 - b2 - b3));
}

// ILdip_massive_Icd (S/dis/longitudinal.py:12-71).
template <typename T>
__device__ __forceinline__ T il_dip_cd(T Q2, T z1, T r, T mf, T xi, T x) {
    const T kz = dis_sqrt(z1 * (T(1) - z1) * Q2 + mf * mf);
    const T a = kz * r;
    if (!(a >= T(1e-7))) return T(0);
    const T m2 = mf * mf, omx = T(1) - x, omxi = T(1) - xi, omz = T(1) - z1;
    const T clm1 = z1 * z1 * omxi / omz * (-xi * xi + x * omxi * (T(1) + omxi * (T(1) + z1 * xi / omz)) / (x * omxi + xi / omz));
    const T clm2 = omz * omz * omxi / z1 * (-xi * xi + x * omxi * (T(1) + omxi * (T(1) + omz * xi / z1)) / (x * omxi + xi / z1));
    const T kap1 = xi * m2 / (omxi * omx * (x * omxi + xi / omz)) * (xi * omx + x * (T(1) - z1 * omxi / omz));
    const T kap2 = xi * m2 / (omxi * omx * (x * omxi + xi / z1)) * (xi * omx + x * (T(1) - omz * omxi / z1));
    const T b1 = bessel_k0(a);
    const T v = b1 * m2 *
                ((b1 - bessel_k0(r * dis_sqrt(kz * kz / omx + kap1))) * clm1 /
                     (omxi * omx * (x * omxi + xi / omz) * (x / omx * kz * kz + kap1)) +
                 (b1 - bessel_k0(r * dis_sqrt(kz * kz / omx + kap2))) * clm2 /
                     (omxi * omx * (x * omxi + xi / z1) * (x / omx * kz * kz + kap2)));
    return T(4) * Q2 * (z1 * omz) * (z1 * omz) * v;
}

// ------------------------------------------------------- qqbar-g (a21, a22)

// _tripole_variables (S/dis/transverse.py:854-909); the longitudinal set
// (S/dis/longitudinal.py:339-365) is the same with k -> j and l -> k.
template <typename T>
struct Tripole {
    T z0, x20x21;
    T qj2, qk2;  // Qbar^2 + mf^2
    T om_j, om_k, lam_j, lam_k;
    T x2j2, x2k2, x3j2, x3k2;  // squares
    T sj, sk;                  // x3^2 + omega x2^2
    T x2j_x3j, x2k_x3k, x2j_x3k, x2k_x3j, x3j_x3k;
};

template <typename T>
__device__ __forceinline__ Tripole<T> tripole_vars(T Q2, T mf, T z1, T z2, T x01sq, T x02sq, T x21sq) {
    Tripole<T> v;
    const T z0 = T(1) - z1 - z2;
    v.z0 = z0;
    v.x20x21 = T(-0.5) * (x01sq - x21sq - x02sq);
    v.qj2 = Q2 * z1 * (T(1) - z1) + mf * mf;
    v.qk2 = Q2 * z0 * (T(1) - z0) + mf * mf;
    const T a = z0 + z2, b = z1 + z2;
    v.om_j = z0 * z2 / (z1 * a * a);
    v.om_k = z1 * z2 / (z0 * b * b);
    v.lam_j = z1 * z2 / z0;
    v.lam_k = z0 * z2 / z1;
    v.x2j2 = x02sq;
    v.x2k2 = x21sq;
    v.x3j2 = z0 * z0 / (a * a) * x02sq + x21sq - T(2) * z0 / a * v.x20x21;
    v.x3k2 = z1 * z1 / (b * b) * x21sq + x02sq - T(2) * z1 / b * v.x20x21;
    v.sj = v.x3j2 + v.om_j * x02sq;
    v.sk = v.x3k2 + v.om_k * x21sq;
    v.x2j_x3j = v.x20x21 - z0 / a * x02sq;
    v.x2k_x3k = -v.x20x21 + z1 / b * x21sq;
    v.x2j_x3k = -x02sq + z1 / b * v.x20x21;
    v.x2k_x3j = x21sq - z0 / a * v.x20x21;
    v.x3j_x3k = z0 / a * x02sq + z1 / b * x21sq - (T(1) + z0 * z1 / (a * b)) * v.x20x21;
    return v;
}

// The combination z2 (z0 - z1)^2 (...) - (...) x20x21 x3j_x3k shared by the F terms.
template <typename T>
__device__ __forceinline__ T tripole_cross(const Tripole<T>& v, T z1, T z2) {
    const T z0 = v.z0;
    return z2 * (z0 - z1) * (z0 - z1) * (v.x2j_x3j * v.x2k_x3k - v.x2k_x3j * v.x2j_x3k) -
           (z1 * (z0 + z2) + z0 * (z1 + z2)) * (z0 * (z0 + z2) + z1 * (z1 + z2)) * v.x20x21 * v.x3j_x3k;
}

// G_integrand_simplified (S/dis/longitudinal.py:303-336) for the four
// (a, b) pairs used: (1,1) -> K0, (1,2) and (2,1) -> K1, (2,2) -> K2.
//   G_ab = y^(-(2-a+b)/2) 2^(a+b-1) omega^(b-1) (num/den)^((a+b-2)/2) K_(a+b-2)(sqrt(num den / y)),
//   num = y lam mf^2 + qbar^2 + mf^2, den = y x3^2 + omega x2^2.
template <typename T>
struct GValues {
    T g11, g12, g21, g22;
};

template <typename T>
__device__ __forceinline__ GValues<T> g_values(T q2, T m2, T x22, T x32, T om, T lam, T y) {
    const T num = y * lam * m2 + q2;
    const T den = y * x32 + om * x22;
    const T arg = dis_sqrt(num * den / y);
    const K01<T> k = bessel_k01(arg);
    const T k2 = k.k0 + T(2) * k.k1 / arg;
    const T ratio = dis_sqrt(num / den);
    const T iy = T(1) / y;
    GValues<T> g;
    g.g11 = T(2) * k.k0 * iy;                      // y^-1 2 K0
    g.g12 = T(4) * om * dis_sqrt(iy) * iy * ratio * k.k1;  // y^-3/2 4 omega (num/den)^1/2 K1
    g.g21 = T(4) * dis_sqrt(iy) * ratio * k.k1;         // y^-1/2 4 (num/den)^1/2 K1
    g.g22 = T(8) * om * iy * (num / den) * k2;       // y^-1 8 omega (num/den) K2
    return g;
}

// The eight transverse (four longitudinal) G moments at one inner y:
// G22bar, G12bar (lam minus lam = 0), G21, G11, for the j and k legs.
template <typename T>
struct TMoments {
    T b22j, b22k, b12j, b12k, g21j, g21k, g11j, g11k;
};

// FP32: g_values with 1/y, 1/sqrt(y) and 1/den shared by the lam and lam = 0
// evaluations of a leg and no divisions (den does not depend on lam).
// The argument's square root is the hardware one (about 1 ulp; the argument is
// a product of squared scales, far from the denormal range) and
// (num/den)^1/2 = num y^-1/2 / arg reuses its reciprocal.
__device__ __forceinline__ GValues<float> g_values_f32(float num, float den, float iden, float om, float iy,
                                                       float siy) {
    const float arg = __builtin_amdgcn_sqrtf(num * den * iy);
    const K01<float> k = bessel_k01(arg);
    const float iarg = dis_rcp(arg);
    const float k2 = k.k0 + 2.0f * k.k1 * iarg;
    const float nd = num * iden;
    const float ratio = num * siy * iarg;
    GValues<float> g;
    g.g11 = 2.0f * k.k0 * iy;
    g.g12 = 4.0f * om * siy * iy * ratio * k.k1;
    g.g21 = 4.0f * siy * ratio * k.k1;
    g.g22 = 8.0f * om * iy * nd * k2;
    return g;
}

template <typename T>
__device__ __forceinline__ TMoments<T> t_moments(const Tripole<T>& v, T m2, T y) {
    if constexpr (sizeof(T) == 4) {
        const float iy = dis_rcp(y), siy = dis_rsqrt(y);
        const float dj = y * v.x3j2 + v.om_j * v.x2j2, dk = y * v.x3k2 + v.om_k * v.x2k2;
        const float idj = dis_rcp(dj), idk = dis_rcp(dk);
        const GValues<float> j = g_values_f32(y * v.lam_j * m2 + v.qj2, dj, idj, v.om_j, iy, siy);
        const GValues<float> j0 = g_values_f32(v.qj2, dj, idj, v.om_j, iy, siy);
        const GValues<float> k = g_values_f32(y * v.lam_k * m2 + v.qk2, dk, idk, v.om_k, iy, siy);
        const GValues<float> k0 = g_values_f32(v.qk2, dk, idk, v.om_k, iy, siy);
        return {j.g22 - j0.g22, k.g22 - k0.g22, j.g12 - j0.g12, k.g12 - k0.g12, j.g21, k.g21, j.g11, k.g11};
    } else {
        const GValues<T> j = g_values(v.qj2, m2, v.x2j2, v.x3j2, v.om_j, v.lam_j, y);
        const GValues<T> j0 = g_values(v.qj2, m2, v.x2j2, v.x3j2, v.om_j, T(0), y);
        const GValues<T> k = g_values(v.qk2, m2, v.x2k2, v.x3k2, v.om_k, v.lam_k, y);
        const GValues<T> k0 = g_values(v.qk2, m2, v.x2k2, v.x3k2, v.om_k, T(0), y);
        return {j.g22 - j0.g22, k.g22 - k0.g22, j.g12 - j0.g12, k.g12 - k0.g12, j.g21, k.g21, j.g11, k.g11};
    }
}

template <typename T>
__device__ __forceinline__ void t_moments_axpy(TMoments<T>& acc, T w, const TMoments<T>& m) {
    acc.b22j += w * m.b22j;
    acc.b22k += w * m.b22k;
    acc.b12j += w * m.b12j;
    acc.b12k += w * m.b12k;
    acc.g21j += w * m.g21j;
    acc.g21k += w * m.g21k;
    acc.g11j += w * m.g11j;
    acc.g11k += w * m.g11k;
}

// Quantities of the transverse tripole that do not depend on the inner y.
template <typename T>
struct TFixed {
    T pj, pk;            // (2 z0 (z0+z2) + z2^2)/(z0+z2)^2 and the k mirror
    T g12j, g12k, g22j, g22k, hj, hk;  // _G12_sing, _G22_sing, _H
    T k1j, k1k, k0j, k0k;  // K_nu(sqrt(s) sqrt(qbar^2 + mf^2))
    T cross;
};

template <typename T>
__device__ __forceinline__ TFixed<T> t_fixed(const Tripole<T>& v, T m2, T z1, T z2) {
    TFixed<T> f;
    const T z0 = v.z0;
    f.pj = (T(2) * z0 * (z0 + z2) + z2 * z2) / ((z0 + z2) * (z0 + z2));
    f.pk = (T(2) * z1 * (z1 + z2) + z2 * z2) / ((z1 + z2) * (z1 + z2));
    const T aj = dis_sqrt(v.qj2 * v.sj), ak = dis_sqrt(v.qk2 * v.sk);
    const K01<T> bj = bessel_k01(aj), bk = bessel_k01(ak);
    f.k0j = bj.k0;
    f.k1j = bj.k1;
    f.k0k = bk.k0;
    f.k1k = bk.k1;
    f.g12j = bj.k0 / v.x2j2;
    f.g12k = bk.k0 / v.x2k2;
    f.g22j = dis_sqrt(v.qj2 / v.sj) * bj.k1 / v.x2j2;
    f.g22k = dis_sqrt(v.qk2 / v.sk) * bk.k1 / v.x2k2;
    const T mj = v.qj2 + m2 * v.lam_j, mk = v.qk2 + m2 * v.lam_k;  // qbar^2 + mf^2 (1 + lam)
    f.hj = T(4) * dis_sqrt(mj / v.sj) * bessel_k1(dis_sqrt(mj * v.sj));
    f.hk = T(4) * dis_sqrt(mk / v.sk) * bessel_k1(dis_sqrt(mk * v.sk));
    f.cross = tripole_cross(v, z1, z2);
    return f;
}

// ITNLOqg_massive_dipole_uvsub (S/dis/transverse.py:308-321, 371-399, 437-463).
template <typename T>
__device__ __forceinline__ T it_qg_uvsub(const Tripole<T>& v, T m2, T z1, T z2, T x01sq) {
    const T z0 = v.z0;
    const T pj = (T(2) * z0 * (z0 + z2) + z2 * z2) / ((z0 + z2) * (z0 + z2));
    const T pk = (T(2) * z1 * (z1 + z2) + z2 * z2) / ((z1 + z2) * (z1 + z2));
    const T eg = T(1.7810724179901979852);  // e^gamma
    const T ej = -exp(-v.x2j2 / (x01sq * eg)) / v.x2j2;
    const T ek = -exp(-v.x2k2 / (x01sq * eg)) / v.x2k2;
    const K01<T> bj = bessel_k01(dis_sqrt(x01sq * v.qj2)), bk = bessel_k01(dis_sqrt(x01sq * v.qk2));
    const T jk = pj * (T(1) - T(2) * z1 * (T(1) - z1)) * v.qj2 * ej * bj.k1 * bj.k1 +
                 pk * (T(1) - T(2) * z0 * (T(1) - z0)) * v.qk2 * ek * bk.k1 * bk.k1;
    const T jkm = m2 * (pj * ej * bj.k0 * bj.k0 + pk * ek * bk.k0 * bk.k0);
    return jk + jkm;
}

// ITNLOqg_massive_tripole_part_I1 (S/dis/transverse.py:324-337, 402-434, 466-585).
template <typename T>
__device__ __forceinline__ T it_qg_i1(const Tripole<T>& v, const TFixed<T>& f, T m2, T z1, T z2) {
    const T z0 = v.z0;
    const T jk = f.pj * (T(1) - T(2) * z1 * (T(1) - z1)) * v.qj2 / v.x2j2 * v.x3j2 / v.sj * f.k1j * f.k1j +
                 f.pk * (T(1) - T(2) * z0 * (T(1) - z0)) * v.qk2 / v.x2k2 * v.x3k2 / v.sk * f.k1k * f.k1k;
    const T jkm = m2 * (f.pj / v.x2j2 * f.k0j * f.k0j + f.pk / v.x2k2 * f.k0k * f.k0k);
    const T a = z0 + z2, b = z1 + z2;
    const T F = T(0.5) * (T(4) / (a * b) * f.cross * f.g22j * f.g22k -
                          a * z1 * z2 / (b * b) * v.x2j_x3j * f.hk * f.g22j +
                          b * z0 * z2 / (a * a) * v.x2k_x3k * f.hj * f.g22k -
                          z0 * z0 * z1 * z2 / (a * a * a) * v.x2j_x3j * f.hj * f.g22j +
                          z1 * z1 * z0 * z2 / (b * b * b) * v.x2k_x3k * f.hk * f.g22k +
                          (z0 * z2) * (z0 * z2) / (T(8) * a * a * a * a) * f.hj * f.hj +
                          (z1 * z2) * (z1 * z2) / (T(8) * b * b * b * b) * f.hk * f.hk);
    const T Fm = T(0.5) * m2 * (-((T(2) * z0 + z2) * (T(2) * z1 + z2) + z2 * z2) * v.x20x21 * T(64) * f.g12j * f.g12k / (T(32) * a * b));
    return jk + jkm + F + Fm;
}

// _IT_tripole_I2 from the y-dependent moments (S/dis/transverse.py:587-734).
template <typename T>
__device__ __forceinline__ T it_qg_i2(const Tripole<T>& v, const TFixed<T>& f, T m2, T z1, T z2, const TMoments<T>& m) {
    const T z0 = v.z0, a = z0 + z2, b = z1 + z2;
    const T jk = f.pj * (T(1) - T(2) * z1 * (T(1) - z1)) * m.b22j * v.x3j2 / T(8) * dis_sqrt(v.qj2 / v.sj) * f.k1j +
                 f.pk * (T(1) - T(2) * z0 * (T(1) - z0)) * m.b22k * v.x3k2 / T(8) * dis_sqrt(v.qk2 / v.sk) * f.k1k;
    const T jkm = m2 * (f.pj * m.b12j / T(4) * f.k0j + f.pk * m.b12k / T(4) * f.k0k);
    const T F = T(0.5) * (f.cross / (T(4) * a * b) * (m.b22k * f.g22j + m.b22j * f.g22k) -
                          a * z1 * z2 / (T(16) * b * b) * v.x2j_x3j * f.hk * m.b22j +
                          b * z0 * z2 / (T(16) * a * a) * v.x2k_x3k * f.hj * m.b22k -
                          z0 * z0 * z1 * z2 / (T(16) * a * a * a) * v.x2j_x3j * f.hj * m.b22j +
                          z1 * z1 * z0 * z2 / (T(16) * b * b * b) * v.x2k_x3k * f.hk * m.b22k);
    const T p22 = (T(2) * z0 + z2) * (T(2) * z1 + z2) + z2 * z2;
    const T Fm = T(0.5) * m2 *
                 (-z0 * z1 * z2 * z2 / (T(16) * a * a * a) * v.x2j_x3j * m.g21j * T(8) * f.g12j +
                  z0 * z1 * z2 * z2 / (T(16) * b * b * b) * v.x2k_x3k * m.g21k * T(8) * f.g12k -
                  p22 * v.x20x21 / (T(32) * a * b) * (m.b12k * T(8) * f.g12j + m.b12j * T(8) * f.g12k) -
                  (z0 * z2) * (z0 * z2) / (T(16) * a * b * b) * v.x2j_x3k * m.g21k * T(8) * f.g12j +
                  (z1 * z2) * (z1 * z2) / (T(16) * a * a * b) * v.x2k_x3j * m.g21j * T(8) * f.g12k -
                  z0 * z1 * z2 * z2 / (T(16) * a * a * a) * v.x2j_x3j * m.g11j * T(16) * f.g22j +
                  z0 * z1 * z2 * z2 / (T(16) * b * b * b) * v.x2k_x3k * m.g11k * T(16) * f.g22k -
                  a * z2 * z2 / (T(16) * b * b) * v.x2j_x3j * m.g11k * T(16) * f.g22j +
                  b * z2 * z2 / (T(16) * a * a) * v.x2k_x3k * m.g11j * T(16) * f.g22k +
                  z0 * z2 * z2 * z2 / (T(4) * a * a * a * a) * f.hj * m.g11j +
                  z1 * z2 * z2 * z2 / (T(4) * b * b * b * b) * f.hk * m.g11k);
    return jk + jkm + F + Fm;
}

// _IT_tripole_I3 from the moments at y_t1 (m1) and y_t2 (m2) (S/dis/transverse.py:737-851).
template <typename T>
__device__ __forceinline__ T it_qg_i3(const Tripole<T>& v, const TFixed<T>& f, T mf2, T z1, T z2, const TMoments<T>& m1,
                                      const TMoments<T>& m2) {
    const T z0 = v.z0, a = z0 + z2, b = z1 + z2;
    const T jkm = mf2 * (f.pj * v.x2j2 / T(64) * m1.b12j * m2.b12j + f.pk * v.x2k2 / T(64) * m1.b12k * m2.b12k);
    const T jk = f.pj * (T(1) - T(2) * z1 * (T(1) - z1)) * v.x3j2 * v.x2j2 / T(256) * m1.b22j * m2.b22j +
                 f.pk * (T(1) - T(2) * z0 * (T(1) - z0)) * v.x3k2 * v.x2k2 / T(256) * m1.b22k * m2.b22k;
    const T F = T(0.5) * m1.b22j * m2.b22k / (T(64) * a * b) * f.cross;
    const T p22 = (T(2) * z0 + z2) * (T(2) * z1 + z2) + z2 * z2;
    const T z2_4 = z2 * z2 * z2 * z2;
    const T ra = z2 / a, rb = z2 / b;
    const T Fm = T(0.5) * mf2 *
                 (z2_4 / (T(64) * a * a * a * a) * (T(4) * z1 * (z1 - T(1)) + T(2)) * v.x3j2 * m1.g21j * m2.g21j -
                  z0 * z1 * z2 * z2 / (T(16) * a * a * a) * v.x2j_x3j * m1.b12j * m2.g21j +
                  z2_4 / (T(64) * b * b * b * b) * (T(4) * z0 * (z0 - T(1)) + T(2)) * v.x3k2 * m1.g21k * m2.g21k +
                  z0 * z1 * z2 * z2 / (T(16) * b * b * b) * v.x2k_x3k * m1.b12k * m2.g21k -
                  p22 * v.x20x21 / (T(32) * a * b) * m1.b12j * m2.b12k +
                  z2_4 / (T(32) * a * a * b * b) * p22 * v.x3j_x3k * m1.g21j * m2.g21k +
                  mf2 / T(8) * (ra * ra) * (ra * ra) * m1.g11j * m2.g11j +
                  mf2 / T(8) * (rb * rb) * (rb * rb) * m1.g11k * m2.g11k -
                  (z0 * z2) * (z0 * z2) / (T(16) * a * b * b) * v.x2j_x3k * m1.b12j * m2.g21k +
                  (z1 * z2) * (z1 * z2) / (T(16) * b * a * a) * v.x2k_x3j * m1.b12k * m2.g21j -
                  z0 * z1 * z2 * z2 / (T(16) * a * a * a) * v.x2j_x3j * m1.g11j * m2.b22j +
                  z0 * z1 * z2 * z2 / (T(16) * b * b * b) * v.x2k_x3k * m1.g11k * m2.b22k -
                  a * z2 * z2 / (T(16) * b * b) * v.x2j_x3j * m1.g11k * m2.b22j +
                  b * z2 * z2 / (T(16) * a * a) * v.x2k_x3k * m1.g11j * m2.b22k);
    return jk + jkm + F + Fm;
}

// ILNLOqg_massive_dipole_uvsub (S/dis/longitudinal.py:186-214).
template <typename T>
__device__ __forceinline__ T il_qg_uvsub(const Tripole<T>& v, T Q2, T z1, T z2, T x01sq) {
    const T z0 = v.z0;
    const T eg = T(1.7810724179901979852);
    const T bk = bessel_k0(dis_sqrt(v.qj2) * dis_sqrt(x01sq)), bl = bessel_k0(dis_sqrt(v.qk2) * dis_sqrt(x01sq));
    const T tk = -z1 * z1 * (T(2) * z0 * (z0 + z2) + z2 * z2) / v.x2j2 * exp(-v.x2j2 / x01sq / eg) * bk * bk;
    const T tl = -z0 * z0 * (T(2) * z1 * (z1 + z2) + z2 * z2) / v.x2k2 * exp(-v.x2k2 / x01sq / eg) * bl * bl;
    return T(4) * Q2 * (tk + tl);
}

// ILNLOqg_massive_tripole_part_I1 (S/dis/longitudinal.py:146-183).
template <typename T>
__device__ __forceinline__ T il_qg_i1(const Tripole<T>& v, const TFixed<T>& f, T Q2, T z1, T z2) {
    const T z0 = v.z0;
    const T tk = z1 * z1 * (T(2) * z0 * (z0 + z2) + z2 * z2) / v.x2j2 * f.k0j * f.k0j;
    const T tl = z0 * z0 * (T(2) * z1 * (z1 + z2) + z2 * z2) / v.x2k2 * f.k0k * f.k0k;
    const T tkl = T(-2) * z0 * z1 * (z0 * (T(1) - z0) + z1 * (T(1) - z1)) * v.x20x21 / (v.x2j2 * v.x2k2) * f.k0j * f.k0k;
    return T(4) * Q2 * (tk + tl + tkl);
}

// ILNLOqg_massive_tripole_part_I2 from the moments (S/dis/longitudinal.py:217-255).
template <typename T>
__device__ __forceinline__ T il_qg_i2(const Tripole<T>& v, const TFixed<T>& f, T Q2, T z1, T z2, const TMoments<T>& m) {
    const T z0 = v.z0;
    const T tk = z1 * z1 * (T(2) * z0 * (z0 + z2) + z2 * z2) / T(4) * m.b12j * f.k0j;
    const T tl = z0 * z0 * (T(2) * z1 * (z1 + z2) + z2 * z2) / T(4) * m.b12k * f.k0k;
    const T tkl = -z0 * z1 * (z0 * (T(1) - z0) + z1 * (T(1) - z1)) * v.x20x21 / T(4) *
                  (m.b12j * f.k0k / v.x2k2 + m.b12k * f.k0j / v.x2j2);
    return T(4) * Q2 * (tk + tl + tkl);
}

// ILNLOqg_massive_tripole_part_I3 from the moments (S/dis/longitudinal.py:258-300).
template <typename T>
__device__ __forceinline__ T il_qg_i3(const Tripole<T>& v, T Q2, T mf2, T z1, T z2, const TMoments<T>& m1,
                                      const TMoments<T>& m2) {
    const T z0 = v.z0;
    const T tk = z1 * z1 * (T(2) * z0 * (z0 + z2) + z2 * z2) * v.x2j2 / T(64) * m1.b12j * m2.b12j;
    const T tl = z0 * z0 * (T(2) * z1 * (z1 + z2) + z2 * z2) * v.x2k2 / T(64) * m1.b12k * m2.b12k;
    const T tkl = -z1 * z0 * (z1 * (T(1) - z1) + z0 * (T(1) - z0)) * v.x20x21 / T(32) * m1.b12j * m2.b12k;
    const T rk = z1 / (z0 + z2), rl = z0 / (z1 + z2);
    const T tmf = mf2 / T(16) * z2 * z2 * z2 * z2 *
                  (rk * rk * m1.g11j * m2.g11j + rl * rl * m1.g11k * m2.g11k - T(2) * rl * rk * m1.g11j * m2.g11k);
    return T(4) * Q2 * (tk + tl + tkl + tmf);
}

}  // namespace dis
}  // namespace nlo
