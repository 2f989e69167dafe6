// Device-side BK physics: coupling, log-log splines held in LDS, the K1
// integrand and the K2 + Kf integrand in FP64 (reference formula) and FP32
// (cancellation-free formula).  Reference prefixes: S/ = torch_nlo_clean/src/
// nlo_torch/, K/ = S/custom_kernels/.
#pragma once

#include <hip/hip_runtime.h>

#include "bessel_j1_i1.hpp"
#include "common.hpp"

namespace nlo {

constexpr double kPi = 3.14159265358979323846;
constexpr double kLn2 = 0.69314718055994530942;
constexpr double kLog2e = 1.44269504088896340736;

// Physics constants derived once on the host from nlo_bk_physics.
struct BkConsts {
    double nc, nf, ksub, minr;
    double b0;          // (11 Nc - 2 Nf) / (12 pi)
    double log_mu0;     // 2 / 0.2 * ln 2.5
    double scale_num;   // 4 C2
    double lambda_qcd;
    double log_scale0;  // ln(4 C2 / Lambda^2)
    double norm;        // 1 / (8 pi^4)
    int symmetrize;
    int rc_lo, rc_nlo, resum_rc;  // NLO_RC_LO_*, NLO_RC_NLO_*, NLO_RESUM_RC_*
    int has_resum;      // Order resums the double logs (all but LO and NLO)
    int slog;           // ... and the single logs
    int fixed;          // every coupling is FIXED_AS (RC_LO or RC_NLO fixed)
    int thresholds;     // NF > 3: one-loop coupling with flavour thresholds
    int production;     // Balitsky LO, smallest resummation, DLOG_SLOG, smooth coupling
    double fixed_as;
    double log_4c2;     // ln(4 C2)
};

// -------------------------------------------------------------- coupling

// S/coupling.py:81-97: alpha = 1 / (b0 * 0.2 * logaddexp(2/0.2 ln 2.5,
// ln(4 C2 / (r^2 Lambda^2)) / 0.2)).
__device__ __forceinline__ double alpha_s64(double r, const BkConsts& c) {
    const double scale = c.scale_num / (r * r * c.lambda_qcd * c.lambda_qcd);
    const double right = log(scale) / 0.2;
    const double hi = fmax(c.log_mu0, right);
    return 1.0 / (c.b0 * (0.2 * (hi + log1p(exp(-fabs(c.log_mu0 - right))))));
}

// bk_alpha_s for NF > 3 (S/coupling.py:23-46): Nf and Lambda switched at
// 4 C2 / r^2 = 1.3^2 and 4.5^2, b0 = 11 - 2 Nf / 3, alpha = 4 pi / (b0 ln(4 C2 /
// (r^2 Lambda^2))), 1 where the logarithm's argument is below 1 or alpha > 1.
__device__ __forceinline__ double alpha_thresholds64(double r, const BkConsts& c) {
    const double dipolescale = c.scale_num / (r * r);
    double nf, lambda;
    if (dipolescale < 1.3 * 1.3) {
        nf = 3.0;
        lambda = 0.146159;
    } else if (dipolescale < 4.5 * 4.5) {
        nf = 4.0;
        lambda = 0.122944;
    } else {
        nf = 5.0;
        lambda = 0.0904389;
    }
    const double b0 = 11.0 - 2.0 * nf / 3.0;
    const double log_arg = c.scale_num / (r * r * (lambda * lambda));
    const double a = 4.0 * kPi / (b0 * log(log_arg));
    return (log_arg < 1.0 || a > 1.0) ? 1.0 : a;
}

// _alpha_s of S/bk_kernels/k1.py:158-169 and S/bk/derivatives.py:775-786:
// FIXED_AS, the smooth coupling (NF <= 3) or the threshold coupling.
__device__ __forceinline__ double alpha_any64(double r, const BkConsts& c) {
    if (c.fixed) return c.fixed_as;
    return c.thresholds ? alpha_thresholds64(r, c) : alpha_s64(r, c);
}

// The same threshold coupling from ln d (FP32 NLO integrand):
// ln(4 C2 / d^2) against 2 ln 1.3 and 2 ln 4.5, ln(arg) = that - 2 ln Lambda.
template <typename T>
__device__ __forceinline__ T alpha_thresholds_log(T log_d, T log_4c2) {
    const T lq = log_4c2 - T(2) * log_d;
    T nf = T(5), l2 = T(-4.806161588812351);  // 2 ln 0.0904389
    if (lq < T(0.5247285289349821)) {         // 2 ln 1.3
        nf = T(3);
        l2 = T(-3.8461204175059556);          // 2 ln 0.146159
    } else if (lq < T(3.0081547935525483)) {  // 2 ln 4.5
        nf = T(4);
        l2 = T(-4.192052623672091);           // 2 ln 0.122944
    }
    const T L = lq - l2;
    const T a = T(4.0 * kPi) / ((T(11) - T(2) * nf / T(3)) * L);
    return (L < T(0) || a > T(1)) ? T(1) : a;
}

// Coupling of the K2 + Kf integrand (_nlo_alpha_s, S/bk/derivatives.py:615-632):
// mode 0 = smooth coupling at the smallest distance (production), 1 = a
// constant per parent (FIXED_AS or alpha(r)), 2 = threshold coupling at the
// smallest distance.
template <typename T>
struct NloCoupling {
    int mode;
    T b0, log_mu0, log_scale0, log_4c2, as_const;
};

// ANYC = false: the production coupling (mode 0) compiled alone, so that the
// production kernels carry no code or registers of the other modes.
template <bool ANYC>
__device__ __forceinline__ double nlo_alpha_from_log(double log_min, const NloCoupling<double>& k) {
    if constexpr (ANYC) {
        if (k.mode == 1) return k.as_const;
        if (k.mode == 2) return alpha_thresholds_log<double>(log_min, k.log_4c2);
    }
    const double right = 5.0 * (k.log_scale0 - 2.0 * log_min);
    const double hi = fmax(k.log_mu0, right);
    return 1.0 / (k.b0 * (0.2 * (hi + log1p(exp(-fabs(k.log_mu0 - right))))));
}

// NloCoupling for a parent: the mode from the physics choice, the constant
// for FIXED_NLO / PARENT_NLO evaluated at r in FP64.
template <typename T>
__device__ __forceinline__ NloCoupling<T> nlo_coupling(double r, const BkConsts& c) {
    NloCoupling<T> k;
    k.b0 = (T)c.b0;
    k.log_mu0 = (T)c.log_mu0;
    k.log_scale0 = (T)c.log_scale0;
    k.log_4c2 = (T)c.log_4c2;
    k.as_const = T(0);
    if (c.fixed || c.rc_nlo == NLO_RC_NLO_FIXED) {
        k.mode = 1;
        k.as_const = (T)c.fixed_as;
    } else if (c.rc_nlo == NLO_RC_NLO_PARENT) {
        k.mode = 1;
        k.as_const = (T)alpha_any64(r, c);
    } else {
        k.mode = c.thresholds ? 2 : 0;
    }
    return k;
}

// Same coupling from ln r (FP32): right = 5 (ln(4 C2 / Lambda^2) - 2 ln r).
__device__ __forceinline__ float alpha_s32_log(float log_r, float b0, float log_mu0,
                                               float log_scale0) {
    const float right = 5.0f * (log_scale0 - 2.0f * log_r);
    const float hi = fmaxf(log_mu0, right);
    return 1.0f / (b0 * (0.2f * (hi + log1pf(expf(-fabsf(log_mu0 - right))))));
}

// --------------------------------------------------------------- splines

// A log-log spline staged in LDS: nodes x[n] = ln r_i and per-interval
// records {a, b, c, d} (structure of arrays would need four LDS reads).
template <typename T>
struct LdsSpline {
    const T* x;
    const T* rec;  // 4 (n - 1)
    int n;
    T x0, inv_h;   // ln r_0 and (n - 1) / (ln r_{n-1} - ln r_0)
    T log_min, log_max;
};

// searchsorted(x, v, right=True) - 1 clamped to [0, n - 2]
// (S/numerics/interpolation.py:55-56): the uniform-grid estimate, then exact
// correction against the nodes (one step on the log-uniform evolution grid).
template <typename T>
__device__ __forceinline__ int spline_interval(const LdsSpline<T>& s, T v) {
    int k = (int)((v - s.x0) * s.inv_h);
    k = k < 0 ? 0 : (k > s.n - 2 ? s.n - 2 : k);
    while (k > 0 && v < s.x[k]) --k;
    while (k < s.n - 2 && v >= s.x[k + 1]) ++k;
    return k;
}

template <typename T>
__device__ __forceinline__ T spline_poly(const LdsSpline<T>& s, T v) {
    const int k = spline_interval(s, v);
    const T dx = v - s.x[k];
    const T* c = s.rec + 4 * k;
    return c[0] + dx * (c[1] + dx * (c[2] + dx * c[3]));
}

// _evaluate_loglog (S/bk/derivatives.py:761-772) on a distance d.
__device__ __forceinline__ double loglog64(const LdsSpline<double>& s, double r_min, double r_max,
                                          double d, double underflow, double overflow) {
    const double ev = fmin(fmax(d, r_min), r_max);
    double out = exp(spline_poly(s, log(ev)));
    out = isfinite(out) ? out : 0.0;
    out = d < r_min ? underflow : out;
    return d > r_max ? overflow : out;
}

// Same from ln d: the clamp and the range tests done in log space.
template <typename T>
__device__ __forceinline__ T loglog_from_log(const LdsSpline<T>& s, T log_d, T underflow,
                                            T overflow) {
    const T lv = fmin(fmax(log_d, s.log_min), s.log_max);
    T out = exp(spline_poly(s, lv));
    out = isfinite(out) ? out : T(0);
    out = log_d < s.log_min ? underflow : out;
    return log_d > s.log_max ? overflow : out;
}

// ------------------------------------------------------- K1 (FP64 only)

// Kernel_lo (S/bk_kernels/k1.py:20-155) for BALITSKY_LO, RESUM_RC_SMALLEST,
// NLO_RESUM_DLOG_SLOG, times the dipole of _lo_integrand
// (S/bk/derivatives.py:445-468), split so that everything that depends only
// on the daughter distance Y = z is computed once per z node (K1Z) and the
// theta loop evaluates the rest (k1_theta_term).  X^2 = (r - z)^2 +
// 4 r z sin^2(theta / 2) (the form K/bk/k1/kernel.cu:478-479 uses; it keeps
// X accurate near X -> 0).  Logarithms of distances are formed once and
// shared by the spline, the coupling and the resummation logs.

// alpha(d) from ln d (S/coupling.py:81-97): ln(4 C2 / (d^2 Lambda^2)) =
// log_scale0 - 2 ln d.  The denominator is returned too, so that ratios of
// couplings need no division: alpha_a / alpha_b = alpha_a * den_b.
// Far from the freezing scale (|mu0 - right| > 40) the smooth-max term
// log1p(e^-|d|) < 4.3e-18 is below half an ulp of hi >= mu0 = 10 ln 2.5, so
// hi + term rounds to hi: skipping the exp and log1p there changes no bit,
// and whole waves (small distances) skip them.
__device__ __forceinline__ double alpha_s64_den(double log_d, const BkConsts& c) {
    const double right = (c.log_scale0 - 2.0 * log_d) * 5.0;  // / 0.2 within one rounding
    const double hi = fmax(c.log_mu0, right);
    const double d = fabs(c.log_mu0 - right);
    double sum = hi;
    if (!(d > 40.0)) sum = hi + log1p(exp(-d));  // NaN keeps the full formula
    return c.b0 * (0.2 * sum);
}

struct K1Parent {
    double r, log_r, n_r, a_r;
    double lo_coef;     // Nc / (2 pi^2) alpha(r)
    double fin_coef;    // alpha(r) Nc / (4 pi) (67/9 - pi^2/3 - 10/9 Nf/Nc)
    double log_ksub;
    double resum_on;    // 1 where r > 1.01 MINR (double-log resummation active)
    const double* j1tab;  // J1 Chebyshev pieces (LDS)
};

struct K1Z {
    double z, log_z, n_y, a_y, den_y, ly, inv_y2;
};

__device__ __forceinline__ double k1_theta_term(const K1Parent& P, const K1Z& Z, double s2,
                                                const LdsSpline<double>& sN, const BkConsts& c) {
    const double r = P.r, z = Z.z;
    const double xsq = (r - z) * (r - z) + 4.0 * r * z * s2;
    if (xsq < 1e-40 || z < 1e-20 || r < 1e-20) return 0.0;
    const double X = sqrt(xsq);
    const double lnX = log(X);
    const double n_x = fmin(loglog_from_log(sN, lnX, 0.0, 1.0), 1.0);
    const double dipole = n_x + Z.n_y - P.n_r - n_x * Z.n_y;
    const double den_x = alpha_s64_den(lnX, c);
    // one division for both reciprocals: q = 1 / (den_x X^2)
    const double q = 1.0 / (den_x * xsq);
    const double a_x = q * xsq;
    const double inv_x2 = q * den_x;
    const double Y = z;
    const double a_min = fmin(X, Y) <= r ? (X <= Y ? a_x : Z.a_y) : P.a_r;
    const double rxy2 = r * r * inv_x2 * Z.inv_y2;  // r^2 / (X^2 Y^2)
    // Kernel_lo: alpha_r (r^2/(X^2 Y^2) + (alpha_Y/alpha_X - 1)/Y^2 + (alpha_X/alpha_Y - 1)/X^2)
    double lo = P.lo_coef * (rxy2 + (Z.a_y * den_x - 1.0) * Z.inv_y2 + (a_x * Z.den_y - 1.0) * inv_x2);
    lo = isfinite(lo) ? lo : 0.0;
    const double xl = 4.0 * (lnX - P.log_r) * Z.ly;
    const double abar = a_min * (c.nc / kPi);  // loop-invariant quotient
    // J1(2a)/a (xl >= 0) or I1(2a)/a (xl < 0), a^2 = abar |xl| (k1.py:108-132):
    // one series in s = -abar xl for |s| <= 16 (no divergence between the two
    // sides), the J1 table or the large-argument I1 beyond
    const double sres = -abar * xl;
    double dlog = 1.0;
    if (P.resum_on != 0.0 && sres != 0.0) {
        if (fabs(sres) <= 16.0) {
            dlog = j1_i1_series(sres);
        } else {
            const double as_x = sqrt(fabs(sres));
            dlog = sres < 0.0 ? j1_over_half_x(2.0 * as_x, P.j1tab) : i1_over_half_x(2.0 * as_x);
        }
    }
    if (!isfinite(dlog)) return 0.0;
    // ln(KSUB r^2 / min(X, Y)^2): the single-log exponent of both slog terms
    const double lmin = X <= Y ? lnX : Z.log_z;
    const double slog_exp = -abar * (11.0 / 12.0) * fabs(P.log_ksub + 2.0 * (P.log_r - lmin));
    const double slog = exp(slog_exp);
    const double lo_single = P.lo_coef * rxy2;
    const double k1fin = lo * P.fin_coef;
    return (dlog * slog * lo - lo_single * slog_exp + k1fin) * dipole;
}

// ------------------------------------------ K1 for every physics choice

// Kernel_lo (S/bk_kernels/k1.py:20-155) for any RC_LO / RESUM_RC / Order
// choice, at valid distances X, Y (logs given).  j1tab: the J1 table in LDS.
__device__ inline double kernel_lo_any(double r, double log_r, double X, double lnX, double Y,
                                       double lnY, const BkConsts& c, const double* j1tab) {
    const double mn = fmin(fmin(X, Y), r);
    const double a_r = alpha_any64(r, c);
    const double rxy2 = (r / (X * Y)) * (r / (X * Y));
    const double pref = c.nc / (2.0 * kPi * kPi);
    double lo, a_scale;
    switch (c.rc_lo) {
        case NLO_RC_LO_SMALLEST:
            a_scale = alpha_any64(mn, c);
            lo = pref * a_scale * rxy2;
            break;
        case NLO_RC_LO_PARENT:
            a_scale = a_r;
            lo = pref * a_r * rxy2;
            break;
        case NLO_RC_LO_FAC: {
            const double ab_r = a_r * c.nc / kPi;
            const double ab_x = alpha_any64(X, c) * c.nc / kPi;
            const double ab_y = alpha_any64(Y, c) * c.nc / kPi;
            a_scale = a_r;
            lo = 1.0 / (2.0 * kPi) /
                 (1.0 / ab_r + (X * X - Y * Y) / (r * r) * (ab_x - ab_y) / (ab_x * ab_y)) * rxy2;
            break;
        }
        case NLO_RC_LO_BEUF: {
            const double reff2 = r * r * pow(Y * Y / (X * X), (X * X - Y * Y) / (r * r));
            a_scale = alpha_any64(sqrt(reff2), c);
            const double q = r / (X * Y + 1e-40);
            lo = pref * a_scale * q * q;
            break;
        }
        default: {  // BALITSKY_LO, FIXED_LO
            const double a_x = alpha_any64(X, c), a_y = alpha_any64(Y, c);
            a_scale = a_r;
            lo = pref * a_r * (rxy2 + (a_y / a_x - 1.0) / (Y * Y) + (a_x / a_y - 1.0) / (X * X));
            break;
        }
    }
    lo = isfinite(lo) ? lo : 0.0;
    if (!c.has_resum) return lo;
    const double a_res = c.resum_rc == NLO_RESUM_RC_PARENT  ? a_r
                         : c.resum_rc == NLO_RESUM_RC_FIXED ? c.fixed_as
                                                            : alpha_any64(mn, c);
    const double abar = a_res * c.nc / kPi;
    const double xl = 4.0 * (lnX - log_r) * (lnY - log_r);
    const double sres = -abar * xl;  // a^2 with the J1 side negative, as in k1_theta_term
    double dlog = 1.0;
    if (r > 1.01 * c.minr && sres != 0.0) {
        if (fabs(sres) <= 16.0) {
            dlog = j1_i1_series(sres);
        } else {
            const double as_x = sqrt(fabs(sres));
            dlog = sres < 0.0 ? j1_over_half_x(2.0 * as_x, j1tab) : i1_over_half_x(2.0 * as_x);
        }
    }
    if (!isfinite(dlog)) return 0.0;
    double slog = 1.0, slog_exp = 0.0;
    if (c.slog) {
        // ln(KSUB r^2 / min(X, Y)^2), shared by both single-log terms
        slog_exp = -abar * (11.0 / 12.0) * fabs(log(c.ksub) + 2.0 * (log_r - fmin(lnX, lnY)));
        slog = exp(slog_exp);
    }
    const double lo_single = pref * a_scale * rxy2;
    const double fin = a_scale * c.nc / (4.0 * kPi) * (67.0 / 9.0 - kPi * kPi / 3.0 - 10.0 / 9.0 * c.nf / c.nc);
    return dlog * slog * lo - lo_single * slog_exp + lo * fin;
}

// Amplitude history for the kinematical constraint: rows N(r_grid, y_k),
// k < *rows, in HBM; the current rapidity is the last row's.
struct K1History {
    const double* y;
    const double* N;
    const double* r;
    const int32_t* rows;  // device row count, read by the kernel (clamped to capacity)
    int n;
    int capacity;
};

// NaturalCubicSpline on six nodes evaluated at v (S/numerics/interpolation.py:10-68).
__device__ inline double natural6(const double* x, const double* y, double v) {
    double h[5], up[4], sol[4];
    for (int i = 0; i < 5; ++i) h[i] = x[i + 1] - x[i];
    for (int i = 0; i < 4; ++i) {
        const double diag = 2.0 * (h[i] + h[i + 1]);
        const double rhs = 6.0 * ((y[i + 2] - y[i + 1]) / h[i + 1] - (y[i + 1] - y[i]) / h[i]);
        const double den = i == 0 ? diag : diag - h[i] * up[i - 1];
        up[i] = h[i + 1] / den;
        sol[i] = (rhs - (i == 0 ? 0.0 : h[i] * sol[i - 1])) / den;
    }
    for (int i = 2; i >= 0; --i) sol[i] -= up[i] * sol[i + 1];
    int k = 0;
    while (k < 4 && x[k + 1] <= v) ++k;
    const double m0 = k == 0 ? 0.0 : sol[k - 1], m1 = k == 4 ? 0.0 : sol[k];
    const double b = (y[k + 1] - y[k]) / h[k] - h[k] * (2.0 * m0 + m1) / 6.0;
    const double dx = v - x[k];
    return y[k] + dx * (b + dx * (m0 / 2.0 + dx * (m1 - m0) / (6.0 * h[k])));
}

// _interpolate_history_N (S/bk/derivatives.py:712-758) at (d, rapidity),
// 0 <= rapidity <= the last row's rapidity; the last row is read through the
// current spline sN.
__device__ inline double history_n(double d, double log_d, double rapidity, const K1History& H,
                                   int rows, const LdsSpline<double>& sN) {
    int lo = 0, hi = rows;  // first row with y > rapidity
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (H.y[mid] <= rapidity) lo = mid + 1; else hi = mid;
    }
    const int yind = lo - 1 < 0 ? 0 : lo - 1;
    double out;
    if (yind == rows - 1) {
        out = loglog_from_log(sN, log_d, 0.0, 1.0);
    } else {
        int a = 0, b = H.n;  // first node with r > d
        while (a < b) {
            const int mid = (a + b) >> 1;
            if (H.r[mid] <= d) a = mid + 1; else b = mid;
        }
        const int rind = a - 1;
        const int start = rind - 3 > 0 ? rind - 3 : 0;
        if (start + 6 > H.n) {
            out = 1.0;
        } else {
            double x[6], yl[6], yu[6];
            for (int i = 0; i < 6; ++i) {
                x[i] = H.r[start + i];
                yl[i] = H.N[(size_t)yind * H.n + start + i];
                yu[i] = H.N[(size_t)(yind + 1) * H.n + start + i];
            }
            double v = d;
            if (v < x[0]) v = x[0] * 1.00001;
            if (v > x[5]) v = x[5] * 0.999999;
            const double nl = natural6(x, yl, v), nu = natural6(x, yu, v);
            const double frac = (rapidity - H.y[yind]) / (H.y[yind + 1] - H.y[yind]);
            out = nl + frac * (nu - nl);
        }
    }
    return fmin(fmax(out, 0.0), 1.0);
}

// The K1 integrand (_lo_integrand, S/bk/derivatives.py:434-468) for any
// physics choice, with the kinematical constraint when H.rows is set.
__device__ inline double k1_theta_term_any(const K1Parent& P, const K1Z& Z, double s2,
                                           const LdsSpline<double>& sN, const BkConsts& c,
                                           const K1History& H, int rows, double rapidity) {
    const double r = P.r, z = Z.z;
    const double xsq = (r - z) * (r - z) + 4.0 * r * z * s2;
    if (xsq < 1e-40 || z < 1e-20 || r < 1e-20) return 0.0;
    const double X = sqrt(xsq);
    const double lnX = log(X);
    const double k = kernel_lo_any(r, P.log_r, X, lnX, z, Z.log_z, c, P.j1tab);
    if (rows == 0) {
        const double n_x = fmin(loglog_from_log(sN, lnX, 0.0, 1.0), 1.0);
        return k * (n_x + Z.n_y - P.n_r - n_x * Z.n_y);
    }
    // rapidity shifted by max(0, ln(min(X, Y)^2 / r^2)); cut below zero
    const double delta = fmax(2.0 * (fmin(lnX, Z.log_z) - P.log_r), 0.0);
    const double shifted = rapidity - delta;
    if (shifted < 0.0) return 0.0;
    const double s02 = 1.0 - history_n(X, lnX, shifted, H, rows, sN);
    const double s12 = 1.0 - history_n(z, Z.log_z, shifted, H, rows, sN);
    return k * (-s02 * s12 + (1.0 - P.n_r));
}

// -------------------------------------------------- K2 + Kf, FP64 direct

struct NloValue64 {
    double total, k2, kf;
};

__device__ __forceinline__ double kernel_nlo64(double r, double X, double Y, double X2, double Y2,
                                               double Z, double z4) {
    // S/bk_kernels/k2.py:24-36
    const double a2 = (X * Y2) * (X * Y2), b2 = (X2 * Y) * (X2 * Y);
    const double diff = a2 - b2;
    const double rz = r * Z;
    const double k = -2.0 / z4 + ((a2 + b2 - 4.0 * rz * rz) / (z4 * diff) + (r * r) * (r * r) / (a2 * diff) +
                                  r * r / ((X * Y2 * Z) * (X * Y2 * Z))) *
                                     2.0 * log(X * Y2 / (X2 * Y));
    return isfinite(k) ? k : 0.0;
}

__device__ __forceinline__ double kernel_nlo_fermion64(double r, double X, double Y, double X2,
                                                       double Y2, double Z, double z4, double nf,
                                                       double nc) {
    // S/bk_kernels/kf.py:28-38
    const double a2 = (X * Y2) * (X * Y2), b2 = (X2 * Y) * (X2 * Y);
    const double rz = r * Z;
    double k = 2.0 / z4 - (a2 + b2 - rz * rz) / (z4 * (a2 - b2)) * 2.0 * log(X * Y2 / (X2 * Y));
    k = k * nf / nc;
    return isfinite(k) ? k : 0.0;
}

// _nlo_integrand (S/bk/derivatives.py:521-612) times exp(2 ln z + 2 ln z')
// (S/bk/derivatives.py:335), with the reference's kernels and S dipoles.
template <bool ANYC>
__device__ NloValue64 nlo_integrand64(double r, double lz, double lz2, double tz, double tz2,
                                      const LdsSpline<double>& sS, double r_min, double r_max,
                                      const BkConsts& c) {
    NloValue64 out{0.0, 0.0, 0.0};
    const double z = exp(lz), z2 = exp(lz2);
    const double s1 = sin(tz / 2.0), s2 = sin(tz2 / 2.0), s3 = sin((tz - tz2) / 2.0);
    const double xsq = (r - z) * (r - z) + 4.0 * r * z * s1 * s1;
    const double x2sq = (r - z2) * (r - z2) + 4.0 * r * z2 * s2 * s2;
    const double zsq = (z - z2) * (z - z2) + 4.0 * z * z2 * s3 * s3;
    const double X = sqrt(xsq), X2 = sqrt(x2sq), Z = sqrt(zsq);
    if (X < 1e-20 || z < 1e-20 || X2 < 1e-20 || z2 < 1e-20 || Z < 1e-20) return out;
    const double Y = z, Y2 = z2;
    const double S_X = loglog64(sS, r_min, r_max, X, 1.0, 0.0);
    const double S_Y = loglog64(sS, r_min, r_max, Y, 1.0, 0.0);
    const double S_X2 = loglog64(sS, r_min, r_max, X2, 1.0, 0.0);
    const double S_Y2 = loglog64(sS, r_min, r_max, Y2, 1.0, 0.0);
    const double S_Z = loglog64(sS, r_min, r_max, Z, 1.0, 0.0);
    const double dmin = fmin(fmin(fmin(r, X), Y), fmin(fmin(X2, Y2), Z));
    // _nlo_alpha_s (S/bk/derivatives.py:615-632)
    double as;
    if constexpr (ANYC)
        as = (c.fixed || c.rc_nlo == NLO_RC_NLO_FIXED) ? c.fixed_as
             : c.rc_nlo == NLO_RC_NLO_PARENT           ? alpha_any64(r, c)
                                                       : alpha_any64(dmin, c);
    else
        as = alpha_s64(dmin, c);
    const double z4 = Z * Z * Z * Z;
    double k = kernel_nlo64(r, X, Y, X2, Y2, Z, z4);
    double ks = kernel_nlo64(r, X2, Y2, X, Y, Z, z4);
    double dip = -(S_X * S_Z * S_Y2 - S_X * S_Y);
    double dips = -(S_X2 * S_Z * S_Y - S_X2 * S_Y2);
    if (fabs(k) > 1e10 && fabs(dip) < 1e-10) { k = 0.0; dip = 0.0; }
    if (fabs(ks) > 1e10 && fabs(dips) < 1e-10) { ks = 0.0; dips = 0.0; }
    double k2r = c.symmetrize ? (k * dip + ks * dips) / 2.0 : k * dip;
    double kfr = 0.0;
    if (c.nf > 0.0) {
        const double kf = kernel_nlo_fermion64(r, X, Y, X2, Y2, Z, z4, c.nf, c.nc);
        const double kfs = kernel_nlo_fermion64(r, X2, Y2, X, Y, Z, z4, c.nf, c.nc);
        const double df = S_Y * (S_X2 - S_X);
        const double dfs = S_Y2 * (S_X - S_X2);
        kfr = -(c.symmetrize ? (kf * df + kfs * dfs) / 2.0 : kf * df);
    }
    const double scale = (as * c.nc) * (as * c.nc) * c.norm * exp(2.0 * lz + 2.0 * lz2);
    k2r *= scale;
    kfr *= scale;
    out.k2 = isfinite(k2r) ? k2r : 0.0;
    out.kf = isfinite(kfr) ? kfr : 0.0;
    out.total = out.k2 + out.kf;
    return out;
}

// ------------------------------------- K2 + Kf, cancellation-free kernels

// g(L) = L coth L - 1 (even).  |L| < 1: L^2 P(L^2) / Q(L^2) with
// L cosh L - sinh L = L^3 P(L^2), sinh L = L Q(L^2), both series positive;
// otherwise |L| + h(|L|) - 1.
template <typename T>
__device__ __forceinline__ T coth_minus_one(T L);

template <>
__device__ __forceinline__ double coth_minus_one<double>(double L) {
    L = fabs(L);
    if (L < 1.0) {
        const double y = L * L;
        const double p =
            ((((((((((1.0 / 1175091669949317120000.0) * y + 1.0 / 2554547108585472000.0) * y +
                    1.0 / 6758061133824000.0) * y + 1.0 / 22230464256000.0) * y +
                  1.0 / 93405312000.0) * y + 1.0 / 518918400.0) * y + 1.0 / 3991680.0) * y +
               1.0 / 45360.0) * y + 1.0 / 840.0) * y + 1.0 / 30.0) * y + 1.0 / 3.0;
        const double q =
            (((((((((((1.0 / 25852016738884976640000.0) * y + 1.0 / 51090942171709440000.0) * y +
                     1.0 / 121645100408832000.0) * y + 1.0 / 355687428096000.0) * y +
                   1.0 / 1307674368000.0) * y + 1.0 / 6227020800.0) * y + 1.0 / 39916800.0) * y +
                1.0 / 362880.0) * y + 1.0 / 5040.0) * y + 1.0 / 120.0) * y + 1.0 / 6.0) * y + 1.0;
        return y * p / q;
    }
    return L + 2.0 * L / expm1(2.0 * L) - 1.0;
}

template <>
__device__ __forceinline__ float coth_minus_one<float>(float L) {
    L = fabsf(L);
    if (L < 1.0f) {
        const float y = L * L;
        const float p = ((((1.0f / 518918400.0f) * y + 1.0f / 3991680.0f) * y + 1.0f / 45360.0f) * y +
                         1.0f / 840.0f) * y + 1.0f / 30.0f;
        const float q = ((((1.0f / 39916800.0f) * y + 1.0f / 362880.0f) * y + 1.0f / 5040.0f) * y +
                         1.0f / 120.0f) * y + 1.0f / 6.0f;
        return y * (p * y + 1.0f / 3.0f) / (q * y + 1.0f);
    }
    return L + 2.0f * L / expm1f(2.0f * L) - 1.0f;
}

// ln(X / Y) with X^2 = (r - z)^2 + 4 r z sin^2(theta / 2), Y = z:
// (X^2 - z^2) / z^2 = rho (rho - 2 cos theta), rho = r / z.
template <typename T>
__device__ __forceinline__ T log_x_over_y(T r, T z, T log_z, T cos_t, T xsq);

template <>
__device__ __forceinline__ double log_x_over_y<double>(double r, double z, double log_z,
                                                       double cos_t, double xsq) {
    const double rho = r / z;
    const double u = rho * (rho - 2.0 * cos_t);
    return fabs(u) < 0.5 ? 0.5 * log1p(u) : 0.5 * log(xsq) - log_z;
}

template <>
__device__ __forceinline__ float log_x_over_y<float>(float r, float z, float log_z, float cos_t,
                                                     float xsq) {
    const float rho = r / z;
    const float u = rho * (rho - 2.0f * cos_t);
    return fabsf(u) < 0.5f ? 0.5f * log1pf(u) : 0.5f * logf(xsq) - log_z;
}

struct NloValue32 {
    float total, k2, kf;
};

// The K2 + Kf integrand of S/bk/derivatives.py:521-612 with the kernels in
// cancellation-free form (multiplied by J = z^2 z'^2):
//   K2 J  = 2 A^2 g - 4 h B^2 + h C^2 + 2 L D^2
//   K2s J = 2 A^2 g - 4 h' D^2 + h' C^2 - 2 L B^2
//   Kf J  = (-2 A^2 g + h B^2) Nf / Nc,  Kfs J = (-2 A^2 g + h' D^2) Nf / Nc
// L = ln(X Y' / (X' Y)), g = L coth L - 1, h = 2L / (e^{2L} - 1),
// h' = h e^{2L}, A = z z' / Z^2, B = r z' / (X' Z), C = r^2 / (X X'),
// D = r z / (X Z); dipoles in N = 1 - S form.  Every quantity is a ratio of
// distances, so FP32 neither overflows nor cancels; no FP64 band is needed.
template <typename T>
struct NloStable {
    T total, k2, kf;
};

template <typename T, bool ANYC>
__device__ __forceinline__ NloStable<T> nlo_integrand_stable(T r, T log_r, T lz, T lz2, T tz, T tz2,
                                                             const LdsSpline<T>& sN,
                                                             const NloCoupling<T>& cpl, T nf_over_nc,
                                                             T nc, T norm, bool symmetrize) {
    NloStable<T> out{T(0), T(0), T(0)};
    const T z = exp(lz), z2 = exp(lz2);
    const T c1 = cos(tz), c2 = cos(tz2);
    const T s1 = sin(T(0.5) * tz), s2 = sin(T(0.5) * tz2), s3 = sin(T(0.5) * (tz - tz2));
    const T xsq = (r - z) * (r - z) + T(4) * r * z * s1 * s1;
    const T x2sq = (r - z2) * (r - z2) + T(4) * r * z2 * s2 * s2;
    const T zsq = (z - z2) * (z - z2) + T(4) * z * z2 * s3 * s3;
    const T tiny = sizeof(T) == 4 ? T(1e-37) : T(1e-40);
    if (!(xsq >= tiny) || !(x2sq >= tiny) || !(zsq >= tiny)) return out;
    const T X = sqrt(xsq), X2 = sqrt(x2sq), Z = sqrt(zsq);
    const T lxy = log_x_over_y<T>(r, z, lz, c1, xsq);
    const T lx2y2 = log_x_over_y<T>(r, z2, lz2, c2, x2sq);
    const T log_x = lz + lxy, log_x2 = lz2 + lx2y2, log_zz = T(0.5) * log(zsq);
    // N at the five distances: r_min..r_max, N -> 0 below and 1 above.
    const T n_x = loglog_from_log(sN, log_x, T(0), T(1));
    const T n_y = loglog_from_log(sN, lz, T(0), T(1));
    const T n_x2 = loglog_from_log(sN, log_x2, T(0), T(1));
    const T n_y2 = loglog_from_log(sN, lz2, T(0), T(1));
    const T n_z = loglog_from_log(sN, log_zz, T(0), T(1));
    const T sx = T(1) - n_x, sx2 = T(1) - n_x2;
    T dip = sx * (n_z + n_y2 - n_y - n_z * n_y2);
    T dips = sx2 * (n_z + n_y - n_y2 - n_z * n_y);

    const T L = lxy - lx2y2;
    const T g = coth_minus_one<T>(L);
    const T e2l = exp(T(2) * L);
    const T h = L == T(0) ? T(1) : T(2) * L / expm1(T(2) * L);
    const T hm = h * e2l;
    const T A = z * z2 / zsq, B = r * z2 / (X2 * Z), C = r * r / (X * X2), D = r * z / (X * Z);
    const T A2g = T(2) * A * A * g;
    T k = A2g - T(4) * h * B * B + h * C * C + T(2) * L * D * D;
    T ks = A2g - T(4) * hm * D * D + hm * C * C - T(2) * L * B * B;
    // cut of S/bk/derivatives.py:589-594: |K2| > 1e10 with |dipole| < 1e-10
    const T cut = T(1e10) * exp(T(2) * (lz + lz2));
    if (fabs(k) > cut && fabs(dip) < T(1e-10)) { k = T(0); dip = T(0); }
    if (fabs(ks) > cut && fabs(dips) < T(1e-10)) { ks = T(0); dips = T(0); }
    T k2r = symmetrize ? (k * dip + ks * dips) * T(0.5) : k * dip;
    T kfr = T(0);
    if (nf_over_nc > T(0)) {
        const T kf = (-A2g + h * B * B) * nf_over_nc;
        const T kfs = (-A2g + hm * D * D) * nf_over_nc;
        const T df = (T(1) - n_y) * (n_x - n_x2);
        const T dfs = (T(1) - n_y2) * (n_x2 - n_x);
        kfr = -(symmetrize ? (kf * df + kfs * dfs) * T(0.5) : kf * df);
    }
    const T log_min = fmin(fmin(fmin(log_r, log_x), lz), fmin(fmin(log_x2, lz2), log_zz));
    const T as = nlo_alpha_from_log<ANYC>(log_min, cpl);
    const T scale = (as * nc) * (as * nc) * norm;
    k2r *= scale;
    kfr *= scale;
    out.k2 = isfinite(k2r) ? k2r : T(0);
    out.kf = isfinite(kfr) ? kfr : T(0);
    out.total = out.k2 + out.kf;
    return out;
}

// ------------------------------------------- FP32 path: fast primitives
//
// The FP32 integrand works in base-2 logarithms: the VEGAS map hands it
// t = log2 z (its bin table is scaled by log2 e), so z = 2^t and every
// logarithm of a distance comes from the hardware log2 with no conversion;
// the spline coordinate, the kernel's L and the coupling take their constants
// in the same units.

constexpr float kLn2f = 0.693147182f;
constexpr float kLog2ef = 1.44269502f;

__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }

__device__ __forceinline__ float flog2(float x) { return __builtin_amdgcn_logf(x); }

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// |sin y| for y in [-pi, pi] (the integrand needs only squares of sines):
// fold to v = min(|y|, pi - |y|), with pi - |y| computed exactly for
// |y| >= pi/2 (Sterbenz) plus the low part of pi, then the Taylor polynomial
// to v^11 (relative error < 6e-8, accurate for small v).
__device__ __forceinline__ float fsin_abs(float y) {
    const float ay = fabsf(y);
    const float v = fminf(ay, (3.14159274f - ay) + -8.74227766e-8f);
    const float v2 = v * v;
    float p = fmaf(v2, -2.50521084e-8f, 2.75573192e-6f);
    p = fmaf(v2, p, -1.98412698e-4f);
    p = fmaf(v2, p, 8.33333333e-3f);
    p = fmaf(v2, p, -1.66666667e-1f);
    return fmaf(v2 * v, p, v);
}

// log2(1 + u) / 2 for |u| < 0.5: atanh(s) log2(e) with s = u / (2 + u),
// |s| < 0.2, the series to s^9 with log2(e) folded into its coefficients.
__device__ __forceinline__ float fhalf_log2_1p_small(float u) {
    const float s = u * frcp(2.0f + u);
    const float s2 = s * s;
    float p = fmaf(s2, 0.160299450f, 0.206099287f);  // log2(e) / 9, / 7
    p = fmaf(s2, p, 0.288539022f);                  // log2(e) / 5
    p = fmaf(s2, p, 0.480898350f);                  // log2(e) / 3
    p = fmaf(s2, p, 1.44269502f);                   // log2(e)
    return s * p;
}

// h(L) = 2L / (e^{2L} - 1) and h(-L) = h(L) e^{2L} for x = 2L given also in
// base 2 (x2 = 2L log2 e, so e^{2L} = 2^x2 with no rounding of the argument):
// Bernoulli series 1 - x/2 + x^2/12 - x^4/720 + x^6/30240 for |x| < 0.25,
// else from e^{2L} (no cancellation there; the second form avoids 1 - e^x
// for x << 0).  x2 is clamped at 124 (e^86).
__device__ __forceinline__ void fh_pair(float x, float x2, float& h, float& hm) {
    if (fabsf(x) < 0.25f) {
        const float xx = x * x;
        float p = fmaf(xx, 1.0f / 30240.0f, -1.0f / 720.0f);
        p = fmaf(xx, p, 1.0f / 12.0f);
        h = fmaf(xx, p, fmaf(x, -0.5f, 1.0f));
        hm = h + x;
    } else {
        const float E = fexp2(fminf(x2, 124.0f));
        h = x * frcp(E - 1.0f);
        hm = h * E;
    }
}

// FP32 spline table: per interval the cubic in the local coordinate
// t = (ln d - ln r_0) / h - k in [0, 1), coefficients pre-multiplied by
// log2(e) so that N = 2^(a + t (b + t (c + t d))): one 16-byte LDS read.
struct FastSpline {
    // rec[0] and rec[R] are constant records for ln d below ln r_0 (N = 2^-126,
    // i.e. 0 to FP32 precision, S = 1) and above ln r_{R-1} (N = 1, S = 0);
    // rec[k], k = 1..R-1, is interval k - 1.  u = log2 d * inv_h2 + off is the
    // position in records: inv_h2 = ln 2 / h, off = 1 - ln r_0 / h.
    const float4* rec;
    float inv_h2, off, umax;  // umax = R
};

// u clamped to [0, umax] by one median-of-three (log2_d is finite: it is the
// log of a distance); for u >= 0 the fraction u - floor(u) is exact and
// (int)u = floor(u), so t and the record index need no floor.
__device__ __forceinline__ float fspline_log2n(const FastSpline& s, float log2_d) {
    const float u = __builtin_amdgcn_fmed3f(fmaf(log2_d, s.inv_h2, s.off), 0.0f, s.umax);
    const float t = __builtin_amdgcn_fractf(u);
    const float4 c = s.rec[(int)u];
    return fmaf(t, fmaf(t, fmaf(t, c.w, c.z), c.y), c.x);
}

__device__ __forceinline__ float fspline_n(const FastSpline& s, float log2_d) {
    return fexp2(fspline_log2n(s, log2_d));
}

// N and S = 1 - N.  Near N = 1 the subtraction would keep only the absolute
// precision of N, so S = -expm1(p ln 2) comes from the exponent p = log2 N
// by its series when |p ln 2| < 1/4 (relative error < 2e-8), written in p:
// expm1(p ln 2) = sum_k (ln 2)^k p^k / k!.
__device__ __forceinline__ void fspline_ns(const FastSpline& s, float log2_d, float& n, float& sv) {
    const float p = fspline_log2n(s, log2_d);
    n = fexp2(p);
    float e = fmaf(p, 1.54035304e-4f, 1.33335581e-3f);  // (ln 2)^6 / 720, (ln 2)^5 / 120
    e = fmaf(p, e, 9.61812911e-3f);                     // (ln 2)^4 / 24
    e = fmaf(p, e, 5.55041087e-2f);                     // (ln 2)^3 / 6
    e = fmaf(p, e, 2.40226507e-1f);                     // (ln 2)^2 / 2
    e = fmaf(p, e, 6.93147181e-1f);                     // ln 2
    sv = fabsf(p) < 0.360673760f ? -(p * e) : 1.0f - n;  // |p ln 2| < 1/4
}

// The FP32 K2 + Kf integrand: the cancellation-free formula of
// nlo_integrand_stable, every transcendental on the fast primitives above,
// cos(theta) = 1 - 2 sin^2(theta / 2), logs of distances (base 2) from the
// log1p forms, N from the local-coordinate spline.  t1 = log2 z and
// t2 = log2 z', log2_r = log2 r.
template <bool ANYC>
__device__ __forceinline__ NloStable<float> nlo_integrand_fp32(
    float r, float log2_r, float t1, float t2, float tz, float tz2, const FastSpline& sp,
    const NloCoupling<float>& cpl, float nf_over_nc, float nc, float norm, bool symmetrize) {
    NloStable<float> out{0.0f, 0.0f, 0.0f};
    const float z = fexp2(t1), z2 = fexp2(t2);
    const float s1 = fsin_abs(0.5f * tz), s2 = fsin_abs(0.5f * tz2);
    const float s3 = fsin_abs(0.5f * (tz - tz2));
    const float xsq = fmaf(4.0f * r * z, s1 * s1, (r - z) * (r - z));
    const float x2sq = fmaf(4.0f * r * z2, s2 * s2, (r - z2) * (r - z2));
    const float zsq = fmaf(4.0f * z * z2, s3 * s3, (z - z2) * (z - z2));
    if (!(xsq >= 1e-37f) || !(x2sq >= 1e-37f) || !(zsq >= 1e-37f)) return out;
    // the distances enter only as reciprocals: 1 / X = rsq(X^2)
    const float iX = frsq(xsq), iX2 = frsq(x2sq), iZ = frsq(zsq);
    // log2(X / Y): (X^2 - z^2) / z^2 = rho (rho - 2 cos t), cos t = 1 - 2 s^2
    const float rho1 = r * frcp(z), rho2 = r * frcp(z2);
    const float u1 = rho1 * (rho1 - 2.0f + 4.0f * s1 * s1);
    const float u2 = rho2 * (rho2 - 2.0f + 4.0f * s2 * s2);
    const float lxy = fabsf(u1) < 0.5f ? fhalf_log2_1p_small(u1) : 0.5f * flog2(xsq) - t1;
    const float lx2y2 = fabsf(u2) < 0.5f ? fhalf_log2_1p_small(u2) : 0.5f * flog2(x2sq) - t2;
    const float log_x = t1 + lxy, log_x2 = t2 + lx2y2, log_zz = 0.5f * flog2(zsq);
    float n_x, s_x, n_y, s_y, n_x2, s_x2, n_y2, s_y2;
    fspline_ns(sp, log_x, n_x, s_x);
    fspline_ns(sp, t1, n_y, s_y);
    fspline_ns(sp, log_x2, n_x2, s_x2);
    fspline_ns(sp, t2, n_y2, s_y2);
    const float n_z = fspline_n(sp, log_zz);
    float dip = s_x * (n_z + n_y2 - n_y - n_z * n_y2);
    float dips = s_x2 * (n_z + n_y - n_y2 - n_z * n_y);

    // L = ln(X Y' / (X' Y)) from its base-2 form
    const float L2 = lxy - lx2y2;
    const float L = L2 * kLn2f;
    float h, hm;
    fh_pair(2.0f * L, 2.0f * L2, h, hm);
    const float aL = fabsf(L);
    float g;
    if (aL < 1.0f) {
        // L coth L - 1 = sum_n 2^2n B_2n L^2n / (2n)!: seven terms reach the
        // rounding of y = L^2 on |L| < 1 (the terms fall by ~1/pi^2)
        const float y = L * L;
        float p = fmaf(y, 2.19259479e-7f, -2.16440428e-6f);
        p = fmaf(y, p, 2.13777992e-5f);
        p = fmaf(y, p, -2.11640212e-4f);
        p = fmaf(y, p, 2.11640212e-3f);
        p = fmaf(y, p, -2.22222222e-2f);
        p = fmaf(y, p, 3.33333333e-1f);
        g = y * p;
    } else {
        g = aL + (L >= 0.0f ? h : hm) - 1.0f;  // |L| + h(|L|) - 1
    }
    const float A = z * z2 * (iZ * iZ);
    const float rZ = r * iZ;
    const float B = rZ * z2 * iX2, C = r * r * (iX * iX2), D = rZ * z * iX;
    const float B2 = B * B, C2 = C * C, D2 = D * D;
    const float A2g = 2.0f * A * A * g;
    float k = A2g - 4.0f * h * B2 + h * C2 + 2.0f * L * D2;
    float ks = A2g - 4.0f * hm * D2 + hm * C2 - 2.0f * L * B2;
    const float zz2 = z * z2;
    const float cut = 1e10f * (zz2 * zz2);  // |K2| > 1e10 (S/bk/derivatives.py:589), J = z^2 z'^2
    if (fabsf(k) > cut && fabsf(dip) < 1e-10f) { k = 0.0f; dip = 0.0f; }
    if (fabsf(ks) > cut && fabsf(dips) < 1e-10f) { ks = 0.0f; dips = 0.0f; }
    float k2r = symmetrize ? (k * dip + ks * dips) * 0.5f : k * dip;
    float kfr = 0.0f;
    if (nf_over_nc > 0.0f) {
        const float kf = (h * B2 - A2g) * nf_over_nc;
        const float kfs = (hm * D2 - A2g) * nf_over_nc;
        const float df = s_y * (n_x - n_x2);
        const float dfs = s_y2 * (n_x2 - n_x);
        kfr = -(symmetrize ? (kf * df + kfs * dfs) * 0.5f : kf * df);
    }
    // smallest distance, base 2
    const float log_min = fminf(fminf(fminf(log2_r, log_x), t1), fminf(fminf(log_x2, t2), log_zz));
    float as;
    if (!ANYC || cpl.mode == 0) {  // smooth coupling at the smallest distance (production)
        // in base 2: right = 5 (ln s0 - 2 ln d) log2 e, mu = ln mu0 log2 e,
        // alpha = 1 / (b0 0.2 ln 2 (max(mu, right) + log2(1 + 2^-|mu - right|)))
        const float right = fmaf(-10.0f, log_min, cpl.log_scale0 * (5.0f * kLog2ef));
        const float mu = cpl.log_mu0 * kLog2ef;
        const float hi = fmaxf(mu, right);
        // 2^-|d| <= 1 is added to 1: the plain log2 is accurate enough here
        const float t = fexp2(-fabsf(mu - right));
        as = frcp((cpl.b0 * (0.2f * kLn2f)) * (hi + flog2(1.0f + t)));
    } else if (cpl.mode == 1) {
        as = cpl.as_const;
    } else {
        as = alpha_thresholds_log<float>(log_min * kLn2f, cpl.log_4c2);
    }
    const float scale = (as * nc) * (as * nc) * norm;
    k2r *= scale;
    kfr *= scale;
    out.k2 = isfinite(k2r) ? k2r : 0.0f;
    out.kf = isfinite(kfr) ? kfr : 0.0f;
    out.total = out.k2 + out.kf;
    return out;
}

}  // namespace nlo
