/*
 * FP64 CPU oracle for the NLO BK right-hand side (test infrastructure only;
 * see nlo_oracle.h).  Each function names the reference lines it restates.
 * S/ = torch_nlo_clean/src/nlo_torch/, K/ = S/custom_kernels/.
 */
#define _DEFAULT_SOURCE 1 /* j1() from the C library */
#include "nlo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------- coupling */

/* _alpha_s (S/bk_kernels/k1.py:158-169, S/bk/derivatives.py:775-786) and
 * bk_alpha_s (S/coupling.py:11-46): FIXED_AS when either coupling choice is
 * fixed; for NF <= 3 the smooth coupling of S/coupling.py:81-97, alpha =
 * 1 / (b0 * 0.2 * logaddexp(2/0.2 ln 2.5, ln(4 C2 / (r^2 Lambda^2)) / 0.2));
 * for NF > 3 one-loop running with the flavour number and Lambda switched at
 * the scales 1.3 and 4.5 (4 C2 / r^2 in GeV^2), capped at 1. */
double orc_alpha_s(double r, const orc_bk_params* p) {
    if (p->rc_lo == ORC_RC_LO_FIXED || p->rc_nlo == ORC_RC_NLO_FIXED) return p->fixed_as;
    if (p->nf <= 3.0) {
        const double b0 = (11.0 * p->nc - 2.0 * p->nf) / (12.0 * M_PI);
        const double left = 2.0 / 0.2 * log(2.5);
        const double scale = 4.0 * p->c2 / (r * r * p->lambda_qcd * p->lambda_qcd);
        const double right = log(scale) / 0.2;
        const double hi = left > right ? left : right;
        const double lae = hi + log1p(exp(-fabs(left - right)));
        return 1.0 / (b0 * (0.2 * lae));
    }
    const double dipolescale = 4.0 * p->c2 / (r * r);
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
    const double log_arg = 4.0 * p->c2 / (r * r * (lambda * lambda));
    const double a = 4.0 * M_PI / (b0 * log(log_arg));
    return (log_arg < 1.0 || a > 1.0) ? 1.0 : a;
}

/* S/bk/initial_conditions.py:29-35. The constant e is truncated to 2.7182818. */
double orc_mv(double r, double qs0sqr, double gamma, double ec, double lambda_qcd) {
    const double exponent =
        pow(r * r * qs0sqr, gamma) / 4.0 * log(1.0 / (r * lambda_qcd) + ec * 2.7182818);
    return exponent < 1e-5 ? exponent : 1.0 - exp(-exponent);
}

/* S/bk/evolution.py:584-592: r_i = MINR * step^i, step = (MAXR/MINR)^(1/RPOINTS). */
void orc_r_grid(double minr, double maxr, int rpoints, double* out) {
    const double step = pow(maxr / minr, 1.0 / (double)rpoints);
    for (int i = 0; i <= rpoints; ++i) out[i] = minr * pow(step, (double)i);
}

/* ----------------------------------------------------------------- splines */

/* S/numerics/interpolation.py:27-48 and :124-149 (natural spline, Thomas). */
void orc_natural_spline(int n, const double* x, const double* y, double* a, double* b,
                        double* c, double* d) {
    const int m = n - 2; /* interior second derivatives */
    double* h = (double*)malloc(sizeof(double) * (size_t)(n - 1));
    double* sec = (double*)calloc((size_t)n, sizeof(double));
    double* upper = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double* sol = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    for (int i = 0; i < n - 1; ++i) h[i] = x[i + 1] - x[i];
    if (m == 1) {
        const double diag = 2.0 * (h[0] + h[1]);
        const double rhs = 6.0 * ((y[2] - y[1]) / h[1] - (y[1] - y[0]) / h[0]);
        sol[0] = rhs / diag;
    } else if (m > 1) {
        /* row i: off-diagonal h[i] couples unknowns i-1, i (interior index). */
        upper[0] = h[1] / (2.0 * (h[0] + h[1]));
        sol[0] = 6.0 * ((y[2] - y[1]) / h[1] - (y[1] - y[0]) / h[0]) / (2.0 * (h[0] + h[1]));
        for (int i = 1; i < m; ++i) {
            const double diag = 2.0 * (h[i] + h[i + 1]);
            const double off = h[i];
            const double rhs =
                6.0 * ((y[i + 2] - y[i + 1]) / h[i + 1] - (y[i + 1] - y[i]) / h[i]);
            const double den = diag - off * upper[i - 1];
            if (i < m - 1) upper[i] = h[i + 1] / den;
            sol[i] = (rhs - off * sol[i - 1]) / den;
        }
        for (int i = m - 2; i >= 0; --i) sol[i] = sol[i] - upper[i] * sol[i + 1];
    }
    for (int i = 0; i < m; ++i) sec[i + 1] = sol[i];
    for (int i = 0; i < n - 1; ++i) {
        a[i] = y[i];
        b[i] = (y[i + 1] - y[i]) / h[i] - h[i] * (2.0 * sec[i] + sec[i + 1]) / 6.0;
        c[i] = sec[i] / 2.0;
        d[i] = (sec[i + 1] - sec[i]) / (6.0 * h[i]);
    }
    free(h);
    free(sec);
    free(upper);
    free(sol);
}

/* S/bk/evolution.py:548-560 (clamps) + S/numerics/interpolation.py:98-103
 * (log y with the -40 floor). */
void orc_evolution_splines(int n, const double* r, const double* N, int force_positive,
                           double* x_out, double* coef_n, double* coef_s) {
    double* ly = (double*)malloc(sizeof(double) * (size_t)n);
    const int m = n - 1;
    for (int i = 0; i < n; ++i) x_out[i] = log(r[i]);
    for (int i = 0; i < n; ++i) {
        double v = N[i] > 1.0 ? 1.0 : N[i];
        if (force_positive && v < 0.0) v = 0.0;
        ly[i] = v > 0.0 ? log(v) : -40.0;
    }
    orc_natural_spline(n, x_out, ly, coef_n, coef_n + m, coef_n + 2 * m, coef_n + 3 * m);
    for (int i = 0; i < n; ++i) {
        double s = 1.0 - N[i];
        if (s < 0.0) s = 0.0;
        if (force_positive && s > 1.0) s = 1.0;
        ly[i] = s > 0.0 ? log(s) : -40.0;
    }
    orc_natural_spline(n, x_out, ly, coef_s, coef_s + m, coef_s + 2 * m, coef_s + 3 * m);
    free(ly);
}

/* S/numerics/interpolation.py:54-60: interval = searchsorted(x, v, right) - 1,
 * clamped to [0, n - 2]. */
double orc_spline_eval(const orc_spline* s, double v) {
    int lo = 0, hi = s->n;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (v >= s->x[mid]) lo = mid + 1;
        else hi = mid;
    }
    int k = lo - 1;
    if (k < 0) k = 0;
    if (k > s->n - 2) k = s->n - 2;
    const double dx = v - s->x[k];
    return s->a[k] + dx * (s->b[k] + dx * (s->c[k] + dx * s->d[k]));
}

/* S/bk/derivatives.py:761-772 (_evaluate_loglog). */
double orc_loglog_eval(const orc_spline* s, double value, double underflow, double overflow) {
    double ev = value < s->r_min ? s->r_min : value;
    ev = ev > s->r_max ? s->r_max : ev;
    double out = exp(orc_spline_eval(s, log(ev)));
    if (!isfinite(out)) out = 0.0;
    if (value < s->r_min) out = underflow;
    if (value > s->r_max) out = overflow;
    return out;
}

/* ------------------------------------------------------------------ Bessel */

/* J1 from the C library (glibc j1). */
double orc_bessel_j1(double x) { return j1(x); }

/* I1 by its power series sum_k (x/2)^(2k+1) / (k! (k+1)!), all terms positive,
 * accumulated in long double (no cancellation). */
double orc_bessel_i1(double x) {
    const long double h = (long double)fabs(x) / 2.0L;
    const long double h2 = h * h;
    long double term = h, sum = h;
    for (int k = 1; k < 400; ++k) {
        term *= h2 / ((long double)k * (long double)(k + 1));
        sum += term;
        if (term < sum * 1e-22L) break;
    }
    return x < 0 ? -(double)sum : (double)sum;
}

/* ---------------------------------------------------------------- K1 (LO) */

static double sqr(double v) { return v * v; }

static double dist_sq(double r, double z, double theta, int cos_form) {
    if (cos_form) return r * r + z * z - 2.0 * r * z * cos(theta);
    const double s = sin(theta / 2.0);
    return (r - z) * (r - z) + 4.0 * r * z * s * s;
}

static int order_has_resummation(int order) {
    return order != ORC_ORDER_LO && order != ORC_ORDER_NLO;
}

static int order_resums_slog(int order) {
    return order == ORC_ORDER_NLO_RESUM_DLOG_SLOG || order == ORC_ORDER_LO_RESUM_DLOG_SLOG;
}

/* Kernel_lo, S/bk_kernels/k1.py:20-155, for every RC_LO / RESUM_RC / Order
 * choice.  The k1fin double-log term carries dlog = 0 for every order that
 * reaches it (all of them resum the double logs). */
static double kernel_lo_impl(double r, double z, double theta, const orc_bk_params* p,
                             double* magnitude) {
    const double Y = z;
    const double Xsqr = dist_sq(r, z, theta, p->geometry_cos);
    int invalid = Xsqr <= 0.0;
    const double X = sqrt(Xsqr > 0.0 ? Xsqr : 0.0);
    invalid = invalid || X < 1e-20 || Y < 1e-20;
    const double sX = invalid ? 1.0 : X;
    const double sY = invalid ? 1.0 : Y;
    double mn = sX < sY ? sX : sY;
    mn = mn < r ? mn : r;
    const double a_r = orc_alpha_s(r, p);
    const double pref = p->nc / (2.0 * M_PI * M_PI);
    const double rxy2 = sqr(r / (sX * sY));
    double lo, lo_abs, a_scale;
    switch (p->rc_lo) {
        case ORC_RC_LO_SMALLEST:
            a_scale = orc_alpha_s(mn, p);
            lo = p->nc * a_scale / (2.0 * M_PI * M_PI) * rxy2;
            lo_abs = fabs(lo);
            break;
        case ORC_RC_LO_PARENT:
            a_scale = a_r;
            lo = p->nc * a_r / (2.0 * M_PI * M_PI) * rxy2;
            lo_abs = fabs(lo);
            break;
        case ORC_RC_LO_FAC: {
            const double ab_r = a_r * p->nc / M_PI;
            const double ab_X = orc_alpha_s(sX, p) * p->nc / M_PI;
            const double ab_Y = orc_alpha_s(sY, p) * p->nc / M_PI;
            a_scale = a_r;
            const double den =
                1.0 / ab_r + (sX * sX - sY * sY) / (r * r) * (ab_X - ab_Y) / (ab_X * ab_Y);
            lo = 1.0 / (2.0 * M_PI) / den * rxy2;
            /* rounding scale: the denominator's terms in absolute value, the
             * differences X^2 - Y^2 and abar_X - abar_Y by their sums */
            lo_abs = fabs(lo) / fabs(den) *
                     (1.0 / ab_r + (sX * sX + sY * sY) / (r * r) * (ab_X + ab_Y) / (ab_X * ab_Y));
            break;
        }
        case ORC_RC_LO_BEUF: {
            const double expo = (sX * sX - sY * sY) / (r * r), base = sY * sY / (sX * sX);
            const double reff2 = r * r * pow(base, expo);
            a_scale = orc_alpha_s(sqrt(reff2), p);
            lo = p->nc * a_scale / (2.0 * M_PI * M_PI) * sqr(r / (sX * sY + 1e-40));
            /* rounding scale: ln r_eff^2 = ln r^2 + expo ln base carries the
             * rounding of X^2 times |expo| and that of X^2 - Y^2 times |ln base|
             * (relative condition numbers; the coupling's log-derivative < 1) */
            const double kx = (r * r + z * z + 2.0 * r * z) / (sX * sX);
            lo_abs = fabs(lo) * (1.0 + 2.0 * fabs(expo) * kx +
                                 fabs(log(base)) * (sX * sX + sY * sY) / (r * r));
            break;
        }
        default: { /* BALITSKY_LO, FIXED_LO */
            const double a_Y = orc_alpha_s(sY, p);
            const double a_X = orc_alpha_s(sX, p);
            a_scale = a_r;
            lo = pref * a_r *
                 (r * r / (sX * sX * sY * sY) + (a_Y / a_X - 1.0) / (sY * sY) +
                  (a_X / a_Y - 1.0) / (sX * sX));
            lo_abs = pref * a_r *
                     (r * r / (sX * sX * sY * sY) + (a_Y / a_X + 1.0) / (sY * sY) +
                      (a_X / a_Y + 1.0) / (sX * sX));
            break;
        }
    }
    if (!isfinite(lo)) lo = 0.0;
    if (invalid) lo = 0.0;
    if (!order_has_resummation(p->order)) {
        if (magnitude) *magnitude = invalid ? 0.0 : (isfinite(lo_abs) ? lo_abs : 0.0);
        return lo;
    }

    double a_res;
    if (p->resum_rc == ORC_RESUM_RC_PARENT)
        a_res = a_r;
    else if (p->resum_rc == ORC_RESUM_RC_FIXED)
        a_res = p->fixed_as;
    else
        a_res = orc_alpha_s(mn, p);
    const double xl = 4.0 * log(sX / r) * log(sY / r);
    const double as_x = sqrt(a_res * p->nc / M_PI * fabs(xl));
    const double safe = as_x != 0.0 ? as_x : 1.0;
    double dlog = xl >= 0.0 ? orc_bessel_j1(2.0 * as_x) / safe : orc_bessel_i1(2.0 * as_x) / safe;
    if (as_x == 0.0) dlog = 1.0;
    if (!(r > 1.01 * p->minr)) dlog = 1.0;
    const int resum_valid = isfinite(dlog);

    double slog = 1.0, slog_exp = 0.0;
    if (order_resums_slog(p->order)) {
        const double minxy = sX < sY ? sX : sY;
        const double abar = a_res * p->nc / M_PI;
        const double A1 = 11.0 / 12.0;
        slog = exp(-abar * A1 * fabs(log(p->ksub * sqr(r / minxy))));
        slog_exp = -abar * A1 * fabs(2.0 * log(sqrt(p->ksub) * r / minxy));
    }
    const double lo_single = a_scale * p->nc / (2.0 * M_PI * M_PI) * rxy2;
    const double subtract = lo_single * slog_exp;
    const double fin = a_scale * p->nc / (4.0 * M_PI) *
                       (67.0 / 9.0 - M_PI * M_PI / 3.0 - 10.0 / 9.0 * p->nf / p->nc);
    const double k1fin = lo * fin;
    const double result = dlog * slog * lo - subtract + k1fin;
    if (magnitude) {
        *magnitude = (invalid || !resum_valid)
                         ? 0.0
                         : fabs(dlog * slog) * lo_abs + fabs(subtract) + fabs(k1fin) +
                               fabs(lo_abs * fin);
    }
    return (invalid || !resum_valid) ? 0.0 : result;
}

double orc_kernel_lo(double r, double z, double theta, const orc_bk_params* p) {
    return kernel_lo_impl(r, z, theta, p, NULL);
}

/* Magnitude of Kernel_lo with every additive piece taken in absolute value
 * (used by tests to scale tolerances where the reference formula cancels). */
double orc_kernel_lo_magnitude(double r, double z, double theta, const orc_bk_params* p) {
    double m = 0.0;
    (void)kernel_lo_impl(r, z, theta, p, &m);
    return m;
}

/* S/bk/derivatives.py:434-468 (_lo_integrand without kinematical constraint). */
double orc_k1_integrand(double r, double z, double theta, const orc_spline* sN,
                        const orc_bk_params* p) {
    const double Xsqr = dist_sq(r, z, theta, p->geometry_cos);
    const int invalid = Xsqr < 1e-40 || z < 1e-20 || r < 1e-20;
    const double X = sqrt(Xsqr > 0.0 ? Xsqr : 0.0);
    const double sX = invalid ? 1.0 : X;
    const double sY = invalid ? 1.0 : z;
    double nX = orc_loglog_eval(sN, sX, 0.0, 1.0);
    double nY = orc_loglog_eval(sN, sY, 0.0, 1.0);
    double nr = orc_loglog_eval(sN, r, 0.0, 1.0);
    nX = nX < 1.0 ? nX : 1.0;
    nY = nY < 1.0 ? nY : 1.0;
    nr = nr < 1.0 ? nr : 1.0;
    const double dip = nX + nY - nr - nX * nY;
    const double v = orc_kernel_lo(r, z, theta, p) * dip;
    return invalid ? 0.0 : v;
}

double orc_k1_integrand_magnitude(double r, double z, double theta, const orc_spline* sN,
                                  const orc_bk_params* p) {
    const double Xsqr = dist_sq(r, z, theta, p->geometry_cos);
    const int invalid = Xsqr < 1e-40 || z < 1e-20 || r < 1e-20;
    if (invalid) return 0.0;
    const double X = sqrt(Xsqr > 0.0 ? Xsqr : 0.0);
    double nX = orc_loglog_eval(sN, X, 0.0, 1.0);
    double nY = orc_loglog_eval(sN, z, 0.0, 1.0);
    double nr = orc_loglog_eval(sN, r, 0.0, 1.0);
    nX = nX < 1.0 ? nX : 1.0;
    nY = nY < 1.0 ? nY : 1.0;
    nr = nr < 1.0 ? nr : 1.0;
    return orc_kernel_lo_magnitude(r, z, theta, p) * (nX + nY + nr + nX * nY);
}

/* NaturalCubicSpline(x, y)(v) on six nodes (S/numerics/interpolation.py:10-68):
 * second derivatives from the Thomas algorithm, interval by
 * searchsorted(right) - 1 clamped to [0, 4]. */
static double natural6(const double* x, const double* y, double v) {
    double h[5], sec[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, up[4], sol[4];
    for (int i = 0; i < 5; ++i) h[i] = x[i + 1] - x[i];
    for (int i = 0; i < 4; ++i) {
        const double diag = 2.0 * (h[i] + h[i + 1]);
        const double rhs = 6.0 * ((y[i + 2] - y[i + 1]) / h[i + 1] - (y[i + 1] - y[i]) / h[i]);
        if (i == 0) {
            up[0] = h[1] / diag;
            sol[0] = rhs / diag;
        } else {
            const double den = diag - h[i] * up[i - 1];
            up[i] = h[i + 1] / den;
            sol[i] = (rhs - h[i] * sol[i - 1]) / den;
        }
    }
    for (int i = 2; i >= 0; --i) sol[i] -= up[i] * sol[i + 1];
    for (int i = 0; i < 4; ++i) sec[i + 1] = sol[i];
    int k = 0;
    while (k < 5 && x[k + 1] <= v) ++k;
    k = k > 4 ? 4 : k;
    const double b = (y[k + 1] - y[k]) / h[k] - h[k] * (2.0 * sec[k] + sec[k + 1]) / 6.0;
    const double c = sec[k] / 2.0, d = (sec[k + 1] - sec[k]) / (6.0 * h[k]);
    const double dx = v - x[k];
    return y[k] + dx * (b + dx * (c + dx * d));
}

double orc_history_N(double r, double rapidity, const orc_spline* sN, const orc_history* h) {
    if (rapidity < 0.0) rapidity = 0.0;
    if (rapidity > h->y[h->rows - 1]) return NAN;
    int yind = -1; /* searchsorted(history_y, rapidity, right=True) - 1 */
    while (yind + 1 < h->rows && h->y[yind + 1] <= rapidity) ++yind;
    yind = yind < 0 ? 0 : yind;
    double result;
    if (yind == h->rows - 1) {
        result = orc_loglog_eval(sN, r, 0.0, 1.0);
    } else {
        int rind = -1;
        while (rind + 1 < h->n && h->r[rind + 1] <= r) ++rind;
        const int start = rind - 3 > 0 ? rind - 3 : 0;
        if (start + 6 > h->n) {
            result = 1.0;
        } else {
            const double* x = h->r + start;
            double v = r;
            if (v < x[0]) v = x[0] * 1.00001;
            if (v > x[5]) v = x[5] * 0.999999;
            const double lower = natural6(x, h->N + (size_t)yind * h->n + start, v);
            const double upper = natural6(x, h->N + (size_t)(yind + 1) * h->n + start, v);
            const double frac = (rapidity - h->y[yind]) / (h->y[yind + 1] - h->y[yind]);
            result = lower + frac * (upper - lower);
        }
    }
    return result < 0.0 ? 0.0 : (result > 1.0 ? 1.0 : result);
}

double orc_k1_integrand_history(double r, double z, double theta, const orc_spline* sN,
                                const orc_bk_params* p, const orc_history* h) {
    const double Xsqr = dist_sq(r, z, theta, p->geometry_cos);
    const int invalid = Xsqr < 1e-40 || z < 1e-20 || r < 1e-20;
    const double X = sqrt(Xsqr > 0.0 ? Xsqr : 0.0);
    const double sX = invalid ? 1.0 : X;
    const double sY = invalid ? 1.0 : z;
    double nr = orc_loglog_eval(sN, r, 0.0, 1.0);
    nr = nr < 1.0 ? nr : 1.0;
    const double m2 = sX * sX < sY * sY ? sX * sX : sY * sY;
    double delta = log(m2 / (r * r));
    delta = delta > 0.0 ? delta : 0.0;
    const double shifted = h->rapidity - delta;
    if (invalid || shifted < 0.0) return 0.0;
    const double s02 = 1.0 - orc_history_N(sX, shifted, sN, h);
    const double s12 = 1.0 - orc_history_N(sY, shifted, sN, h);
    return orc_kernel_lo(r, z, theta, p) * (-s02 * s12 + (1.0 - nr));
}

/* K/bk/k1/kernel.cu:2040-2092: sum over panels [x_i, x_{i+1}] of the log-r
 * grid, GL(nr) in log z and GL(nt) in v in [0, 1] with theta = pi v^2,
 * integrand * 2 z^2 * (half_width w_r) * (w_t pi v). */
void orc_k1_fixed_history(int parents, const double* parent_r, const orc_spline* sN, int nr,
                          const double* gr_x, const double* gr_w, int nt, const double* gt_x,
                          const double* gt_w, const orc_bk_params* p, const orc_history* h,
                          double* out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int pi = 0; pi < parents; ++pi) {
        const double r = parent_r[pi];
        long double total = 0.0L;
        for (int panel = 0; panel < sN->n - 1; ++panel) {
            const double lower = sN->x[panel], upper = sN->x[panel + 1];
            const double center = (lower + upper) / 2.0, half = (upper - lower) / 2.0;
            for (int i = 0; i < nr; ++i) {
                const double log_z = center + half * gr_x[i];
                const double z = exp(log_z);
                const double wr = half * gr_w[i];
                for (int j = 0; j < nt; ++j) {
                    const double v = (gt_x[j] + 1.0) / 2.0;
                    const double theta = M_PI * v * v;
                    const double wt = gt_w[j] * M_PI * v;
                    const double f = h ? orc_k1_integrand_history(r, z, theta, sN, p, h)
                                       : orc_k1_integrand(r, z, theta, sN, p);
                    total += (long double)(f * (2.0 * z * z) * wr * wt);
                }
            }
        }
        out[pi] = (double)total;
    }
}

void orc_k1_fixed(int parents, const double* parent_r, const orc_spline* sN, int nr,
                  const double* gr_x, const double* gr_w, int nt, const double* gt_x,
                  const double* gt_w, const orc_bk_params* p, double* out) {
    orc_k1_fixed_history(parents, parent_r, sN, nr, gr_x, gr_w, nt, gt_x, gt_w, p, NULL, out);
}

/* ------------------------------------------------ adaptive Gauss-Kronrod */

/* QUADPACK's 21-point Kronrod extension of the 10-point Gauss rule: nodes
 * xgk[0..9] (xgk[1, 3, .., 9] are the Gauss nodes), the centre is node 10. */
static const double kXgk[10] = {
    0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
    0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
    0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
    0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
    0.294392862701460198131126603103866, 0.148874338981631210884826001129720};
static const double kWgk[10] = {
    0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
    0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
    0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
    0.123491976262065851077958109831074, 0.134709217311473325928054001771707,
    0.142775938577060080797094273138717, 0.147739104901338491374841515972068};
static const double kWgkCenter = 0.149445554002916905664936468389821;
static const double kWg[5] = {
    0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
    0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
    0.295524224714752870173892994651338};

void orc_gk21(orc_fn f, void* ctx, double a, double b, double* estimate, double* error) {
    const double center = (a + b) / 2.0, half = (b - a) / 2.0;
    double neg[10], pos[10];
    for (int i = 0; i < 10; ++i) {
        neg[i] = f(center - half * kXgk[i], ctx);
    }
    const double fc = f(center, ctx);
    for (int i = 0; i < 10; ++i) {
        pos[i] = f(center + half * kXgk[i], ctx);
    }
    double kron = 0.0, gauss = 0.0;
    for (int i = 0; i < 10; ++i) kron += (neg[i] + pos[i]) * kWgk[i];
    kron += kWgkCenter * fc;
    for (int i = 0; i < 5; ++i) gauss += (neg[2 * i + 1] + pos[2 * i + 1]) * kWg[i];
    const double mean = kron / 2.0;
    double sabs = 0.0, sasc = 0.0;
    for (int i = 0; i < 10; ++i) {
        sabs += (fabs(neg[i]) + fabs(pos[i])) * kWgk[i];
        sasc += (fabs(neg[i] - mean) + fabs(pos[i] - mean)) * kWgk[i];
    }
    const double resabs = half * (sabs + kWgkCenter * fabs(fc));
    const double resasc = half * (sasc + kWgkCenter * fabs(fc - mean));
    double e = half * fabs(kron - gauss);
    const double scale = pow(200.0 * e / (resasc > 0.0 ? resasc : 1.0), 1.5);
    if (resasc > 0.0 && e > 0.0) e = resasc * (scale < 1.0 ? scale : 1.0);
    const double floor_e = 50.0 * 2.220446049250313e-16 * resabs;
    *estimate = half * kron;
    *error = e > floor_e ? e : floor_e;
}

double orc_adaptive_gk21(orc_fn f, void* ctx, double a, double b, double epsrel, int max_intervals,
                         double* error, int* converged, long* evaluations) {
    double* lo = (double*)malloc(sizeof(double) * 4 * (size_t)(max_intervals + 1));
    double* hi = lo + max_intervals + 1;
    double* est = hi + max_intervals + 1;
    double* err = est + max_intervals + 1;
    int count = 1;
    long nodes = 21;
    lo[0] = a;
    hi[0] = b;
    orc_gk21(f, ctx, a, b, &est[0], &err[0]);
    double value = 0.0, total = 0.0;
    int done = 0;
    while (count < max_intervals) {
        value = 0.0;
        total = 0.0;
        for (int i = 0; i < count; ++i) {
            value += est[i];
            total += err[i];
        }
        if (total <= epsrel * fabs(value)) {
            done = 1;
            break;
        }
        int split = 0;
        for (int i = 1; i < count; ++i)
            if (err[i] > err[split]) split = i;
        const double l = lo[split], u = hi[split], mid = (l + u) / 2.0;
        double e0, r0, e1, r1;
        orc_gk21(f, ctx, l, mid, &e0, &r0);
        orc_gk21(f, ctx, mid, u, &e1, &r1);
        for (int i = count; i > split + 1; --i) {
            lo[i] = lo[i - 1];
            hi[i] = hi[i - 1];
            est[i] = est[i - 1];
            err[i] = err[i - 1];
        }
        hi[split] = mid;
        est[split] = e0;
        err[split] = r0;
        lo[split + 1] = mid;
        hi[split + 1] = u;
        est[split + 1] = e1;
        err[split + 1] = r1;
        ++count;
        nodes += 42;
    }
    if (!done) {
        value = 0.0;
        total = 0.0;
        for (int i = 0; i < count; ++i) {
            value += est[i];
            total += err[i];
        }
        done = total <= epsrel * fabs(value);
    }
    free(lo);
    if (error) *error = total;
    if (converged) *converged = done;
    if (evaluations) *evaluations = nodes;
    return value;
}

typedef struct {
    double r, z;
    const orc_spline* sN;
    const orc_bk_params* p;
    const orc_history* h;
    int tint;
    double epsrel;
    long evals;
} k1_ctx;

static double k1_theta_fn(double theta, void* ctx) {
    const k1_ctx* c = (const k1_ctx*)ctx;
    return c->h ? orc_k1_integrand_history(c->r, c->z, theta, c->sN, c->p, c->h)
                : orc_k1_integrand(c->r, c->z, theta, c->sN, c->p);
}

/* integrate_theta (S/bk/derivatives.py:61-101): 2 z^2 x the theta integral. */
static double k1_logz_fn(double log_z, void* ctx) {
    k1_ctx* c = (k1_ctx*)ctx;
    c->z = exp(log_z);
    long n = 0;
    const double v = orc_adaptive_gk21(k1_theta_fn, c, 0.0, M_PI, c->epsrel, c->tint, NULL, NULL, &n);
    c->evals += n;
    return v * exp(2.0 * log_z) * 2.0;
}

void orc_k1_adaptive(int parents, const double* parent_r, const orc_spline* sN, int rint, int tint,
                     double epsrel, const orc_bk_params* p, const orc_history* h, double* out,
                     double* err, long* evals) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int pi = 0; pi < parents; ++pi) {
        k1_ctx c = {parent_r[pi], 0.0, sN, p, h, tint, epsrel, 0};
        double e = 0.0;
        out[pi] = orc_adaptive_gk21(k1_logz_fn, &c, sN->x[0], sN->x[sN->n - 1], epsrel, rint, &e,
                                    NULL, NULL);
        if (err) err[pi] = e;
        if (evals) evals[pi] = c.evals;
    }
}

/* Same rule over |integrand| magnitudes: the rounding scale of K1. */
void orc_k1_fixed_magnitude(int parents, const double* parent_r, const orc_spline* sN, int nr,
                            const double* gr_x, const double* gr_w, int nt, const double* gt_x,
                            const double* gt_w, const orc_bk_params* p, double* out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int pi = 0; pi < parents; ++pi) {
        const double r = parent_r[pi];
        long double total = 0.0L;
        for (int panel = 0; panel < sN->n - 1; ++panel) {
            const double lower = sN->x[panel], upper = sN->x[panel + 1];
            const double center = (lower + upper) / 2.0, half = (upper - lower) / 2.0;
            for (int i = 0; i < nr; ++i) {
                const double z = exp(center + half * gr_x[i]);
                for (int j = 0; j < nt; ++j) {
                    const double v = (gt_x[j] + 1.0) / 2.0;
                    total += (long double)(orc_k1_integrand_magnitude(r, z, M_PI * v * v, sN, p) *
                                           (2.0 * z * z) * (half * gr_w[i]) * (gt_w[j] * M_PI * v));
                }
            }
        }
        out[pi] = (double)total;
    }
}

/* ----------------------------------------------------------- K2, Kf (NLO) */

/* S/bk_kernels/k2.py:8-39. */
double orc_kernel_nlo(double r, double X, double Y, double X2, double Y2, double zmz2) {
    if (X < 1e-20 || X2 < 1e-20 || Y < 1e-20 || Y2 < 1e-20 || zmz2 < 1e-20) return 0.0;
    const double XY2sq = sqr(X * Y2), X2Ysq = sqr(X2 * Y);
    const double diff = XY2sq - X2Ysq;
    const double z4 = pow(zmz2, 4.0);
    double k = -2.0 / z4;
    k = k + ((XY2sq + X2Ysq - 4.0 * sqr(r * zmz2)) / (z4 * diff) + pow(r, 4.0) / (XY2sq * diff) +
             r * r / sqr(X * Y2 * zmz2)) *
                2.0 * log(X * Y2 / (X2 * Y));
    return isfinite(k) ? k : 0.0;
}

/* Rounding-error scale of Kernel_nlo / Kernel_nlo_fermion: every additive
 * term in absolute value, the Delta-divided ones amplified by the relative
 * size of Delta = (X Y2)^2 - (X2 Y)^2 (its subtraction loses that many digits). */
static double kernel_nlo_scale(double r, double X, double Y, double X2, double Y2, double zmz2,
                               int fermion) {
    if (X < 1e-20 || X2 < 1e-20 || Y < 1e-20 || Y2 < 1e-20 || zmz2 < 1e-20) return 0.0;
    const double XY2sq = sqr(X * Y2), X2Ysq = sqr(X2 * Y);
    const double diff = XY2sq - X2Ysq;
    const double rel = fabs(diff) / (XY2sq + X2Ysq);
    const double amp = 1.0 / (rel > 1e-300 ? rel : 1e-300);
    const double z4 = pow(zmz2, 4.0);
    const double L = fabs(2.0 * log(X * Y2 / (X2 * Y)));
    double m;
    if (fermion) {
        m = 2.0 / z4 + fabs((XY2sq + X2Ysq + sqr(r * zmz2)) / (z4 * diff)) * L * amp;
    } else {
        m = 2.0 / z4 + r * r / sqr(X * Y2 * zmz2) * L +
            (fabs((XY2sq + X2Ysq + 4.0 * sqr(r * zmz2)) / (z4 * diff)) +
             fabs(pow(r, 4.0) / (XY2sq * diff))) * L * amp;
    }
    return isfinite(m) ? m : 0.0;
}

/* S/bk_kernels/kf.py:10-41. */
double orc_kernel_nlo_fermion(double r, double X, double Y, double X2, double Y2, double zmz2,
                              double nf, double nc) {
    if (X < 1e-20 || X2 < 1e-20 || Y < 1e-20 || Y2 < 1e-20 || zmz2 < 1e-20) return 0.0;
    const double XY2sq = sqr(X * Y2), X2Ysq = sqr(X2 * Y);
    double k = 2.0 / pow(zmz2, 4.0);
    k = k - ((XY2sq + X2Ysq - sqr(r * zmz2)) / (pow(zmz2, 4.0) * (XY2sq - X2Ysq)) * 2.0 *
             log(X * Y2 / (X2 * Y)));
    k = k * nf / nc;
    return isfinite(k) ? k : 0.0;
}

/* Kernels times J = z^2 z'^2 in a form free of the 1/Delta cancellation.
 * With a = X Y', b = X' Y, L = ln(a / b), Delta = a^2 - b^2 = b^2 (e^{2L} - 1):
 *   (a^2 + b^2) 2L / Delta = g(L) + 1,  g(L) = L coth L - 1,
 *   2L / Delta = h(L) / b^2,             h(L) = 2L / (e^{2L} - 1),  h(-L) = h(L) e^{2L},
 * so (S/bk_kernels/k2.py:24-36, kf.py:28-38)
 *   K2 J  = 2 A^2 g - 4 h B^2 + h C^2 + 2 L D^2
 *   K2s J = 2 A^2 g - 4 h(-L) D^2 + h(-L) C^2 - 2 L B^2
 *   Kf J  = (-2 A^2 g + h B^2) Nf / Nc
 *   Kfs J = (-2 A^2 g + h(-L) D^2) Nf / Nc
 * with the scale-free ratios A = z z'/Z^2, B = r z'/(X' Z), C = r^2/(X X'),
 * D = r z/(X Z).  g uses L^2 P(L^2) / Q(L^2) for |L| < 1, where
 * L cosh L - sinh L = L^3 P(L^2), sinh L = L Q(L^2) (positive series). */
static const double kGP[11] = {1.0 / 3.0,
                               1.0 / 30.0,
                               1.0 / 840.0,
                               1.0 / 45360.0,
                               1.0 / 3991680.0,
                               1.0 / 518918400.0,
                               1.0 / 93405312000.0,
                               1.0 / 22230464256000.0,
                               1.0 / 6758061133824000.0,
                               1.0 / 2554547108585472000.0,
                               1.0 / 1175091669949317120000.0};
static const double kGQ[12] = {1.0,
                               1.0 / 6.0,
                               1.0 / 120.0,
                               1.0 / 5040.0,
                               1.0 / 362880.0,
                               1.0 / 39916800.0,
                               1.0 / 6227020800.0,
                               1.0 / 1307674368000.0,
                               1.0 / 355687428096000.0,
                               1.0 / 121645100408832000.0,
                               1.0 / 51090942171709440000.0,
                               1.0 / 25852016738884976640000.0};

double orc_coth_minus_one(double L) {
    L = fabs(L); /* even function */
    if (L < 1.0) {
        const double y = L * L;
        double pp = kGP[10], qq = kGQ[11];
        for (int i = 9; i >= 0; --i) pp = pp * y + kGP[i];
        for (int i = 10; i >= 0; --i) qq = qq * y + kGQ[i];
        return y * pp / qq;
    }
    return L + 2.0 * L / expm1(2.0 * L) - 1.0;
}

double orc_h_ratio(double L) { return L == 0.0 ? 1.0 : 2.0 * L / expm1(2.0 * L); }

/* ln(X / Y) for X^2 = (r - z)^2 + 4 r z sin^2(theta/2), Y = z:
 * (X^2 - z^2) / z^2 = rho (rho - 2 cos theta), rho = r / z. */
double orc_log_x_over_y(double r, double z, double theta, double xsq) {
    const double rho = r / z;
    const double u = rho * (rho - 2.0 * cos(theta));
    if (fabs(u) < 0.5) return 0.5 * log1p(u);
    return 0.5 * log(xsq) - log(z);
}

void orc_nlo_kernels_stable(double r, double z, double theta_z, double z2, double theta_z2,
                            double nf, double nc, double* out /* kJ, ksJ, kfJ, kfsJ */) {
    const double s1 = sin(theta_z / 2.0), s2 = sin(theta_z2 / 2.0);
    const double s3 = sin((theta_z - theta_z2) / 2.0);
    const double Xsq = (r - z) * (r - z) + 4.0 * r * z * s1 * s1;
    const double X2sq = (r - z2) * (r - z2) + 4.0 * r * z2 * s2 * s2;
    const double Zsq = (z - z2) * (z - z2) + 4.0 * z * z2 * s3 * s3;
    const double X = sqrt(Xsq), X2 = sqrt(X2sq), Z = sqrt(Zsq);
    const double L = orc_log_x_over_y(r, z, theta_z, Xsq) - orc_log_x_over_y(r, z2, theta_z2, X2sq);
    const double g = orc_coth_minus_one(L);
    const double h = orc_h_ratio(L);
    const double hm = h * exp(2.0 * L); /* h(-L), no cancellation for either sign */
    const double A = z * z2 / Zsq, B = r * z2 / (X2 * Z), C = r * r / (X * X2), D = r * z / (X * Z);
    out[0] = 2.0 * A * A * g - 4.0 * h * B * B + h * C * C + 2.0 * L * D * D;
    out[1] = 2.0 * A * A * g - 4.0 * hm * D * D + hm * C * C - 2.0 * L * B * B;
    out[2] = (-2.0 * A * A * g + h * B * B) * nf / nc;
    out[3] = (-2.0 * A * A * g + hm * D * D) * nf / nc;
}

/* _nlo_alpha_s (S/bk/derivatives.py:615-632): FIXED_AS, the parent's
 * coupling or the coupling at the smallest of the six distances. */
static double nlo_alpha_s(double r, double smallest, const orc_bk_params* p) {
    if (p->rc_nlo == ORC_RC_NLO_FIXED) return p->fixed_as;
    if (p->rc_nlo == ORC_RC_NLO_PARENT) return orc_alpha_s(r, p);
    return orc_alpha_s(smallest, p);
}

/* S/bk/derivatives.py:490-612 (_nlo_integrand, eager branch) times the
 * exp(2 log z + 2 log z') Jacobian of S/bk/derivatives.py:322-335. */
static double nlo_impl(double r, double log_z, double log_z2, double theta_z,
                       double theta_z2, const orc_spline* sS, const orc_bk_params* p,
                       double* k2_part, double* kf_part, double* magnitude) {
    const double z = exp(log_z), z2 = exp(log_z2);
    const double Xsqr = dist_sq(r, z, theta_z, p->geometry_cos);
    const double X2sqr = dist_sq(r, z2, theta_z2, p->geometry_cos);
    const double zzsqr = dist_sq(z, z2, theta_z - theta_z2, p->geometry_cos);
    const int negative = Xsqr < 0.0 || X2sqr < 0.0 || zzsqr < 0.0;
    const double X = sqrt(Xsqr > 0.0 ? Xsqr : 0.0);
    const double X2 = sqrt(X2sqr > 0.0 ? X2sqr : 0.0);
    const double zz = sqrt(zzsqr > 0.0 ? zzsqr : 0.0);
    const int invalid =
        negative || X < 1e-20 || z < 1e-20 || X2 < 1e-20 || z2 < 1e-20 || zz < 1e-20;
    if (k2_part) *k2_part = 0.0;
    if (kf_part) *kf_part = 0.0;
    if (magnitude) *magnitude = 0.0;
    if (invalid) return 0.0;
    const double Y = z, Y2 = z2;
    const double S_X = orc_loglog_eval(sS, X, 1.0, 0.0);
    const double S_Y = orc_loglog_eval(sS, Y, 1.0, 0.0);
    const double S_X2 = orc_loglog_eval(sS, X2, 1.0, 0.0);
    const double S_Y2 = orc_loglog_eval(sS, Y2, 1.0, 0.0);
    const double S_zz = orc_loglog_eval(sS, zz, 1.0, 0.0);

    double mn = r;
    const double cand[5] = {X, Y, X2, Y2, zz};
    for (int i = 0; i < 5; ++i) mn = cand[i] < mn ? cand[i] : mn;
    const double as = nlo_alpha_s(r, mn, p);

    double k = orc_kernel_nlo(r, X, Y, X2, Y2, zz);
    double ks = orc_kernel_nlo(r, X2, Y2, X, Y, zz);
    double dip = -(S_X * S_zz * S_Y2 - S_X * S_Y);
    double dips = -(S_X2 * S_zz * S_Y - S_X2 * S_Y2);
    if (fabs(k) > 1e10 && fabs(dip) < 1e-10) { k = 0.0; dip = 0.0; }
    if (fabs(ks) > 1e10 && fabs(dips) < 1e-10) { ks = 0.0; dips = 0.0; }
    double k2r = p->symmetrize ? (k * dip + ks * dips) / 2.0 : k * dip;
    double kfr = 0.0;
    double mag = (kernel_nlo_scale(r, X, Y, X2, Y2, zz, 0) *
                      (fabs(S_X * S_zz * S_Y2) + fabs(S_X * S_Y)) +
                  kernel_nlo_scale(r, X2, Y2, X, Y, zz, 0) *
                      (fabs(S_X2 * S_zz * S_Y) + fabs(S_X2 * S_Y2)));
    if (p->nf > 0.0) {
        const double kf = orc_kernel_nlo_fermion(r, X, Y, X2, Y2, zz, p->nf, p->nc);
        const double kfs = orc_kernel_nlo_fermion(r, X2, Y2, X, Y, zz, p->nf, p->nc);
        const double df = S_Y * (S_X2 - S_X);
        const double dfs = S_Y2 * (S_X - S_X2);
        kfr = -(p->symmetrize ? (kf * df + kfs * dfs) / 2.0 : kf * df);
        mag += p->nf / p->nc *
               (kernel_nlo_scale(r, X, Y, X2, Y2, zz, 1) * fabs(S_Y) * (fabs(S_X2) + fabs(S_X)) +
                kernel_nlo_scale(r, X2, Y2, X, Y, zz, 1) * fabs(S_Y2) * (fabs(S_X) + fabs(S_X2)));
    }
    const double scale = sqr(as * p->nc) / (8.0 * pow(M_PI, 4.0)) * exp(2.0 * log_z + 2.0 * log_z2);
    k2r *= scale;
    kfr *= scale;
    if (!isfinite(k2r)) k2r = 0.0;
    if (!isfinite(kfr)) kfr = 0.0;
    if (k2_part) *k2_part = k2r;
    if (kf_part) *kf_part = kfr;
    if (magnitude) *magnitude = isfinite(mag * scale) ? mag * scale : 0.0;
    return k2r + kfr;
}

double orc_nlo_integrand_log(double r, double log_z, double log_z2, double theta_z,
                             double theta_z2, const orc_spline* sS, const orc_bk_params* p,
                             double* k2_part, double* kf_part) {
    return nlo_impl(r, log_z, log_z2, theta_z, theta_z2, sS, p, k2_part, kf_part, NULL);
}

/* Sum of the absolute values of the kernel x dipole products of the NLO
 * integrand (tolerance scale for tests where the reference formula cancels). */
double orc_nlo_integrand_magnitude(double r, double log_z, double log_z2, double theta_z,
                                   double theta_z2, const orc_spline* sS,
                                   const orc_bk_params* p) {
    double m = 0.0;
    (void)nlo_impl(r, log_z, log_z2, theta_z, theta_z2, sS, p, NULL, NULL, &m);
    return m;
}

/* N from the log-log N spline at ln d: clamp and range tests in log space
 * (below the grid 0, above it 1; S/bk/derivatives.py:761-772). */
static double n_from_log(const orc_spline* sN, double log_d) {
    const double lo = sN->x[0], hi = sN->x[sN->n - 1];
    double lv = log_d < lo ? lo : log_d;
    lv = lv > hi ? hi : lv;
    double out = exp(orc_spline_eval(sN, lv));
    if (!isfinite(out)) out = 0.0;
    if (log_d < lo) out = 0.0;
    if (log_d > hi) out = 1.0;
    return out;
}

/* The K2 + Kf integrand of S/bk/derivatives.py:521-612 evaluated without
 * catastrophic cancellation: distances and their logs from the sin^2 forms,
 * kernels times J = z^2 z'^2 from orc_nlo_kernels_stable, N from the log-log
 * N spline and the dipole combinations in N = 1 - S form
 *   dip  = S_X (N_Z + N_Y' - N_Y - N_Z N_Y'),  dip' = S_X' (N_Z + N_Y - N_Y' - N_Z N_Y),
 *   dipf = S_Y (N_X - N_X'),                   dipf' = S_Y' (N_X' - N_X),
 * the coupling at the smallest distance from ln d.  Same result as the
 * reference formula in exact arithmetic, up to the N- vs S-spline choice. */
double orc_nlo_integrand_stable(double r, double log_z, double log_z2, double theta_z,
                                double theta_z2, const orc_spline* sN, const orc_bk_params* p,
                                double* k2_part, double* kf_part) {
    if (k2_part) *k2_part = 0.0;
    if (kf_part) *kf_part = 0.0;
    const double z = exp(log_z), z2 = exp(log_z2);
    const double s1 = sin(0.5 * theta_z), s2 = sin(0.5 * theta_z2);
    const double s3 = sin(0.5 * (theta_z - theta_z2));
    const double xsq = (r - z) * (r - z) + 4.0 * r * z * s1 * s1;
    const double x2sq = (r - z2) * (r - z2) + 4.0 * r * z2 * s2 * s2;
    const double zsq = (z - z2) * (z - z2) + 4.0 * z * z2 * s3 * s3;
    if (!(xsq >= 1e-40) || !(x2sq >= 1e-40) || !(zsq >= 1e-40)) return 0.0;
    const double lxy = orc_log_x_over_y(r, z, theta_z, xsq);
    const double lx2y2 = orc_log_x_over_y(r, z2, theta_z2, x2sq);
    const double log_x = log_z + lxy, log_x2 = log_z2 + lx2y2, log_zz = 0.5 * log(zsq);
    const double n_x = n_from_log(sN, log_x), n_y = n_from_log(sN, log_z);
    const double n_x2 = n_from_log(sN, log_x2), n_y2 = n_from_log(sN, log_z2);
    const double n_z = n_from_log(sN, log_zz);
    double dip = (1.0 - n_x) * (n_z + n_y2 - n_y - n_z * n_y2);
    double dips = (1.0 - n_x2) * (n_z + n_y - n_y2 - n_z * n_y);
    double kk[4];
    orc_nlo_kernels_stable(r, z, theta_z, z2, theta_z2, p->nf, p->nc, kk);
    double k = kk[0], ks = kk[1];
    const double cut = 1e10 * exp(2.0 * (log_z + log_z2)); /* |K2| > 1e10 */
    if (fabs(k) > cut && fabs(dip) < 1e-10) { k = 0.0; dip = 0.0; }
    if (fabs(ks) > cut && fabs(dips) < 1e-10) { ks = 0.0; dips = 0.0; }
    double k2r = p->symmetrize ? (k * dip + ks * dips) / 2.0 : k * dip;
    double kfr = 0.0;
    if (p->nf > 0.0) {
        const double df = (1.0 - n_y) * (n_x - n_x2);
        const double dfs = (1.0 - n_y2) * (n_x2 - n_x);
        kfr = -(p->symmetrize ? (kk[2] * df + kk[3] * dfs) / 2.0 : kk[2] * df);
    }
    double lmin = log(r);
    const double cand[5] = {log_x, log_z, log_x2, log_z2, log_zz};
    for (int i = 0; i < 5; ++i) lmin = cand[i] < lmin ? cand[i] : lmin;
    double as;
    if (p->rc_nlo != ORC_RC_NLO_SMALLEST || p->rc_lo == ORC_RC_LO_FIXED || p->nf > 3.0) {
        as = nlo_alpha_s(r, exp(lmin), p);
    } else { /* the smooth coupling from ln d */
        const double b0 = (11.0 * p->nc - 2.0 * p->nf) / (12.0 * M_PI);
        const double left = 2.0 / 0.2 * log(2.5);
        const double right = 5.0 * (log(4.0 * p->c2 / (p->lambda_qcd * p->lambda_qcd)) - 2.0 * lmin);
        const double hi = left > right ? left : right;
        as = 1.0 / (b0 * (0.2 * (hi + log1p(exp(-fabs(left - right))))));
    }
    const double scale = sqr(as * p->nc) / (8.0 * pow(M_PI, 4.0));
    k2r *= scale;
    kfr *= scale;
    if (!isfinite(k2r)) k2r = 0.0;
    if (!isfinite(kfr)) kfr = 0.0;
    if (k2_part) *k2_part = k2r;
    if (kf_part) *kf_part = kfr;
    return k2r + kfr;
}

/* ------------------------------------------------------------------- Sobol */

/* Gray-code Sobol (SobolEngine): q_n = shift ^ XOR_{b in gray(n)} v_b. */
void orc_sobol_u32(int dim, const uint32_t* direction, const uint32_t* shift, int64_t first,
                   int64_t n, uint32_t* out) {
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t idx = (uint64_t)(first + i);
        const uint64_t g = idx ^ (idx >> 1);
        for (int d = 0; d < dim; ++d) {
            uint32_t q = shift[d];
            for (int b = 0; b < 30; ++b)
                if ((g >> b) & 1u) q ^= direction[d * 30 + b];
            out[i * dim + d] = q;
        }
    }
}

/* ------------------------------------------------------------------- VEGAS */

/* u = q / 2^30; bin = min(floor(u B), B - 1); frac = u B - bin
 * (S/dis/fixed_sobol.py:100-113); x = lo + width (edge_l + (edge_r - edge_l) frac),
 * jac = prod_d B (edge_r - edge_l) * width_d. */
static void map_one(int dim, int bins, const uint32_t* q, const double* edges, const double* lo,
                    const double* width, double* x, double* jac, int32_t* bin) {
    double j = 1.0;
    for (int d = 0; d < dim; ++d) {
        const double u = (double)q[d] / 1073741824.0;
        const double scaled = u * (double)bins;
        int b = (int)floor(scaled);
        if (b > bins - 1) b = bins - 1;
        const double frac = scaled - (double)b;
        const double left = edges[d * (bins + 1) + b], right = edges[d * (bins + 1) + b + 1];
        const double w = right - left;
        x[d] = lo[d] + width[d] * (left + w * frac);
        j *= (double)bins * w;
        if (bin) bin[d] = b;
    }
    double vol = 1.0;
    for (int d = 0; d < dim; ++d) vol *= width[d];
    *jac = vol * j;
}

void orc_vegas_map(int dim, int bins, int64_t n, const uint32_t* q, const double* edges,
                   const double* lo, const double* width, double* x, double* jac, int32_t* bin) {
    for (int64_t i = 0; i < n; ++i)
        map_one(dim, bins, q + i * dim, edges, lo, width, x + i * dim, jac + i,
                bin ? bin + i * dim : NULL);
}

/* S/numerics/integration.py:828-866 (eager branch). */
void orc_vegas_adapt(int dim, int bins, const double* edges, const double* hist,
                     const double* count, double eps, double* out) {
    double* imp = (double*)malloc(sizeof(double) * (size_t)bins);
    double* sm = (double*)malloc(sizeof(double) * (size_t)bins);
    double* cum = (double*)malloc(sizeof(double) * (size_t)(bins + 1));
    for (int d = 0; d < dim; ++d) {
        const double* h = hist + d * bins;
        const double* c = count + d * bins;
        const double* e = edges + d * (bins + 1);
        double* o = out + d * (bins + 1);
        for (int i = 0; i < bins; ++i) imp[i] = h[i] / (c[i] > 1.0 ? c[i] : 1.0);
        for (int i = 1; i < bins - 1; ++i) sm[i] = (imp[i - 1] + imp[i] + imp[i + 1]) / 3.0;
        sm[0] = (imp[0] + imp[1]) / 2.0;
        sm[bins - 1] = (imp[bins - 2] + imp[bins - 1]) / 2.0;
        for (int i = 0; i < bins; ++i) sm[i] += eps;
        cum[0] = 0.0;
        for (int i = 0; i < bins; ++i) cum[i + 1] = cum[i] + sm[i];
        const double total = cum[bins];
        for (int t = 0; t <= bins; ++t) {
            /* torch.linspace(0, total, bins + 1) on the CPU: i * step for the
             * lower half, fma(-step, steps - 1 - i, end) for the upper half
             * (matched bit for bit against torch.linspace). */
            const double step = total / (double)bins;
            const double target = t < (bins + 1) / 2 ? (double)t * step
                                                     : fma(-step, (double)(bins - t), total);
            int lo = 0, hi = bins + 1;
            while (lo < hi) {
                const int mid = (lo + hi) / 2;
                if (target >= cum[mid]) lo = mid + 1;
                else hi = mid;
            }
            int k = lo - 1;
            if (k < 0) k = 0;
            if (k > bins - 1) k = bins - 1;
            const double frac = (target - cum[k]) / sm[k];
            o[t] = e[k] + frac * (e[k + 1] - e[k]);
        }
        o[0] = 0.0;
        o[bins] = 1.0;
    }
    free(imp);
    free(sm);
    free(cum);
}

/* Per-parent VEGAS sums of the K2 + Kf integrand (S/numerics/integration.py:
 * 392-428 with Sobol points in place of torch.rand / torch.randint). */
void orc_bk_nlo_sums(int parents, const double* parent_r, const orc_spline* sS, int bins,
                     const double* edges, const double* lo, const double* width,
                     const uint32_t* direction, const uint32_t* shift, const uint32_t* first_q,
                     int64_t first, int64_t n, const orc_bk_params* p, double* sums, double* hist,
                     double* count, double* magnitude) {
    const int dim = 4;
#pragma omp parallel for schedule(dynamic, 1)
    for (int pi = 0; pi < parents; ++pi) {
        const double* e = edges + (size_t)pi * dim * (bins + 1);
        long double s = 0, s2 = 0, sk2 = 0, skf = 0, sm = 0;
        long double* lh = hist ? (long double*)calloc((size_t)dim * bins, sizeof(long double)) : NULL;
        double* lc = count ? count + (size_t)pi * dim * bins : NULL;
        if (lc) memset(lc, 0, sizeof(double) * (size_t)dim * bins);
        for (int64_t i = 0; i < n; ++i) {
            uint32_t q[4];
            const int64_t idx = first + i;
            if (idx == 0 && first_q) {
                for (int d = 0; d < dim; ++d) q[d] = first_q[d];
            } else {
                orc_sobol_u32(dim, direction, shift, idx, 1, q);
            }
            double x[4], jac;
            int32_t b[4];
            map_one(dim, bins, q, e, lo, width, x, &jac, b);
            double k2, kf;
            const double f =
                p->stable
                    ? orc_nlo_integrand_stable(parent_r[pi], x[0], x[1], x[2], x[3], sS, p, &k2, &kf)
                    : orc_nlo_integrand_log(parent_r[pi], x[0], x[1], x[2], x[3], sS, p, &k2, &kf);
            const double w = f * jac;
            if (magnitude && !p->stable)
                sm += (long double)(orc_nlo_integrand_magnitude(parent_r[pi], x[0], x[1], x[2], x[3],
                                                                sS, p) *
                                    fabs(jac));
            s += w;
            s2 += (long double)w * w;
            sk2 += (long double)(k2 * jac);
            skf += (long double)(kf * jac);
            if (lh)
                for (int d = 0; d < dim; ++d) lh[d * bins + b[d]] += fabs(w);
            if (lc)
                for (int d = 0; d < dim; ++d) lc[d * bins + b[d]] += 1.0;
        }
        sums[4 * pi + 0] = (double)s;
        sums[4 * pi + 1] = (double)s2;
        sums[4 * pi + 2] = (double)sk2;
        sums[4 * pi + 3] = (double)skf;
        if (magnitude) magnitude[pi] = (double)sm;
        if (lh) {
            for (int k = 0; k < dim * bins; ++k) hist[(size_t)pi * dim * bins + k] = (double)lh[k];
            free(lh);
        }
    }
}
