/*
 * FP64 CPU oracle for the NLO BK right-hand side and rapidity driver.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the reference the HIP path is
 * tested against: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  The product library (nlo_bk_optimized_amd/) never links,
 * loads or calls it.
 *
 * Every function restates the reference algorithm from its source text and
 * cites the reference file:line it follows (prefix S/ = torch_nlo_clean/src/
 * nlo_torch/, K/ = S/custom_kernels/).  Arithmetic is IEEE double; sums over
 * samples are accumulated in long double so that the oracle's own rounding
 * is far below the 1e-10 parity tolerance.
 */
#ifndef NLO_ORACLE_H
#define NLO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physics choices of BKConfig (S/bk/config.py:9-64, :66-101).  The choice
 * codes are those of the C interface (include/nlo_gfx950.h); zero is the
 * production choice in every field. */
#define ORC_RC_LO_BALITSKY 0
#define ORC_RC_LO_FIXED 1
#define ORC_RC_LO_PARENT 2
#define ORC_RC_LO_SMALLEST 3
#define ORC_RC_LO_FAC 4
#define ORC_RC_LO_BEUF 5
#define ORC_RC_NLO_SMALLEST 0
#define ORC_RC_NLO_FIXED 1
#define ORC_RC_NLO_PARENT 2
#define ORC_RESUM_RC_SMALLEST 0
#define ORC_RESUM_RC_PARENT 1
#define ORC_RESUM_RC_FIXED 2
#define ORC_ORDER_NLO_RESUM_DLOG_SLOG 0
#define ORC_ORDER_NLO_RESUM_DLOG 1
#define ORC_ORDER_NLO 2
#define ORC_ORDER_LO_RESUM_DLOG_SLOG 3
#define ORC_ORDER_LO_RESUM_DLOG 4
#define ORC_ORDER_LO 5

typedef struct orc_bk_params {
    double nc;          /* NC */
    double nf;          /* NF (0, 3: smooth coupling; 5: flavour thresholds) */
    double lambda_qcd;  /* LambdaQCD */
    double c2;          /* C2 */
    double ksub;        /* KSUB */
    double minr;        /* MINR (double-log resummation switch r > 1.01 MINR) */
    int symmetrize;     /* SYMMETRIZE_Z_Z2_INTEGRATION */
    int geometry_cos;   /* 1: X^2 = r^2 + z^2 - 2 r z cos(theta) (torch code);
                           0: X^2 = (r - z)^2 + 4 r z sin^2(theta / 2) */
    int stable;         /* orc_bk_nlo_sums: 0 = reference integrand on the S
                           spline, 1 = orc_nlo_integrand_stable on the N spline */
    int rc_lo;          /* ORC_RC_LO_* */
    int rc_nlo;         /* ORC_RC_NLO_* */
    int resum_rc;       /* ORC_RESUM_RC_* */
    int order;          /* ORC_ORDER_* */
    int reserved;
    double fixed_as;    /* FIXED_AS */
} orc_bk_params;

/* Amplitude history for the kinematical constraint (S/bk/derivatives.py:712-758):
 * rows N(r_grid, y_k), k < rows, and the rapidity of the current derivative. */
typedef struct orc_history {
    int rows;
    int n;               /* r grid points */
    const double* r;     /* r grid */
    const double* y;     /* rows rapidities, increasing */
    const double* N;     /* rows x n */
    double rapidity;
} orc_history;

/* Natural cubic spline of log(y) against log(r) (S/numerics/interpolation.py:95-121). */
typedef struct orc_spline {
    int n;              /* nodes */
    double r_min;       /* r_grid[0]  (clamp/underflow limit) */
    double r_max;       /* r_grid[-1] (clamp/overflow limit)  */
    const double* x;    /* log r, n values */
    const double* a;    /* n - 1 coefficients each */
    const double* b;
    const double* c;
    const double* d;
} orc_spline;

double orc_alpha_s(double r, const orc_bk_params* p);
double orc_mv(double r, double qs0sqr, double gamma, double ec, double lambda_qcd);
void orc_r_grid(double minr, double maxr, int rpoints, double* out);

void orc_natural_spline(int n, const double* x, const double* y,
                        double* a, double* b, double* c, double* d);
/* Evolution splines of N and S = 1 - N (S/bk/evolution.py:548-560).
 * x_out: n log r; coef_n / coef_s: 4 * (n - 1) values laid out a|b|c|d. */
void orc_evolution_splines(int n, const double* r, const double* N, int force_positive,
                           double* x_out, double* coef_n, double* coef_s);
double orc_spline_eval(const orc_spline* s, double log_value);
double orc_loglog_eval(const orc_spline* s, double value, double underflow, double overflow);

double orc_bessel_j1(double x);
double orc_bessel_i1(double x);

double orc_kernel_lo(double r, double z, double theta, const orc_bk_params* p);
double orc_k1_integrand(double r, double z, double theta, const orc_spline* sN,
                        const orc_bk_params* p);
/* Tolerance scales: the same expressions with every additive piece in
 * absolute value (only tests use them). */
double orc_kernel_lo_magnitude(double r, double z, double theta, const orc_bk_params* p);
double orc_k1_integrand_magnitude(double r, double z, double theta, const orc_spline* sN,
                                  const orc_bk_params* p);
double orc_nlo_integrand_magnitude(double r, double log_z, double log_z2, double theta_z,
                                   double theta_z2, const orc_spline* sS,
                                   const orc_bk_params* p);
double orc_kernel_nlo(double r, double X, double Y, double X2, double Y2, double zmz2);
double orc_kernel_nlo_fermion(double r, double X, double Y, double X2, double Y2,
                              double zmz2, double nf, double nc);
/* Returns the full integrand (with the exp(2 log z + 2 log z') Jacobian of
 * S/bk/derivatives.py:335); k2_part / kf_part receive the two pieces. */
double orc_nlo_integrand_log(double r, double log_z, double log_z2, double theta_z,
                             double theta_z2, const orc_spline* sS, const orc_bk_params* p,
                             double* k2_part, double* kf_part);

/* Cancellation-free form of the same integrand (N spline, N-form dipoles,
 * kernels from orc_nlo_kernels_stable), as the FP32 GPU path evaluates it. */
double orc_nlo_integrand_stable(double r, double log_z, double log_z2, double theta_z,
                                double theta_z2, const orc_spline* sN, const orc_bk_params* p,
                                double* k2_part, double* kf_part);
/* K2 J, K2s J, Kf J, Kfs J (J = z^2 z'^2) without the 1/Delta cancellation. */
void orc_nlo_kernels_stable(double r, double z, double theta_z, double z2, double theta_z2,
                            double nf, double nc, double* out);
double orc_coth_minus_one(double L);
double orc_h_ratio(double L);
double orc_log_x_over_y(double r, double z, double theta, double xsq);

/* K1 on the fixed production rule: (n-1) log-r panels x GL(nr) x GL(nt) in
 * v with theta = pi v^2 (K/bk/k1/kernel.cu:2040-2092). */
void orc_k1_fixed(int parents, const double* parent_r, const orc_spline* sN,
                  int nr, const double* gr_x, const double* gr_w,
                  int nt, const double* gt_x, const double* gt_w,
                  const orc_bk_params* p, double* out);

/* The same rule with the kinematical constraint: the integrand of
 * orc_k1_integrand_history at h->rapidity. */
void orc_k1_fixed_history(int parents, const double* parent_r, const orc_spline* sN,
                          int nr, const double* gr_x, const double* gr_w,
                          int nt, const double* gt_x, const double* gt_w,
                          const orc_bk_params* p, const orc_history* h, double* out);

/* _interpolate_history_N (S/bk/derivatives.py:712-758): N at (r, rapidity)
 * from the history rows; the last row is read through sN, the spline of the
 * current amplitude.  NaN if rapidity is above the last row. */
double orc_history_N(double r, double rapidity, const orc_spline* sN, const orc_history* h);

/* _lo_integrand with KINEMATICAL_CONSTRAINT (S/bk/derivatives.py:445-468). */
double orc_k1_integrand_history(double r, double z, double theta, const orc_spline* sN,
                                const orc_bk_params* p, const orc_history* h);

/* Gauss-Kronrod 21 on [a, b] as _gauss_kronrod_intervals
 * (S/numerics/integration.py:624-746) forms it: the Kronrod estimate and the
 * error min(1, (200 |K - G| / resasc)^1.5) resasc, at least 50 eps resabs. */
typedef double (*orc_fn)(double x, void* ctx);
void orc_gk21(orc_fn f, void* ctx, double a, double b, double* estimate, double* error);

/* adaptive_gauss_kronrod_1d (S/numerics/integration.py:43-92) with rule 21
 * and epsabs 0: bisect the interval of largest error (the first of equal
 * ones) in an ordered list until the summed error <= epsrel |summed
 * estimate| or max_intervals intervals.  error, converged, evaluations may be
 * NULL. */
double orc_adaptive_gk21(orc_fn f, void* ctx, double a, double b, double epsrel, int max_intervals,
                         double* error, int* converged, long* evaluations);

/* rapidity_derivative_lo (S/bk/derivatives.py:43-110): the adaptive rule in
 * log z over [ln r_0, ln r_{n-1}] (rint intervals) of 2 z^2 times the
 * adaptive rule in theta over [0, pi] (tint intervals), both to epsrel.  h
 * (may be NULL) switches the kinematical constraint on.  err and evals (may
 * be NULL) receive the outer error estimate and the integrand evaluations. */
void orc_k1_adaptive(int parents, const double* parent_r, const orc_spline* sN, int rint, int tint,
                     double epsrel, const orc_bk_params* p, const orc_history* h, double* out,
                     double* err, long* evals);

void orc_k1_fixed_magnitude(int parents, const double* parent_r, const orc_spline* sN,
                            int nr, const double* gr_x, const double* gr_w,
                            int nt, const double* gt_x, const double* gt_w,
                            const orc_bk_params* p, double* out);

/* Scrambled Sobol integer states (gray-code order, SobolEngine semantics). */
void orc_sobol_u32(int dim, const uint32_t* direction /* dim x 30 */, const uint32_t* shift,
                   int64_t first, int64_t n, uint32_t* out /* n x dim */);

/* VEGAS map of integer Sobol states through an importance grid
 * (S/dis/fixed_sobol.py:100-113, S/numerics/integration.py:392-396). */
void orc_vegas_map(int dim, int bins, int64_t n, const uint32_t* q, const double* edges,
                   const double* lo, const double* width, double* x, double* jac,
                   int32_t* bin);

/* Grid refinement (S/numerics/integration.py:804-866). eps: dtype epsilon. */
void orc_vegas_adapt(int dim, int bins, const double* edges, const double* hist,
                     const double* count, double eps, double* out);

/* Per-parent K2 + Kf VEGAS sums over Sobol points [first, first + n)
 * (sS: the S spline for p->stable = 0, the N spline for p->stable = 1):
 * sums[4 p + k] = (sum w, sum w^2, sum w_K2, sum w_Kf).  hist / count
 * (parents x 4 x bins, may be NULL) receive sum |w| and hits per bin. */
void orc_bk_nlo_sums(int parents, const double* parent_r, const orc_spline* sS, int bins,
                     const double* edges /* parents x 4 x (bins + 1) */,
                     const double* lo, const double* width, const uint32_t* direction,
                     const uint32_t* shift, const uint32_t* first_q, int64_t first, int64_t n,
                     const orc_bk_params* p, double* sums, double* hist, double* count,
                     double* magnitude /* parents, may be NULL: sum |jac| x rounding scale of
                                          the reference integrand (p->stable = 0 only) */);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
