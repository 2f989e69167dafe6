/*
 * nlo_gfx950 — C ABI of the MI355X (gfx950) NLO BK right-hand side, Sobol /
 * VEGAS sampling and rapidity driver.
 *
 * Conventions (all entry points):
 *   - pointers named d_* are device pointers; the caller owns every buffer,
 *     including scratch ("workspace"), whose size is queried first;
 *   - every launch goes on the stream argument (hipStream_t); nothing
 *     synchronises the host except the adaptive RK2 step of the rapidity
 *     driver (one error ratio read per attempt); there is no hidden global
 *     state except a thread-local last-error string;
 *   - return value: NLO_OK (0) or an NLO_ERR_* code; nlo_last_error() gives
 *     the message.  The Python layer turns a non-zero code into RuntimeError,
 *     as the reference's TORCH_CHECK failures do.
 *
 * Reference prefixes: S/ = torch_nlo_clean/src/nlo_torch/,
 * K/ = S/custom_kernels/.  Each entry point names the reference interface it
 * replaces.
 */
#ifndef NLO_GFX950_H
#define NLO_GFX950_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NLO_OK 0
#define NLO_ERR_INVALID 1     /* bad argument (shape, size, null pointer) */
#define NLO_ERR_HIP 2         /* a HIP runtime call failed */
#define NLO_ERR_WORKSPACE 3   /* workspace too small */
#define NLO_ERR_UNSUPPORTED 4 /* configuration outside the implemented set */

#define NLO_API_VERSION 3

/* ------------------------------------------------------------------ misc */

int nlo_api_version(void);
/* Message of the last failed call on this host thread ("" if none). */
const char* nlo_last_error(void);
/* Gauss-Legendre nodes (ascending) and weights on [-1, 1], computed on the
 * host by Newton iteration in long double.  Replaces the reference's
 * numpy.polynomial.legendre.leggauss tables (K/quadrature.py:18). */
int nlo_gauss_legendre(int32_t n, double* x, double* w);

/* --------------------------------------------------------------- physics */

/* BKConfig physics choices (S/bk/config.py:9-64) and constants
 * (S/bk/config.py:66-101).  Zero is the production choice in every code
 * field: Balitsky LO coupling, smallest-distance NLO and resummation
 * couplings, NLO_RESUM_DLOG_SLOG.  RC_LO and RC_NLO are fixed together or not
 * at all (S/bk/config.py:119-122); with them every coupling is FIXED_AS.
 * NF = 0 or 3 selects the smooth coupling (S/coupling.py:81-97), NF = 5 the
 * one-loop coupling with flavour thresholds (S/coupling.py:23-46). */
#define NLO_RC_LO_BALITSKY 0
#define NLO_RC_LO_FIXED 1
#define NLO_RC_LO_PARENT 2
#define NLO_RC_LO_SMALLEST 3
#define NLO_RC_LO_FAC 4
#define NLO_RC_LO_BEUF 5
#define NLO_RC_NLO_SMALLEST 0
#define NLO_RC_NLO_FIXED 1
#define NLO_RC_NLO_PARENT 2
#define NLO_RESUM_RC_SMALLEST 0
#define NLO_RESUM_RC_PARENT 1
#define NLO_RESUM_RC_FIXED 2
#define NLO_ORDER_NLO_RESUM_DLOG_SLOG 0
#define NLO_ORDER_NLO_RESUM_DLOG 1
#define NLO_ORDER_NLO 2
#define NLO_ORDER_LO_RESUM_DLOG_SLOG 3
#define NLO_ORDER_LO_RESUM_DLOG 4
#define NLO_ORDER_LO 5

typedef struct nlo_bk_physics {
    double nc;          /* NC */
    double nf;          /* NF: 0, 3 or 5 */
    double lambda_qcd;  /* LambdaQCD (smooth coupling) */
    double c2;          /* C2 */
    double ksub;        /* KSUB */
    double minr;        /* MINR (double-log resummation only for r > 1.01 MINR) */
    int32_t symmetrize; /* SYMMETRIZE_Z_Z2_INTEGRATION */
    int32_t order;      /* NLO_ORDER_*: Order */
    int32_t rc_lo;      /* NLO_RC_LO_*: RC_LO */
    int32_t rc_nlo;     /* NLO_RC_NLO_*: RC_NLO */
    int32_t resum_rc;   /* NLO_RESUM_RC_*: RESUM_RC */
    int32_t reserved;   /* 0 */
    double fixed_as;    /* FIXED_AS */
} nlo_bk_physics;

/* ----------------------------------------------------------------- Sobol */

#define NLO_SOBOL_MAXDIM 8
#define NLO_SOBOL_BITS 30

/* Scrambled Sobol state of torch.quasirandom.SobolEngine(dim, scramble=True,
 * seed): direction = SobolEngine.sobolstate, shift = SobolEngine.shift,
 * first = integer state of the float64 first point (float32(shift / 2^30)
 * widened, torch/quasirandom.py _first_point).  Passed by value to kernels. */
typedef struct nlo_sobol_table {
    int32_t dim;
    uint32_t direction[NLO_SOBOL_MAXDIM][NLO_SOBOL_BITS];
    uint32_t shift[NLO_SOBOL_MAXDIM];
    uint32_t first[NLO_SOBOL_MAXDIM];
} nlo_sobol_table;

#define NLO_SOBOL_U32 0 /* raw 30-bit integer states (row 0 = shift)        */
#define NLO_SOBOL_F32 1 /* float(q) * 2^-30, SobolEngine.draw(dtype=float32) */
#define NLO_SOBOL_F64 2 /* q * 2^-30, row 0 = first * 2^-30 (dtype=float64)  */

/* Points [first_index, first_index + n) of the gray-code sequence, row-major
 * n x dim.  Replaces SobolEngine.draw_base2 / make_sobol_points
 * (S/dis/fixed_sobol.py:66-97). */
int nlo_sobol_draw(const nlo_sobol_table* table, int64_t first_index, int64_t n,
                   int32_t out_kind, void* d_out, hipStream_t stream);

/* ----------------------------------------------------------------- VEGAS */

/* Map unit-cube points through an importance grid: bin = min(floor(u B), B - 1),
 * frac = u B - bin, x = lo + width (e[bin] + (e[bin+1] - e[bin]) frac),
 * jac = prod_d B (e[bin+1] - e[bin]) width_d.  Bins are bit-exact (integer
 * arithmetic on q = u 2^30 for Sobol input).  Replaces map_importance_grid
 * (S/dis/fixed_sobol.py:100-113) and the sample step of vegas()
 * (S/numerics/integration.py:392-396).  d_bin may be NULL; bins <= 255. */
int nlo_vegas_map(int32_t dim, int32_t bins, int64_t n, const double* d_u,
                  const double* d_edges, const double* d_lo, const double* d_width,
                  double* d_x, double* d_jac, uint8_t* d_bin, hipStream_t stream);

/* Deterministic per-bin sum |w| and hit count over n samples
 * (hist, count: dim x bins, double).  Per-block 64-bit fixed-point sums
 * (exact, order-free) followed by a fixed-order tree over blocks.  Replaces
 * vegas_histogram (K/integration.cu:773-861). */
size_t nlo_vegas_histogram_workspace(int32_t dim, int32_t bins, int64_t n);
int nlo_vegas_histogram(int32_t dim, int32_t bins, int64_t n, const uint8_t* d_bin,
                        const double* d_absw, double* d_hist, double* d_count,
                        void* d_workspace, size_t workspace_bytes, hipStream_t stream);

/* Grid refinement for `rows` independent grids (rows x dim x (bins + 1)):
 * importance = hist / max(count, 1), 3-point smoothing (2-point at the ends),
 * + eps, cumulative sum, linspace targets, right-bisection, linear edge
 * interpolation, edges 0 and 1 pinned.  Replaces _adapt_vegas_edges
 * (S/numerics/integration.py:804-866) / vegas_adapt_edges (K/integration.cu:259-377). */
int nlo_vegas_adapt(int32_t rows, int32_t dim, int32_t bins, const double* d_edges,
                    const double* d_hist, const double* d_count, double eps,
                    double* d_edges_out, hipStream_t stream);

/* --------------------------------------------------------------- splines */

/* Size in doubles (or floats) of one spline table for R r-nodes:
 * [r_min, r_max, log r (R), N records (R-1) x {a,b,c,d}, S records (R-1) x {a,b,c,d}]. */
#define NLO_SPLINE_WORDS(R) (2 + (R) + 8 * ((R) - 1))

/* Natural cubic log-log splines of N (clamped to [0,1] / force_positive) and
 * S = 1 - N on log r, one FP64 Thomas solve per row (d_table64: records
 * {a, b, c, d} of ln N = a + b dx + c dx^2 + d dx^3, dx = ln r - ln r_i).
 * d_table32 (may be NULL) holds the same cubics for the FP32 path in the
 * local coordinate t = (ln r - ln r_0) / h - i, h = (ln r_{R-1} - ln r_0)/(R-1),
 * multiplied by log2 e, so that N = 2^(a' + t (b' + t (c' + t d'))); it
 * assumes the log-uniform r grid of the evolution.  Replaces
 * evolution_loglog_splines_mixed (K/interpolation.cu:374-451) and
 * _evolution_interpolators (S/bk/evolution.py:536-560). */
int nlo_bk_spline_build(int32_t R, const double* d_r, const double* d_N, int32_t force_positive,
                        double* d_table64, float* d_table32, hipStream_t stream);

/* ------------------------------------------------------------- BK terms */

/* NLO_PREC_FP64: the reference's K2/Kf formula (S/bk_kernels/k2.py, kf.py)
 *   in FP64, FP64 sums: the plain FP64 path.
 * NLO_PREC_FP32: FP32 integrand with K2/Kf rewritten without the 1/Delta
 *   cancellation (Delta = (X Y')^2 - (X' Y)^2) and dipoles in N = 1 - S
 *   form, so no FP64 band is needed; per-lane and cross-lane sums in FP64.
 * NLO_PREC_FP64_STABLE: the rewritten kernels in FP64 (the accurate FP64
 *   oracle on the GPU). */
#define NLO_PREC_FP64 0
#define NLO_PREC_FP32 1
#define NLO_PREC_FP64_STABLE 2
/* OR'd into nlo_bk_nlo's precision: only the first column (sum w) is formed
 * and the other three read 0, as in the evolver's right-hand side (the mean
 * is all dN/dy needs); not with histograms. */
#define NLO_BK_SUM_ONLY 0x100

/* K1 on the production fixed rule for every parent r_i, i = 1..R-1, with
 * N_i <= 0.99999 (others: 0): (R-1) log-r panels x GL(nr) in log z x GL(nt)
 * in v, theta = pi v^2, FP64.  d_k1: R doubles.  Replaces
 * mixed_fixed_grid_integrals (K/bk/k1/kernel.cu:2321-2418) /
 * _batched_k1_component (S/bk/evolution.py:491-533). */
int nlo_bk_k1(int32_t R, const double* d_r, const double* d_N, const double* d_table64,
              const nlo_bk_physics* physics, int32_t nr, int32_t nt, double* d_k1,
              hipStream_t stream);

/* The same with the kinematical constraint (KINEMATICAL_CONSTRAINT,
 * S/bk/derivatives.py:445-468): the dipole S_01 - S_02 S_12 with S_02, S_12
 * read from the amplitude history at the rapidity y - max(0, ln(min(X, Y)^2 /
 * r^2)) (_interpolate_history_N, S/bk/derivatives.py:712-758), where y is the
 * last history row's rapidity.  d_hist_y: capacity doubles, increasing,
 * d_hist_N: capacity x R amplitudes on the grid d_r, *d_rows (device int32,
 * read by the kernel and clamped to [1, capacity]) the rows in use; d_N must
 * be the last row (the current amplitude). */
int nlo_bk_k1_history(int32_t R, const double* d_r, const double* d_N, const double* d_table64,
                      const nlo_bk_physics* physics, int32_t nr, int32_t nt,
                      const double* d_hist_y, const double* d_hist_N, const int32_t* d_rows,
                      int32_t capacity, double* d_k1, hipStream_t stream);

/* K1 for every parent by the nested adaptive Gauss-Kronrod 21 rule of
 * rapidity_derivative_lo (S/bk/derivatives.py:43-110, with
 * adaptive_gauss_kronrod_1d and _gauss_kronrod_intervals of
 * S/numerics/integration.py:43-92, :624-746): in ln z over
 * [ln r_0, ln r_{R-1}] with at most rint intervals (RINTPOINTS) of 2 z^2
 * times the theta integral over [0, pi] with at most tint intervals
 * (THETAINTPOINTS), both to epsrel (INTACCURACY); rint, tint <= 100.  The
 * reference takes this rule wherever its fixed rule is not validated
 * (S/bk/evolution.py:491-533, S/bk/derivatives.py:145-166).  The history
 * pointers are all NULL, or all set for the kinematical constraint as in
 * nlo_bk_k1_history.  d_err (R, may be NULL): the summed outer error
 * estimate; d_evals (R, may be NULL): integrand evaluations.  Inactive
 * parents get 0. */
int nlo_bk_k1_adaptive(int32_t R, const double* d_r, const double* d_N, const double* d_table64,
                       const nlo_bk_physics* physics, int32_t rint, int32_t tint, double epsrel,
                       const double* d_hist_y, const double* d_hist_N, const int32_t* d_rows,
                       int32_t capacity, double* d_k1, double* d_err, int64_t* d_evals,
                       hipStream_t stream);

/* K2 + Kf VEGAS sums for every active parent over Sobol points
 * [first_index, first_index + n) mapped through the parent's grid
 * (d_edges: R x 4 x (bins + 1); row 0 unused) on the box
 * (log z, log z') in [ln(r_0/2), ln(2 r_{R-1})], angles in [0, 2 pi]
 * (S/bk/derivatives.py:424-431).  d_sums: R x 4 doubles
 * (sum w, sum w^2, sum w K2-part, sum w Kf-part); inactive parents get 0.
 * d_hist / d_count (R x 4 x bins, may be NULL) receive the adaptation
 * histograms.  precision: NLO_PREC_*, optionally | NLO_BK_SUM_ONLY.
 * Replaces bk_nlo_mixed_vegas_summaries (K/bk/k2_kf/kernel.cu:877-1062)
 * batched over parents. */
size_t nlo_bk_nlo_workspace(int32_t R, int32_t bins, int64_t n, int32_t want_hist);
int nlo_bk_nlo(int32_t R, const double* d_r, const double* d_N, const double* d_table64,
               const float* d_table32, const double* d_edges, int32_t bins,
               const nlo_sobol_table* sobol, int64_t first_index, int64_t n, int32_t precision,
               const nlo_bk_physics* physics, double* d_sums, double* d_hist, double* d_count,
               void* d_workspace, size_t workspace_bytes, hipStream_t stream);

/* nlo_bk_nlo without a workspace and without any allocation: no counterpart
 * in the reference; for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as nlo_bk_nlo without histograms
 * (precision may carry NLO_BK_SUM_ONLY); the per-block partial sums live in a
 * static device array of NLO_BK_NOALLOC_SCRATCH_BYTES, one per GPU, held until
 * the process ends.  A workload needing more (nlo_bk_nlo_workspace(R, bins,
 * n, 0) > NLO_BK_NOALLOC_SCRATCH_BYTES, e.g. R = 1024 with n = 2^22) is
 * refused with NLO_ERR_WORKSPACE.  Calls that overlap on one GPU (other
 * streams, other threads) share that array and race; serialise them. */
#define NLO_BK_NOALLOC_SCRATCH_BYTES (16u << 20)
int nlo_bk_nlo_noalloc(int32_t R, const double* d_r, const double* d_N, const double* d_table64,
                       const float* d_table32, const double* d_edges, int32_t bins,
                       const nlo_sobol_table* sobol, int64_t first_index, int64_t n, int32_t precision,
                       const nlo_bk_physics* physics, double* d_sums, hipStream_t stream);

/* dN/dy = K1 + sum_w / n for active parents, 0 elsewhere, non-finite -> 0
 * (S/bk/evolution.py:340-351).  d_k2kf (R, may be NULL) receives sum_w / n. */
int nlo_bk_rhs_finalize(int32_t R, const double* d_N, const double* d_k1, const double* d_sums,
                        int64_t n, double* d_dndy, double* d_k2kf, hipStream_t stream);

/* ------------------------------------------------------- rapidity driver */

#define NLO_RK_EULER 0    /* N + h K(N)                                          */
#define NLO_RK_MIDPOINT 1 /* N + h K(N + h/2 K(N))  (S/bk/evolution.py:265-276)  */
#define NLO_RK_KUTTA3 2   /* (K1 + 4 K2 + K3)/6      (S/bk/evolution.py:190-236) */
#define NLO_RK_CLASSIC4 3 /* classical fourth order                              */
#define NLO_RK_ADAPTIVE2 4 /* midpoint RK2 with step doubling and step-size control
                              (_adaptive_rk2_to, S/bk/evolution.py:239-326)      */

#define NLO_K1_FIXED 0
#define NLO_K1_ADAPTIVE 1

typedef struct nlo_bk_evolver_desc {
    int32_t R;            /* r nodes */
    int32_t bins;         /* VEGAS bins per dimension */
    int64_t n_samples;    /* Sobol points per parent per right-hand side */
    int64_t first_index;  /* first Sobol index of the production point set */
    int32_t precision;    /* NLO_PREC_* */
    int32_t rk_method;    /* NLO_RK_* */
    double step;          /* rapidity step h */
    int32_t force_positive; /* clamp N to [0, 1] after each step */
    int32_t nr, nt;       /* K1 Gauss-Legendre orders (8, 24) */
    int32_t nlo;          /* include K2 + Kf */
    nlo_bk_physics physics;
    nlo_sobol_table sobol;
    /* NLO_RK_ADAPTIVE2: DE_SOLVER_ABSERR, DE_SOLVER_RELERR and the machine
     * epsilon of the caller's dtype (step underflow test) */
    double abserr, relerr, eps;
    /* kinematical constraint (NLO_RK_EULER only): the evolver appends each
     * step's amplitude and rapidity to the history, which K1 reads.  The
     * caller fills row 0 (initial rapidity and amplitude); capacity rows. */
    int32_t kinematical;
    int32_t history_capacity;
    double* d_history_y;  /* capacity doubles */
    double* d_history_N;  /* capacity x R doubles */
    /* K1 rule: NLO_K1_FIXED (the nr x nt Gauss-Legendre rule per log-r panel)
     * or NLO_K1_ADAPTIVE (nlo_bk_k1_adaptive with k1_rint, k1_tint,
     * k1_epsrel) */
    int32_t k1_rule;
    int32_t k1_rint, k1_tint;
    int32_t k1_reserved;
    double k1_epsrel;
} nlo_bk_evolver_desc;

typedef struct nlo_bk_evolver nlo_bk_evolver;

/* One RK step per call, replayed from a hipGraph captured at creation.  The
 * evolver state N (R doubles) lives in the workspace; nlo_bk_evolver_state
 * returns its device address.  d_edges (R x 4 x (bins + 1)) is read by every
 * step (frozen grids) and must outlive the evolver. */
size_t nlo_bk_evolver_workspace(const nlo_bk_evolver_desc* desc);
int nlo_bk_evolver_create(const nlo_bk_evolver_desc* desc, const double* d_r,
                          const double* d_edges, void* d_workspace, size_t workspace_bytes,
                          hipStream_t stream, nlo_bk_evolver** out);
double* nlo_bk_evolver_state(nlo_bk_evolver* ev);
/* Device int32 flag, non-zero once any step produced a non-finite N
 * (S/bk/evolution.py:152-153 raises FloatingPointError). */
int32_t* nlo_bk_evolver_nonfinite_flag(nlo_bk_evolver* ev);
/* Device uint64: number of (parent, right-hand side) evaluations with an
 * active parent since creation; times n_samples it is the exact count of
 * K2 + Kf integrand samples evaluated. */
uint64_t* nlo_bk_evolver_active_evaluations(nlo_bk_evolver* ev);
/* Last step's dN/dy stages: d_k1 (R) and d_k2kf (R) of the final stage. */
const double* nlo_bk_evolver_k1(nlo_bk_evolver* ev);
const double* nlo_bk_evolver_k2kf(nlo_bk_evolver* ev);
/* One output step of the rapidity grid: y -> y + step.  Fixed-step methods
 * replay the captured graph asynchronously on `stream`.  NLO_RK_ADAPTIVE2
 * runs the step-size controller on the host: per attempt one graph of four
 * right-hand sides and the error reduction, then a synchronisation to read
 * the error ratio (as the reference's per-attempt .item()); it returns
 * NLO_ERR_INVALID with "step underflow" in nlo_last_error when the step
 * collapses.  With the kinematical constraint the call fails once the
 * history is full. */
int nlo_bk_evolver_step(nlo_bk_evolver* ev, hipStream_t stream);
/* Controller state of NLO_RK_ADAPTIVE2: current rapidity, next step size and
 * attempt counts (any pointer may be NULL).  For fixed-step methods y
 * advances by step per call and h = step. */
int nlo_bk_evolver_adaptive_state(const nlo_bk_evolver* ev, double* y, double* h, int64_t* accepted,
                                  int64_t* rejected);
int nlo_bk_evolver_destroy(nlo_bk_evolver* ev);

/* ------------------------------------------------------------- batches */

/* Many evolutions advanced together (the batched workload of 64 runs of
 * uneven grid sizes): run i has h_R[i] r nodes; d_r (sum R) and d_edges
 * (sum R x 4 x (bins + 1), row 0 of each run unused) concatenate the runs'
 * grids and frozen VEGAS grids.  Every run shares the physics, Sobol table,
 * samples per parent and step.  Per RK stage: one spline-build launch (a
 * block per run), one K1 launch (a block per node), one persistent K2 + Kf
 * launch whose blocks take (parent, chunk) items from a device counter, one
 * fixed-order reduce and one finalize, captured as one hipGraph per step.
 * spt (samples per lane: 4, 8, 16) and grid_blocks (persistent grid) may be
 * given or left 0: the first nlo_bk_batch_step then times the candidates on
 * the state the caller set and keeps the fastest (nlo_bk_batch_tuning reports
 * the choice).  Fixed-step methods only.  Replaces running solve_bk
 * (S/bk/evolution.py:33-58) once per initial condition. */
typedef struct nlo_bk_batch_desc {
    int32_t runs;
    int32_t bins;
    int64_t n_samples;
    int64_t first_index;
    int32_t precision;
    int32_t rk_method;      /* NLO_RK_EULER .. NLO_RK_CLASSIC4 */
    double step;
    int32_t force_positive;
    int32_t nr, nt;
    int32_t nlo;
    int32_t spt;            /* 0 = tune */
    int32_t grid_blocks;    /* 0 = tune */
    nlo_bk_physics physics;
    nlo_sobol_table sobol;
    int32_t k1_rule;        /* NLO_K1_FIXED or NLO_K1_ADAPTIVE (as in nlo_bk_evolver_desc) */
    int32_t k1_rint, k1_tint;
    int32_t k1_reserved;
    double k1_epsrel;
} nlo_bk_batch_desc;

typedef struct nlo_bk_batch nlo_bk_batch;

size_t nlo_bk_batch_workspace(const nlo_bk_batch_desc* desc, const int32_t* h_R);
int nlo_bk_batch_create(const nlo_bk_batch_desc* desc, const int32_t* h_R, const double* d_r,
                        const double* d_edges, void* d_workspace, size_t workspace_bytes,
                        hipStream_t stream, nlo_bk_batch** out);
/* Concatenated state N (sum R doubles, run order), non-finite flag, active
 * (parent, right-hand side) evaluation counter. */
double* nlo_bk_batch_state(nlo_bk_batch* batch);
int32_t* nlo_bk_batch_nonfinite_flag(nlo_bk_batch* batch);
uint64_t* nlo_bk_batch_active_evaluations(nlo_bk_batch* batch);
int nlo_bk_batch_step(nlo_bk_batch* batch, hipStream_t stream);
/* The K2 + Kf launch shape in use and its measured time per launch. */
int nlo_bk_batch_tuning(const nlo_bk_batch* batch, int32_t* spt, int32_t* grid_blocks,
                        double* ms_per_launch);
int nlo_bk_batch_destroy(nlo_bk_batch* batch);

/* ================================================================== DIS */

/* Special functions on device arrays (n doubles in, n doubles out), in FP64
 * (NLO_PREC_FP64) or computed in FP32 (NLO_PREC_FP32).  Replace the
 * reference's bessel_K0/K1/K2 and Li2 (S/numerics/special.py:18-97). */
#define NLO_FN_K0 0
#define NLO_FN_K1 1
#define NLO_FN_K2 2
#define NLO_FN_LI2 3
int nlo_dis_special(int32_t fn, int64_t n, const double* d_x, int32_t precision, double* d_out,
                    hipStream_t stream);

#define NLO_DIPOLE_GBW 0      /* S/dipole/amplitude.py:22-46 */
#define NLO_DIPOLE_BK_TABLE 1 /* S/dipole/amplitude.py:49-270 */

/* A dipole amplitude N(r, Y).  For a BK table the device arrays hold r[R]
 * (logarithmic grid), Y[T] = table rapidities + max(0, ln(1/x0)) and N[T x R];
 * `initialized` selects the spline set up by nlo_dis_bk_initialize at y_init
 * (with its coefficients d_init_coef, 4 (R - 1) doubles, and
 * maxr_interpolate) wherever BKDipole.interpolator_initialized would. */
typedef struct nlo_dis_dipole {
    int32_t kind;
    int32_t initialized;
    int32_t bilinear; /* BK table: AmplitudeInterpolation.LINEAR_LINEAR (else SPLINE_LINEAR) */
    int32_t reserved;
    double qs0sqr, lambda, gamma, x0; /* GBW */
    int32_t R, T;
    const double* d_r;
    const double* d_Y;
    const double* d_N;
    double y_init, maxr_interpolate;
    const double* d_init_coef;
    /* BK table, SPLINE_LINEAR: the local splines of nlo_dis_bk_local_splines
     * (T (R - 1) x 4 doubles); required by the integrand, VEGAS and batch
     * entry points; nlo_dis_dipole_amplitude solves each local spline per
     * query when it is NULL */
    const double* d_local_coef;
} nlo_dis_dipole;

/* N(r_i, Y_i) for n device points (BKDipole.dipole_amplitude /
 * GBW.dipole_amplitude, S/dipole/amplitude.py:38-43, 133-168). */
int nlo_dis_dipole_amplitude(const nlo_dis_dipole* dipole, int64_t n, const double* d_r,
                             const double* d_Y, double* d_out, hipStream_t stream);

/* The local splines of BKDipole._local_spline_amplitude
 * (S/dipole/amplitude.py:204-240) for every table row y and r interval i:
 * the cubic, in t = r - r_i, of the natural spline through the 7 nodes
 * around interval i (the last 9 at the top of the grid) of row y, on
 * interval i; d_coef: T (R - 1) x 4 doubles.  The spline is linear in its
 * data, so the amplitude at (r, Y) between rows y and y + 1 is the same
 * linear interpolation of the two rows' cubics, and a query needs no solve. */
int nlo_dis_bk_local_splines(const nlo_dis_dipole* dipole, double* d_coef, hipStream_t stream);

/* BKDipole.initialize_interpolation (S/dipole/amplitude.py:111-131, 242-260)
 * at rapidity Y: the natural spline in r through the amplitude at the table
 * nodes (first and last moved inwards by 1e-4) into d_coef (4 (R - 1)
 * doubles) and maxr_interpolate (or -1) into d_maxr (one double). */
int nlo_dis_bk_initialize(const nlo_dis_dipole* dipole, double Y, double* d_coef, double* d_maxr,
                          hipStream_t stream);

/* Integrand kinds (contributions) on the unit cube; dimension in brackets.
 * LO sums every quark of the descriptor; the others use quark 0. */
#define NLO_DIS_LO 0             /* [2] S/dis/observables.py:345-352, S/dis/lo.py:12-51 */
#define NLO_DIS_DIP_T0 1         /* [2] integrand_dip_massive "T0" (S/dis/observables.py:107-172) */
#define NLO_DIS_DIP_T1 2         /* [3] "T1"; FP64 also on the FP32 path (K1/K0 differences) */
#define NLO_DIS_DIP_T2 3         /* [4] "T2" */
#define NLO_DIS_DIP_L_CONST 4    /* [2] "Omega_L_const" */
#define NLO_DIS_DIP_L_AB 5       /* [3] "ab"; FP64 also on the FP32 path (K0 differences) */
#define NLO_DIS_DIP_L_CD 6       /* [4] "cd" */
#define NLO_DIS_QG_I1 7          /* [5] integrand_qgunsub_massive "I1" (:175-250); FP64 also on the
                                    FP32 path: its UV-subtracted sum cancels to O(z2) as z2 -> 0 */
#define NLO_DIS_QG_I2 8          /* [6] "I2" */
#define NLO_DIS_QG_I3 9          /* [7] "I3" */
#define NLO_DIS_QG_I2_NESTED 10  /* [5] integrand_qg_nested_massive "I2" (:253-331) */
#define NLO_DIS_QG_I3_NESTED 11  /* [5] "I3" */
#define NLO_DIS_KINDS 12
#define NLO_DIS_MAX_QUARKS 8

typedef struct nlo_dis_desc {
    double Q2, xbj;
    int32_t transverse;  /* 1: T, 0: L */
    int32_t n_quarks;
    double mass[NLO_DIS_MAX_QUARKS];
    double charge[NLO_DIS_MAX_QUARKS];
    double maxr;                 /* DISConfig.maxr */
    int32_t finite_nc;           /* NcScheme.FiniteNC */
    int32_t parent_coupling;     /* RunningCouplingScheme.PARENT */
    int32_t smooth_coupling;     /* RunningCouplingIRScheme.SMOOTH */
    int32_t mass_dependent_x;    /* HeavyQuarkX.MassDependentX */
    double c2_alpha, max_alpha;  /* C2_alpha, max_alpha_s_freeze */
    int32_t nf_alphas;           /* active flavours of the coupling */
    int32_t nested_points;       /* inner Gauss-Legendre points of the nested I2/I3 (8..128) */
    nlo_dis_dipole dipole;
} nlo_dis_desc;

/* Unit-cube dimension of a contribution, or -1. */
int nlo_dis_dimensions(int32_t contribution);

/* The integrand at n device points x (n x dim doubles, row-major) into d_out
 * (n doubles).  Replaces the reference's integrand callables (dis_lo_integrand,
 * dis_dipole_integrand, the I1 / nested I2-I3 kernels in K/dis/). */
int nlo_dis_integrand(const nlo_dis_desc* desc, int32_t contribution, int64_t n, const double* d_x,
                      int32_t precision, double* d_out, hipStream_t stream);

/* The nested I2 and I3 at n device points x (n x 5 doubles) from one pass
 * over the inner rule, into d_i2 and d_i3 (n doubles each); quark 0 of desc.
 * Replaces dis_nested_gbw_i2_i3_integrand_cuda / dis_nested_i2_i3_cuda
 * (K/dis/i2_i3/extension.py), the pair integrand_qg_nested_massive
 * (S/dis/observables.py:253-331) evaluates one at a time. */
int nlo_dis_qg_i23(const nlo_dis_desc* desc, int64_t n, const double* d_x, int32_t precision,
                   double* d_i2, double* d_i3, hipStream_t stream);

/* One VEGAS pass: n scrambled-Sobol points (indices first_index ..
 * first_index + n - 1 of `sobol`, dimension = the contribution's) mapped
 * through d_edges (dim x (bins + 1), on [0, 1]), the integrand evaluated and
 * w = f x Jacobian summed in FP64 with a fixed reduction order: d_sums[0] =
 * sum w, d_sums[1] = sum w^2.  d_hist / d_count (dim x bins, may both be NULL)
 * receive sum |w| and the sample count per bin.  first_index % 64 == 0.
 * Replaces the sampling, integrand and histogram steps of vegas()
 * (S/numerics/integration.py:237-499) for the DIS integrands. */
size_t nlo_dis_vegas_workspace(int32_t contribution, int32_t bins, int64_t n);
int nlo_dis_vegas(const nlo_dis_desc* desc, int32_t contribution, int32_t bins, const double* d_edges,
                  const nlo_sobol_table* sobol, int64_t first_index, int64_t n, int32_t precision,
                  double* d_sums, double* d_hist, double* d_count, void* d_workspace,
                  size_t workspace_bytes, hipStream_t stream);

/* Tensor Gauss-Legendre on the unit square for a two-dimensional
 * contribution: d_out[0] with `points` nodes per axis, d_out[1] with
 * max(2, points / 2) (tensor_gauss_legendre, S/numerics/integration.py:95-117). */
int nlo_dis_gauss_legendre(const nlo_dis_desc* desc, int32_t contribution, int32_t points,
                           int32_t precision, double* d_out, hipStream_t stream);

/* ------------------------------------------------------------ DIS batches */

/* Many DIS integrals in one call: the integrals of any number of (Q^2, x)
 * points, polarisations, quarks and terms, each with the schedule the
 * reference gives it (vegas() with its own grid, samples per iteration,
 * iteration limits and seed: S/numerics/integration.py:500-589 as driven by
 * _vegas_with_budget, S/dis/observables.py:661-686; or the tensor
 * Gauss-Legendre rule with its half-order error, :95-117).  Each pass over
 * all VEGAS integrals is one launch per contribution kind from a device
 * work-list of (integral, block); one iteration (passes, fixed-order
 * histogram reductions, grid refinement, estimate / variance, the
 * inverse-variance combination and the convergence test) is one captured
 * graph replayed on the device with no host synchronisation; converged
 * integrals skip the remaining iterations.  Replaces the per-point Python
 * loops of photon_proton_cross_section_d2b / F2 (S/dis/observables.py:387-624)
 * batched over points. */
#define NLO_DIS_METHOD_VEGAS 0
#define NLO_DIS_METHOD_GAUSS_LEGENDRE 1

typedef struct nlo_dis_item {
    nlo_dis_desc desc;        /* point, polarisation, quark(s), couplings, dipole */
    int32_t contribution;     /* NLO_DIS_* */
    int32_t method;           /* NLO_DIS_METHOD_* (Gauss-Legendre: 2-D contributions) */
    int64_t n;                /* VEGAS: samples per iteration (>= 2); Gauss-Legendre: nodes per axis (4..128) */
    int32_t sobol_table;      /* VEGAS: index into nlo_dis_batch_desc.tables (the contribution's dimension) */
    int32_t max_iterations;   /* VEGAS: 1..16 */
    int32_t min_iterations;   /* VEGAS: 1..max_iterations */
    int32_t reserved;
    int64_t first_index;      /* VEGAS: first Sobol index of iteration 0 (a multiple of 64); iteration k
                                 starts at first_index + k * (n rounded up to a multiple of 64) */
} nlo_dis_item;

typedef struct nlo_dis_batch_desc {
    int32_t n_items;
    int32_t n_tables;
    const nlo_dis_item* items;      /* host array */
    const nlo_sobol_table* tables;  /* host array */
    const double* d_initial_edges;  /* device, bins + 1: every VEGAS grid's starting edges per dimension */
    int32_t bins;                   /* VEGAS bins per dimension, 2..255 */
    int32_t precision;              /* NLO_PREC_FP32 or NLO_PREC_FP64 (the dtype of Q^2); grid smoothing
                                       and the variance floor take that dtype's epsilon and tiny */
    double epsrel, epsabs;          /* converged: error <= max(epsabs, epsrel |value|) */
} nlo_dis_batch_desc;

size_t nlo_dis_batch_workspace(const nlo_dis_batch_desc* desc);
/* d_results: n_items x 4 doubles (value, error, integrand evaluations,
 * converged 0 / 1).  Returns after the stream has finished the batch. */
int nlo_dis_batch_run(const nlo_dis_batch_desc* desc, double* d_results, void* d_workspace,
                      size_t workspace_bytes, hipStream_t stream);

/* ---------------------------------------- GBW fixed-Sobol qqbar-g estimator */

/* Paired-endpoint map of unit-cube points (dis_sobol_endpoint_map,
 * K/dis/sampling/kernel.cu:15-81, and its Python form
 * paired_endpoint_inputs, S/dis/fixed_sobol.py:116-158): each of n points
 * (n x 5, row-major) is mapped through the importance grid d_edges
 * (5 x (bins + 1)) and split into the sectors of the gluon plane nearer
 * endpoint 0 (d_first) and endpoint 1 (d_second, re-expressed about endpoint
 * 0), with weights inverse density x sector Jacobian.  Element type float
 * (NLO_PREC_FP32, the reference's dtype) or double (NLO_PREC_FP64). */
int nlo_dis_endpoint_map(int32_t precision, int64_t n, const void* d_points, const void* d_edges,
                         int32_t bins, double maxr, int32_t fold_angle, void* d_first,
                         void* d_first_weight, void* d_second, void* d_second_weight,
                         hipStream_t stream);

#define NLO_FIXED_SOBOL_MAX_ROUNDS 8
#define NLO_FIXED_SOBOL_MAX_REPLICATES 32

/* One estimator: one quark's qqbar-g term (I1 + I2 + I3) at one point and
 * polarisation.  GBW dipole; quark 0 of desc; nested_points inner nodes. */
typedef struct nlo_dis_fixed_sobol_item {
    nlo_dis_desc desc;
    int32_t table_set; /* its point sets: tables[table_set * (training_rounds + replicates) + k],
                          training round k < training_rounds, then the replicates */
    int32_t reserved;
} nlo_dis_fixed_sobol_item;

typedef struct nlo_dis_fixed_sobol_desc {
    int32_t n_items;
    int32_t n_tables;                     /* a multiple of training_rounds + replicates */
    const nlo_dis_fixed_sobol_item* items; /* host array */
    const nlo_sobol_table* tables;        /* host array of 5-D tables (make_sobol_points: seeds
                                             seed + k for training, seed + 10000 + k for replicates) */
    int32_t training_samples;             /* power of two (65 536) */
    int32_t training_rounds;              /* 1..NLO_FIXED_SOBOL_MAX_ROUNDS (3); 0 with d_grid */
    int32_t samples_per_replicate;        /* power of two (32 768) */
    int32_t replicates;                   /* 2..NLO_FIXED_SOBOL_MAX_REPLICATES (7) */
    int32_t bins;                         /* grid bins per dimension, 2..255 (32) */
    int32_t fold_angle;                   /* 1: the angle folded onto [0, pi) */
    int32_t precision;                    /* NLO_PREC_FP32 (the reference's float32 path; I1 in FP64)
                                             or NLO_PREC_FP64 */
    int32_t reserved;
    double i1_weight;                     /* weight of I1 in the training score (1.0) */
    double epsrel;                        /* converged: error <= epsrel |value| */
    const double* d_grid;                 /* NULL, or device n_items x 5 x (bins + 1) frozen grids: no
                                             training (training_rounds must be 0; evaluations count the
                                             replicates only), integrate_triple_fixed_sobol alone */
} nlo_dis_fixed_sobol_desc;

/* gbw_qg_fixed_sobol (S/dis/fixed_sobol.py:274-368) for every item at once:
 * learn_triple_importance_grid (:161-210) over the training sets, then
 * integrate_triple_fixed_sobol (:213-271) over the replicates.  d_results:
 * n_items x 4 x 4 doubles, rows I1, I2, I3 and their per-replicate sum, each
 * (mean, standard error over replicates, integrand evaluations, converged
 * 0 / 1).  d_edges (may be NULL): n_items x 5 x (bins + 1), the learned
 * grids.  Returns after the stream has finished. */
size_t nlo_dis_fixed_sobol_workspace(const nlo_dis_fixed_sobol_desc* desc);
int nlo_dis_fixed_sobol(const nlo_dis_fixed_sobol_desc* desc, double* d_results, double* d_edges,
                        void* d_workspace, size_t workspace_bytes, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif
