// BK right-hand side on gfx950: spline build, K1 fixed rule, K2 + Kf VEGAS
// sums with in-kernel Sobol points, deterministic reductions, RHS finalize and
// RK stage updates.  Reference prefixes: S/ = torch_nlo_clean/src/nlo_torch/,
// K/ = S/custom_kernels/.
#include <cmath>
#include <mutex>
#include <type_traits>

#include "bk_device.hpp"
#include "bk_internal.hpp"

namespace nlo {

BkConsts make_consts(const nlo_bk_physics& p) {
    BkConsts c{};
    c.nc = p.nc;
    c.nf = p.nf;
    c.ksub = p.ksub;
    c.minr = p.minr;
    c.b0 = (11.0 * p.nc - 2.0 * p.nf) / (12.0 * kPi);
    c.log_mu0 = 2.0 / 0.2 * std::log(2.5);
    c.scale_num = 4.0 * p.c2;
    c.lambda_qcd = p.lambda_qcd;
    c.log_scale0 = std::log(4.0 * p.c2 / (p.lambda_qcd * p.lambda_qcd));
    c.norm = 1.0 / (8.0 * kPi * kPi * kPi * kPi);
    c.symmetrize = p.symmetrize;
    c.rc_lo = p.rc_lo;
    c.rc_nlo = p.rc_nlo;
    c.resum_rc = p.resum_rc;
    c.has_resum = p.order != NLO_ORDER_LO && p.order != NLO_ORDER_NLO;
    c.slog = p.order == NLO_ORDER_NLO_RESUM_DLOG_SLOG || p.order == NLO_ORDER_LO_RESUM_DLOG_SLOG;
    c.fixed = p.rc_lo == NLO_RC_LO_FIXED || p.rc_nlo == NLO_RC_NLO_FIXED;
    c.thresholds = p.nf > 3.0;
    c.fixed_as = p.fixed_as;
    c.log_4c2 = std::log(4.0 * p.c2);
    c.production = p.rc_lo == NLO_RC_LO_BALITSKY && p.resum_rc == NLO_RESUM_RC_SMALLEST &&
                   (p.order == NLO_ORDER_NLO_RESUM_DLOG_SLOG || p.order == NLO_ORDER_LO_RESUM_DLOG_SLOG) &&
                   !c.fixed && !c.thresholds;
    return c;
}

int validate_physics(const nlo_bk_physics* p) {
    NLO_REQUIRE(p, "null physics");
    NLO_REQUIRE(p->nc > 0 && p->lambda_qcd > 0 && p->c2 > 0 && p->ksub > 0 && p->minr > 0,
                "physics constants must be positive");
    NLO_REQUIRE(p->nf == 0.0 || p->nf == 3.0 || p->nf == 5.0, "NF=%g: NF must be 0, 3 or 5", p->nf);
    NLO_REQUIRE(p->rc_lo >= NLO_RC_LO_BALITSKY && p->rc_lo <= NLO_RC_LO_BEUF, "unknown RC_LO code %d",
                p->rc_lo);
    NLO_REQUIRE(p->rc_nlo >= NLO_RC_NLO_SMALLEST && p->rc_nlo <= NLO_RC_NLO_PARENT,
                "unknown RC_NLO code %d", p->rc_nlo);
    NLO_REQUIRE(p->resum_rc >= NLO_RESUM_RC_SMALLEST && p->resum_rc <= NLO_RESUM_RC_FIXED,
                "unknown RESUM_RC code %d", p->resum_rc);
    NLO_REQUIRE(p->order >= NLO_ORDER_NLO_RESUM_DLOG_SLOG && p->order <= NLO_ORDER_LO,
                "unknown Order code %d", p->order);
    NLO_REQUIRE((p->rc_lo == NLO_RC_LO_FIXED) == (p->rc_nlo == NLO_RC_NLO_FIXED),
                "fixed coupling must be selected for both RC_LO and RC_NLO");
    NLO_REQUIRE(!(p->rc_lo == NLO_RC_LO_FIXED || p->resum_rc == NLO_RESUM_RC_FIXED) || p->fixed_as > 0,
                "FIXED_AS must be positive");
    return NLO_OK;
}

namespace {

// ------------------------------------------------------------ LDS staging

template <typename T>
__device__ void stage_spline(const T* table, int R, int row, T* x, T* rec, LdsSpline<T>& s) {
    // table: [r_min, r_max, ln r (R), N records 4 (R-1), S records 4 (R-1)]
    const T* lx = table + 2;
    const T* src = table + 2 + R + (size_t)row * 4 * (R - 1);
    for (int i = threadIdx.x; i < R; i += blockDim.x) x[i] = lx[i];
    for (int i = threadIdx.x; i < 4 * (R - 1); i += blockDim.x) rec[i] = src[i];
    s.x = x;
    s.rec = rec;
    s.n = R;
    s.x0 = lx[0];
    s.inv_h = T(R - 1) / (lx[R - 1] - lx[0]);
    s.log_min = lx[0];
    s.log_max = lx[R - 1];
}

// -------------------------------------------------------------- splines

// S/numerics/interpolation.py:27-48, :124-149 and S/bk/evolution.py:548-560.
// The interior second derivatives M_1..M_{R-2} of the N and S rows solve one
// tridiagonal system each (natural ends, same matrix); the block solves both
// by parallel cyclic reduction: every row i of a_i M_{i-1} + b_i M_i +
// c_i M_{i+1} = d_i is combined with rows i - s and i + s (s = 1, 2, 4, ..),
// which removes its neighbours, until it holds M_i alone (log2(R) steps with
// two FP64 divisions per row instead of R dependent steps).  The matrix is
// diagonally dominant, so no pivoting is needed.
constexpr int kSplineThreads = 1024;
constexpr int kSplineRows = (kMaxR + kSplineThreads - 1) / kSplineThreads;  // rows per thread

__device__ __forceinline__ void spline_build_body(int R, const double* __restrict__ r,
                                                  const double* __restrict__ N, int force_positive,
                                                  double* __restrict__ t64, float* __restrict__ t32) {
    extern __shared__ double sh[];
    double* x = sh;
    double* ly = x + R;        // 2 R
    double* so = ly + 2 * R;   // 2 R: right-hand sides, then solutions
    double* pa = so + 2 * R;   // R: sub-diagonal
    double* pb = pa + R;       // R: diagonal
    double* pc = pb + R;       // R: super-diagonal
    for (int i = threadIdx.x; i < R; i += blockDim.x) {
        x[i] = log(r[i]);
        double v = N[i] > 1.0 ? 1.0 : N[i];
        if (force_positive && v < 0.0) v = 0.0;
        ly[i] = v > 0.0 ? log(v) : -40.0;
        double s = 1.0 - N[i];
        if (s < 0.0) s = 0.0;
        if (force_positive && s > 1.0) s = 1.0;
        ly[R + i] = s > 0.0 ? log(s) : -40.0;
    }
    __syncthreads();
    const int m = R - 2;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const double hi = x[i + 1] - x[i], hi1 = x[i + 2] - x[i + 1];
        pa[i] = i > 0 ? hi : 0.0;
        pb[i] = 2.0 * (hi + hi1);
        pc[i] = i < m - 1 ? hi1 : 0.0;
        for (int row = 0; row < 2; ++row) {
            const double* y = ly + row * R;
            so[row * R + i] = 6.0 * ((y[i + 2] - y[i + 1]) / hi1 - (y[i + 1] - y[i]) / hi);
        }
    }
    __syncthreads();
    for (int st = 1; st < m; st <<= 1) {
        double na[kSplineRows], nb[kSplineRows], nc[kSplineRows], n0[kSplineRows], n1[kSplineRows];
#pragma unroll
        for (int k = 0; k < kSplineRows; ++k) {
            const int i = threadIdx.x + k * kSplineThreads;
            if (i >= m) continue;
            double a = 0.0, b = pb[i], c = 0.0, d0 = so[i], d1 = so[R + i];
            if (i - st >= 0) {
                const double al = -pa[i] / pb[i - st];
                a = al * pa[i - st];
                b += al * pc[i - st];
                d0 += al * so[i - st];
                d1 += al * so[R + i - st];
            }
            if (i + st < m) {
                const double ga = -pc[i] / pb[i + st];
                c = ga * pc[i + st];
                b += ga * pa[i + st];
                d0 += ga * so[i + st];
                d1 += ga * so[R + i + st];
            }
            na[k] = a;
            nb[k] = b;
            nc[k] = c;
            n0[k] = d0;
            n1[k] = d1;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kSplineRows; ++k) {
            const int i = threadIdx.x + k * kSplineThreads;
            if (i >= m) continue;
            pa[i] = na[k];
            pb[i] = nb[k];
            pc[i] = nc[k];
            so[i] = n0[k];
            so[R + i] = n1[k];
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        so[i] = so[i] / pb[i];
        so[R + i] = so[R + i] / pb[i];
    }
    __syncthreads();
    const int iv = R - 1;
    for (int item = threadIdx.x; item < 2 * iv; item += blockDim.x) {
        const int row = item / iv, i = item - row * iv;
        const double* y = ly + row * R;
        const double* sol = so + row * R;
        const double h = x[i + 1] - x[i];
        const double sl = i == 0 ? 0.0 : sol[i - 1];
        const double sr = i == iv - 1 ? 0.0 : sol[i];
        const double a = y[i];
        const double b = (y[i + 1] - y[i]) / h - h * (2.0 * sl + sr) / 6.0;
        const double c = sl / 2.0;
        const double d = (sr - sl) / (6.0 * h);
        const size_t o = 2 + R + (size_t)row * 4 * iv + 4 * (size_t)i;
        t64[o] = a;
        t64[o + 1] = b;
        t64[o + 2] = c;
        t64[o + 3] = d;
        if (t32) {
            // local coordinate t = (ln d - ln r_0) / h_avg - i and base-2
            // exponent: N = 2^(a' + t (b' + t (c' + t d')))
            const double hav = (x[R - 1] - x[0]) / (double)(R - 1);
            const double l2e = 1.4426950408889634;
            t32[o] = (float)(a * l2e);
            t32[o + 1] = (float)(b * hav * l2e);
            t32[o + 2] = (float)(c * hav * hav * l2e);
            t32[o + 3] = (float)(d * hav * hav * hav * l2e);
        }
    }
    for (int i = threadIdx.x; i < R; i += blockDim.x) {
        t64[2 + i] = x[i];
        if (t32) t32[2 + i] = (float)x[i];
    }
    if (threadIdx.x == 0) {
        t64[0] = r[0];
        t64[1] = r[R - 1];
        if (t32) {
            t32[0] = (float)r[0];
            t32[1] = (float)r[R - 1];
        }
    }
}

__global__ void __launch_bounds__(kSplineThreads) spline_build_kernel(int R, const double* __restrict__ r,
                                                          const double* __restrict__ N,
                                                          int force_positive, double* __restrict__ t64,
                                                          float* __restrict__ t32) {
    spline_build_body(R, r, N, force_positive, t64, t32);
}

// ------------------------------------------------------------------- K1

struct GaussRule {
    int nr, nt;
    double xr[kMaxGaussR], wr[kMaxGaussR];
    double xt[kMaxGaussT], wt[kMaxGaussT];
};

// K/bk/k1/kernel.cu:2040-2092 (production rule) for parent p = blockIdx.x:
// panels [x_i, x_{i+1}] of the log-r grid, GL(nr) in ln z, GL(nt) in v with
// theta = pi v^2, weight half_width w_r * w_t pi v, integrand * 2 z^2.  Each
// work item is one z node (panel, i) with one half of the nt theta nodes,
// whose sin^2(theta/2) and weights sit in LDS.
constexpr int kK1Threads = 1024;

// ANY = false: the production physics with the hoisted per-z-node terms;
// ANY = true: every coupling / order choice and the kinematical constraint
// (H.rows set), the full Kernel_lo per theta node.
template <bool ANY>
__device__ __forceinline__ void k1_body(int R, const double* __restrict__ r, const double* __restrict__ N,
                                        const double* __restrict__ t64, const BkConsts& c,
                                        const GaussRule& g, const K1History& H, int p,
                                        double* __restrict__ k1) {
    extern __shared__ double sh[];
    __shared__ double red[kK1Threads / 64][2];
    __shared__ double gxr[kMaxGaussR], gwr[kMaxGaussR], s2t[kMaxGaussT], wtt[kMaxGaussT];
    __shared__ double j1tab[kJ1Pieces * (kJ1Deg + 1)];
    if (p == 0 || !(N[p] <= 0.99999)) {  // reference active set: index >= 1, N <= 0.99999
        if (threadIdx.x == 0) k1[p] = 0.0;
        return;
    }
    for (int i = threadIdx.x; i < kJ1Pieces * (kJ1Deg + 1); i += blockDim.x) j1tab[i] = kJ1Cheb[i];
    LdsSpline<double> s;
    stage_spline<double>(t64, R, 0, sh, sh + R, s);
    for (int i = threadIdx.x; i < g.nr; i += blockDim.x) {
        gxr[i] = g.xr[i];
        gwr[i] = g.wr[i];
    }
    for (int j = threadIdx.x; j < g.nt; j += blockDim.x) {
        const double v = (g.xt[j] + 1.0) / 2.0;
        const double sn = sin(kPi * v * v / 2.0);
        s2t[j] = sn * sn;
        wtt[j] = g.wt[j] * kPi * v;
    }
    __syncthreads();
    const double r_min = t64[0], r_max = t64[1];
    K1Parent P;
    P.r = r[p];
    P.log_r = log(P.r);
    P.n_r = fmin(loglog64(s, r_min, r_max, P.r, 0.0, 1.0), 1.0);
    P.a_r = ANY ? 0.0 : alpha_s64(P.r, c);
    P.lo_coef = c.nc / (2.0 * kPi * kPi) * P.a_r;
    P.fin_coef = P.a_r * c.nc / (4.0 * kPi) * (67.0 / 9.0 - kPi * kPi / 3.0 - 10.0 / 9.0 * c.nf / c.nc);
    P.log_ksub = log(c.ksub);
    P.resum_on = P.r > 1.01 * c.minr ? 1.0 : 0.0;
    P.j1tab = j1tab;
    const int znodes = (R - 1) * g.nr;
    // kinematical constraint: history rows and the current rapidity (last row)
    const int rows = (ANY && H.rows) ? min(max(*H.rows, 1), H.capacity) : 0;
    const double rapidity = rows > 0 ? H.y[rows - 1] : 0.0;
    // each (z node, half of the theta nodes) is one work item: twice the
    // waves of one item per z node, so the FP64 latency is better hidden
    const int th = (g.nt + 1) / 2;
    double total = 0.0, total_err = 0.0;  // compensated
    for (int item = threadIdx.x; item < 2 * znodes; item += blockDim.x) {
        const int zi = item >> 1;
        const int j0 = (item & 1) * th, j1 = min(j0 + th, g.nt);
        const int panel = zi / g.nr;
        const int i = zi - panel * g.nr;
        const double lower = s.x[panel], upper = s.x[panel + 1];
        const double center = (lower + upper) / 2.0, half = (upper - lower) / 2.0;
        K1Z Z;
        Z.log_z = center + half * gxr[i];
        Z.z = exp(Z.log_z);
        Z.n_y = fmin(loglog_from_log(s, Z.log_z, 0.0, 1.0), 1.0);
        double acc = 0.0;
        if constexpr (ANY) {
            for (int j = j0; j < j1; ++j)
                acc += k1_theta_term_any(P, Z, s2t[j], s, c, H, rows, rapidity) * wtt[j];
        } else {
            Z.den_y = alpha_s64_den(Z.log_z, c);
            Z.a_y = 1.0 / Z.den_y;
            Z.ly = Z.log_z - P.log_r;
            Z.inv_y2 = 1.0 / (Z.z * Z.z);
            for (int j = j0; j < j1; ++j) acc += k1_theta_term(P, Z, s2t[j], s, c) * wtt[j];
        }
        comp_add(total, total_err, acc * ((2.0 * Z.z * Z.z) * (half * gwr[i])));
    }
    wave_comp_sum(total, total_err);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6][0] = total;
        red[threadIdx.x >> 6][1] = total_err;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0][0], e = red[0][1];
        for (int w = 1; w < kK1Threads / 64; ++w) comp_merge(t, e, red[w][0], red[w][1]);
        k1[p] = t + e;
    }
}

template <bool ANY>
__global__ void __launch_bounds__(kK1Threads) k1_kernel(int R, const double* __restrict__ r,
                                                        const double* __restrict__ N,
                                                        const double* __restrict__ t64, BkConsts c,
                                                        GaussRule g, K1History H,
                                                        double* __restrict__ k1) {
    k1_body<ANY>(R, r, N, t64, c, g, H, blockIdx.x, k1);
}

// ------------------------------------------------------ K1, adaptive rule

// rapidity_derivative_lo (S/bk/derivatives.py:43-110): adaptive
// Gauss-Kronrod 21 (adaptive_gauss_kronrod_1d, S/numerics/integration.py:
// 43-92; the rule and its error of :624-746) in ln z over [ln r_0, ln r_{R-1}]
// of 2 z^2 times the adaptive rule in theta over [0, pi].  One block per
// parent: thread 0 keeps the outer interval list; every outer node is an
// inner integral, taken by the next free wave, whose 64 lanes evaluate the
// 21 (first interval) or 42 (two halves) theta nodes of each bisection step
// while every lane runs the same list bookkeeping on LDS.  Interval lists
// stay in order and sums run in list order, as in the reference.
constexpr int kK1aWaves = 8;

// QUADPACK's 21-point Kronrod extension of the 10-point Gauss rule: nodes
// kGkX[0..9] (kGkX[1, 3, .., 9] are the Gauss nodes), centre node last.
__device__ constexpr double kGkX[10] = {
    0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
    0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
    0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
    0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
    0.294392862701460198131126603103866, 0.148874338981631210884826001129720};
__device__ constexpr double kGkW[10] = {
    0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
    0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
    0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
    0.123491976262065851077958109831074, 0.134709217311473325928054001771707,
    0.142775938577060080797094273138717, 0.147739104901338491374841515972068};
constexpr double kGkWCenter = 0.149445554002916905664936468389821;
__device__ constexpr double kGW[5] = {
    0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
    0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
    0.295524224714752870173892994651338};

// Node k (0..20) of [a, b]: centre - h x_k, k < 10; centre, k = 10; centre + h x_{k-11}.
__device__ __forceinline__ double gk_node(double a, double b, int k) {
    const double center = (a + b) / 2.0, half = (b - a) / 2.0;
    if (k < 10) return center - half * kGkX[k];
    if (k == 10) return center;
    return center + half * kGkX[k - 11];
}

// The estimate and error of _gauss_kronrod_intervals from the 21 values v
// (node order of gk_node): min(1, (200 |K - G| / resasc)^1.5) resasc, at
// least 50 eps resabs.
__device__ __forceinline__ void gk_from_values(double a, double b, const double* v, double& est,
                                               double& err) {
    const double half = (b - a) / 2.0;
    const double fc = v[10];
    double kron = 0.0, gauss = 0.0;
    for (int i = 0; i < 10; ++i) kron += (v[i] + v[11 + i]) * kGkW[i];
    kron += kGkWCenter * fc;
    for (int i = 0; i < 5; ++i) gauss += (v[2 * i + 1] + v[12 + 2 * i]) * kGW[i];
    const double mean = kron / 2.0;
    double sabs = 0.0, sasc = 0.0;
    for (int i = 0; i < 10; ++i) {
        sabs += (fabs(v[i]) + fabs(v[11 + i])) * kGkW[i];
        sasc += (fabs(v[i] - mean) + fabs(v[11 + i] - mean)) * kGkW[i];
    }
    const double resabs = half * (sabs + kGkWCenter * fabs(fc));
    const double resasc = half * (sasc + kGkWCenter * fabs(fc - mean));
    double e = half * fabs(kron - gauss);
    // (200 |K - G| / resasc)^1.5 as q sqrt(q): within an ulp of pow(q, 1.5)
    // at the cost of one square root
    const double q = 200.0 * e / (resasc > 0.0 ? resasc : 1.0);
    const double scale = q * sqrt(q);
    if (resasc > 0.0 && e > 0.0) e = resasc * fmin(scale, 1.0);
    est = half * kron;
    err = fmax(e, 50.0 * 2.220446049250313e-16 * resabs);
}

// Ordered interval list in LDS: lower, upper, estimate, error (cap entries each).
struct GkList {
    double* lo;
    double* hi;
    double* est;
    double* err;
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The summed estimate and error of a list in list order, and its first
// interval of largest error, by a whole wave: lanes 0-31 sum the estimates,
// lanes 32-63 the errors and find the first largest one; lanes 0 and 32 hand
// the results to the wave.
__device__ __forceinline__ void gk_scan_wave(const GkList& L, int count, double& value, double& error,
                                             int& split, int lane) {
    const bool upper = lane >= 32;
    const double* x = upper ? L.err : L.est;
    double acc = 0.0, best = L.err[0];
    int arg = 0;
    for (int i = 0; i < count; ++i) {
        const double xi = x[i];
        acc += xi;
        if (xi > best) {  // err[i] > err[arg] on the upper half
            best = xi;
            arg = i;
        }
    }
    value = __shfl(acc, 0, 64);
    error = __shfl(acc, 32, 64);
    split = __shfl(arg, 32, 64);
}

// The rule on [a, mid] (values v[0..20]) and on [mid, b] (v[21..41]) by a
// whole wave: lanes 0-31 sum the first, lanes 32-63 the second; lanes 0 and
// 32 hand the results to the wave.
__device__ __forceinline__ void gk_halves_wave(double a, double mid, double b, const double* v, double& e0,
                                               double& r0, double& e1, double& r1, int lane) {
    const bool upper = lane >= 32;
    double eh, rh;
    gk_from_values(upper ? mid : a, upper ? b : mid, upper ? v + 21 : v, eh, rh);
    e0 = __shfl(eh, 0, 64);
    r0 = __shfl(rh, 0, 64);
    e1 = __shfl(eh, 32, 64);
    r1 = __shfl(rh, 32, 64);
}

// Replace entry `split` of a list of `count` (< cap <= 128) entries by its
// halves [lo, mid] and [mid, hi] (estimates e0, e1; errors r0, r1), moving the
// entries after it one place up.  Every lane of one wave calls it: each lane
// reads up to two of the entries to move, the fence orders all reads before
// the writes, then lane 0 writes the two halves.
__device__ __forceinline__ void gk_insert_wave(const GkList& L, int count, int split, double mid,
                                               double e0, double r0, double e1, double r1, int lane) {
    const double l = L.lo[split], u = L.hi[split];
    double m[2][4];
    int idx[2];
    for (int k = 0; k < 2; ++k) {
        idx[k] = split + 1 + lane + 64 * k;
        if (idx[k] < count) {
            m[k][0] = L.lo[idx[k]];
            m[k][1] = L.hi[idx[k]];
            m[k][2] = L.est[idx[k]];
            m[k][3] = L.err[idx[k]];
        }
    }
    wave_sync();
    for (int k = 0; k < 2; ++k) {
        if (idx[k] < count) {
            L.lo[idx[k] + 1] = m[k][0];
            L.hi[idx[k] + 1] = m[k][1];
            L.est[idx[k] + 1] = m[k][2];
            L.err[idx[k] + 1] = m[k][3];
        }
    }
    if (lane == 0) {
        L.lo[split] = l;
        L.hi[split] = mid;
        L.est[split] = e0;
        L.err[split] = r0;
        L.lo[split + 1] = mid;
        L.hi[split + 1] = u;
        L.est[split + 1] = e1;
        L.err[split + 1] = r1;
    }
}

// The per-daughter terms of the K1 integrand at z = e^{log_z} (K1Z).
template <bool ANY>
__device__ __forceinline__ K1Z k1_daughter(const K1Parent& P, double log_z, const LdsSpline<double>& s,
                                           const BkConsts& c) {
    K1Z Z;
    Z.log_z = log_z;
    Z.z = exp(log_z);
    Z.n_y = fmin(loglog_from_log(s, Z.log_z, 0.0, 1.0), 1.0);
    if constexpr (!ANY) {
        Z.den_y = alpha_s64_den(Z.log_z, c);
        Z.a_y = 1.0 / Z.den_y;
        Z.ly = Z.log_z - P.log_r;
        Z.inv_y2 = 1.0 / (Z.z * Z.z);
    }
    return Z;
}

// _lo_integrand (S/bk/derivatives.py:434-468) at angle theta.
template <bool ANY>
__device__ __forceinline__ double k1_theta_value(const K1Parent& P, const K1Z& Z, double theta,
                                                 const LdsSpline<double>& s, const BkConsts& c,
                                                 const K1History& H, int rows, double rapidity) {
    const double sn = sin(0.5 * theta);
    if constexpr (ANY)
        return k1_theta_term_any(P, Z, sn * sn, s, c, H, rows, rapidity);
    else
        return k1_theta_term(P, Z, sn * sn, s, c);
}

// The theta integral at one daughter size z (terms Z) for parent P, by one
// wave, from the rule on [0, pi] of the block's first pass (estimate e,
// error r); each bisection
// of [a, b] evaluates the 21 nodes of [a, mid] (lanes 0-20) and of [mid, b]
// (lanes 21-41).
template <bool ANY>
__device__ double k1_theta_adaptive(const K1Parent& P, const K1Z& Z, double e, double r,
                                    const LdsSpline<double>& s, const BkConsts& c, const K1History& H,
                                    int rows, double rapidity, int tint, double epsrel, const GkList& L,
                                    double* v, int lane, long long& evals) {
    if (lane == 0) {
        L.lo[0] = 0.0;
        L.hi[0] = kPi;
        L.est[0] = e;
        L.err[0] = r;
    }
    evals += 21;
    int count = 1, split = 0;
    double value = e;
    for (;;) {
        wave_sync();
        double error;
        gk_scan_wave(L, count, value, error, split, lane);
        if (count >= tint || error <= epsrel * fabs(value)) break;
        const double a = L.lo[split], b = L.hi[split], mid = (a + b) / 2.0;
        if (lane < 42) {
            const double x = lane < 21 ? gk_node(a, mid, lane) : gk_node(mid, b, lane - 21);
            v[lane] = k1_theta_value<ANY>(P, Z, x, s, c, H, rows, rapidity);
        }
        wave_sync();
        double e0, r0, e1, r1;
        gk_halves_wave(a, mid, b, v, e0, r0, e1, r1, lane);
        gk_insert_wave(L, count, split, mid, e0, r0, e1, r1, lane);
        ++count;
        evals += 42;
    }
    wave_sync();
    return value;
}

// LDS of the adaptive K1 kernel (doubles): spline (R nodes + 4 (R - 1)
// coefficients), J1 table, outer list (4 cap), outer nodes and values (2 x
// 42), the first-pass values of the pending inner integrals (42 x 21), their
// daughter terms (42 K1Z) and first-pass estimates and errors (2 x 42), then
// per wave a list (4 cap) and 42 node values.
constexpr int kK1ZWords = sizeof(K1Z) / sizeof(double);
static_assert(sizeof(K1Z) == kK1ZWords * sizeof(double), "K1Z is a record of doubles");
__host__ __device__ inline size_t k1a_lds_doubles(int R, int cap) {
    return (size_t)R + 4 * (size_t)(R - 1) + kJ1Pieces * (kJ1Deg + 1) + 4 * (size_t)cap + 84 + 42 * 21 +
           42 * kK1ZWords + 84 + (size_t)kK1aWaves * (4 * (size_t)cap + 42);
}

// Parent p of one evolution (arrays of that evolution).
template <bool ANY>
__device__ __forceinline__ void k1_adaptive_body(int R, const double* __restrict__ r,
                                                 const double* __restrict__ N,
                                                 const double* __restrict__ t64, const BkConsts& c,
                                                 const K1History& H, int rint, int tint, int cap,
                                                 double epsrel, int p, double* __restrict__ k1,
                                                 double* __restrict__ k1_err,
                                                 long long* __restrict__ k1_evals) {
    extern __shared__ double sh[];
    __shared__ int task_next, n_tasks, o_done, n_open;
    __shared__ int open_task[42];  // the inner integrals the first pass leaves open
    __shared__ unsigned long long ev_total;
    if (p == 0 || !(N[p] <= 0.99999)) {  // reference active set: index >= 1, N <= 0.99999
        if (threadIdx.x == 0) {
            k1[p] = 0.0;
            if (k1_err) k1_err[p] = 0.0;
            if (k1_evals) k1_evals[p] = 0;
        }
        return;
    }
    double* j1tab = sh + R + 4 * (R - 1);
    double* o_base = j1tab + kJ1Pieces * (kJ1Deg + 1);
    const GkList O{o_base, o_base + cap, o_base + 2 * cap, o_base + 3 * cap};
    double* o_node = o_base + 4 * cap;  // 42 ln z nodes
    double* o_val = o_node + 42;        // 42 outer integrand values
    double* f_val = o_val + 42;         // 42 x 21 first-pass theta values
    K1Z* z_tab = reinterpret_cast<K1Z*>(f_val + 42 * 21);  // daughter terms of the 42 nodes
    double* f_est = f_val + 42 * 21 + 42 * kK1ZWords;       // first-pass estimates
    double* f_err = f_est + 42;                              // and errors
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double* w_base = f_err + 42 + (size_t)wave * (4 * cap + 42);
    const GkList W{w_base, w_base + cap, w_base + 2 * cap, w_base + 3 * cap};
    double* w_val = w_base + 4 * cap;

    for (int i = threadIdx.x; i < kJ1Pieces * (kJ1Deg + 1); i += blockDim.x) j1tab[i] = kJ1Cheb[i];
    LdsSpline<double> s;
    stage_spline<double>(t64, R, 0, sh, sh + R, s);
    const double a0 = t64[2], b0 = t64[2 + R - 1];  // ln r_0, ln r_{R-1}
    if (threadIdx.x == 0) {
        for (int k = 0; k < 21; ++k) o_node[k] = gk_node(a0, b0, k);
        n_tasks = 21;
        task_next = 0;
        o_done = 0;
        ev_total = 0ull;
    }
    __syncthreads();
    K1Parent P;
    P.r = r[p];
    P.log_r = log(P.r);
    P.n_r = fmin(loglog64(s, t64[0], t64[1], P.r, 0.0, 1.0), 1.0);
    P.a_r = ANY ? 0.0 : alpha_s64(P.r, c);
    P.lo_coef = c.nc / (2.0 * kPi * kPi) * P.a_r;
    P.fin_coef = P.a_r * c.nc / (4.0 * kPi) * (67.0 / 9.0 - kPi * kPi / 3.0 - 10.0 / 9.0 * c.nf / c.nc);
    P.log_ksub = log(c.ksub);
    P.resum_on = P.r > 1.01 * c.minr ? 1.0 : 0.0;
    P.j1tab = j1tab;
    const int rows = (ANY && H.rows) ? min(max(*H.rows, 1), H.capacity) : 0;
    const double rapidity = rows > 0 ? H.y[rows - 1] : 0.0;

    int count = 0;  // outer intervals (wave 0 keeps the outer list; its lanes hold equal copies)
    int split = 0;
    double value = 0.0, error = 0.0;
    long long evals = 0;
    for (int step = 0; step < rint; ++step) {
        // the daughter terms of every pending inner integral, once
        for (int t = threadIdx.x; t < n_tasks; t += blockDim.x) z_tab[t] = k1_daughter<ANY>(P, o_node[t], s, c);
        if (threadIdx.x == 0) n_open = 0;
        __syncthreads();
        // first pass on [0, pi] of every pending inner integral, over the
        // whole block (n_tasks x 21 theta nodes)
        for (int i = threadIdx.x; i < n_tasks * 21; i += blockDim.x) {
            const int t = i / 21, k = i - 21 * t;
            const K1Z Z = z_tab[t];
            f_val[i] = k1_theta_value<ANY>(P, Z, gk_node(0.0, kPi, k), s, c, H, rows, rapidity);
        }
        __syncthreads();
        // the rule on [0, pi] of every pending inner integral, one thread each;
        // an integral that meets its tolerance there (the first test of
        // k1_theta_adaptive, on the list's sums 0 + e and 0 + r) is done, the
        // others go to the bisections
        for (int t = threadIdx.x; t < n_tasks; t += blockDim.x) {
            double e, r;
            gk_from_values(0.0, kPi, f_val + 21 * t, e, r);
            const double value = 0.0 + e, err = 0.0 + r;
            if (1 >= tint || err <= epsrel * fabs(value)) {
                o_val[t] = value * exp(2.0 * o_node[t]) * 2.0;
                atomicAdd(&ev_total, 21ull);
            } else {
                f_est[t] = e;
                f_err[t] = r;
                open_task[atomicAdd(&n_open, 1)] = t;
            }
        }
        __syncthreads();
        // the bisections of each inner integral, one wave each
        for (;;) {
            int t = 0;
            if (lane == 0) t = atomicAdd(&task_next, 1);
            t = __shfl(t, 0, 64);
            if (t >= n_open) break;
            t = open_task[t];
            const double lz = o_node[t];
            const K1Z Z = z_tab[t];
            const double inner = k1_theta_adaptive<ANY>(P, Z, f_est[t], f_err[t], s, c, H, rows, rapidity,
                                                        tint, epsrel, W, w_val, lane, evals);
            if (lane == 0) o_val[t] = inner * exp(2.0 * lz) * 2.0;
        }
        __syncthreads();
        if (wave == 0) {
            if (count == 0) {
                double e, rr;
                gk_from_values(a0, b0, o_val, e, rr);
                if (lane == 0) {
                    O.lo[0] = a0;
                    O.hi[0] = b0;
                    O.est[0] = e;
                    O.err[0] = rr;
                }
                count = 1;
            } else {
                const double a = O.lo[split], b = O.hi[split], mid = (a + b) / 2.0;
                double e0, r0, e1, r1;
                gk_halves_wave(a, mid, b, o_val, e0, r0, e1, r1, lane);
                gk_insert_wave(O, count, split, mid, e0, r0, e1, r1, lane);
                ++count;
            }
            wave_sync();
            gk_scan_wave(O, count, value, error, split, lane);
            if (count >= rint || error <= epsrel * fabs(value)) {
                if (lane == 0) o_done = 1;
            } else {
                const double a = O.lo[split], b = O.hi[split], mid = (a + b) / 2.0;
                if (lane < 21) {
                    o_node[lane] = gk_node(a, mid, lane);
                    o_node[21 + lane] = gk_node(mid, b, lane);
                }
                if (lane == 0) {
                    n_tasks = 42;
                    task_next = 0;
                }
            }
        }
        __syncthreads();
        if (o_done) break;
    }
    if (lane == 0) atomicAdd(&ev_total, (unsigned long long)evals);
    __syncthreads();
    if (threadIdx.x == 0) {
        k1[p] = value;
        if (k1_err) k1_err[p] = error;
        if (k1_evals) k1_evals[p] = (long long)ev_total;
    }
}

template <bool ANY>
__global__ void __launch_bounds__(kK1aWaves * 64)
    k1_adaptive_kernel(int R, const double* __restrict__ r, const double* __restrict__ N,
                       const double* __restrict__ t64, BkConsts c, K1History H, int rint, int tint,
                       int cap, double epsrel, double* __restrict__ k1, double* __restrict__ k1_err,
                       long long* __restrict__ k1_evals) {
    k1_adaptive_body<ANY>(R, r, N, t64, c, H, rint, tint, cap, epsrel, blockIdx.x, k1, k1_err, k1_evals);
}

// ------------------------------------------------------------- K2 + Kf

// LDS layout of the NLO kernel (bytes), shared by host sizing and device.
struct NloLayout {
    size_t x, rec, edges, dir, absw, bins, hist, cnt, total;
};

__host__ __device__ inline NloLayout nlo_layout(int R, int bins, int spt, bool hist, size_t tsize) {
    NloLayout L{};
    size_t o = 0;
    L.x = o;
    if (tsize == sizeof(double)) o = align_up(o + (size_t)R * tsize, 16);  // FP32 needs no nodes
    L.rec = o;  // FP64: R - 1 records; FP32: R + 1 (with the two range records)
    o = align_up(o + (size_t)4 * (tsize == sizeof(double) ? R - 1 : R + 1) * tsize, 16);
    L.edges = o;  // FP64: 4 x (bins + 1) edges; FP32: 4 x bins (left, width) pairs
    o = align_up(o + (size_t)4 * (bins + 1) * sizeof(double), 16);
    L.dir = o;
    o = align_up(o + (size_t)4 * 30 * sizeof(uint32_t), 16);
    if (hist) {
        L.absw = o;
        o = align_up(o + (size_t)256 * spt * sizeof(double), 16);
        L.bins = o;
        o = align_up(o + (size_t)256 * spt * sizeof(uint32_t), 16);
        L.hist = o;
        o = align_up(o + (size_t)4 * bins * sizeof(unsigned long long), 16);
        L.cnt = o;
        o = align_up(o + (size_t)4 * bins * sizeof(unsigned int), 16);
    }
    L.total = o;
    return L;
}

struct NloArgs {
    int R, bins, spt, bpp;
    int64_t n, first;
    const double* r;
    const double* N;
    const double* t64;
    const float* t32;
    const double* edges;
    double* part;       // (R - 1) x bpp x 4
    double* part_hist;  // (R - 1) x bpp x 4 x bins
    double* part_cnt;
    BkConsts c;
    uint32_t dir[4][30];
    uint32_t shift[4];
    uint32_t first_q[4];
};

// One evolution's arrays as an NLO block reads them (a single run: NloArgs'
// own; a batch: the run's slices of the concatenated arrays).
struct RunView {
    int R;
    const double* r;
    const double* N;
    const double* t64;
    const float* t32;
    const double* edges;
    double* part;
};

// Per (parent, block): spt samples per lane.  Sample ordinal
// s = (blk * spt * 4 + j * 4 + wave) * 64 + lane, Sobol index first + s, so
// lanes of a wave hold 64 consecutive indices and the gray-code state splits
// into a wave-uniform high part and a per-lane low part.
// FULL = false (the evolver): only sum w is accumulated; the sum of w^2 and
// the K2 / Kf split read as 0.
// ANYC: any coupling choice (NloCoupling modes); false = the production
// coupling only (smooth, smallest distance), with no other code paths.
template <int MODE, bool HIST, bool FULL, bool ANYC>
__device__ __forceinline__ void nlo_body(const NloArgs& a, const RunView& v, int p, int blk) {
    using T = typename std::conditional<MODE == NLO_PREC_FP32, float, double>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[4][FULL ? 4 : 1][2];  // [wave][sum][value, error]
    __shared__ double wmax[4];
    const size_t pidx = (size_t)(p - 1) * a.bpp + blk;
    const NloLayout lay = nlo_layout(v.R, a.bins, a.spt, HIST, sizeof(T));
    if (!(v.N[p] <= 0.99999)) {
        if (threadIdx.x < 4) v.part[pidx * 4 + threadIdx.x] = 0.0;
        if (HIST)
            for (int k = threadIdx.x; k < 4 * a.bins; k += blockDim.x) {
                a.part_hist[pidx * 4 * a.bins + k] = 0.0;
                a.part_cnt[pidx * 4 * a.bins + k] = 0.0;
            }
        return;
    }
    unsigned char* sx = smem + lay.x;
    unsigned char* srec = smem + lay.rec;
    T* E = reinterpret_cast<T*>(smem + lay.edges);
    uint32_t* dir = reinterpret_cast<uint32_t*>(smem + lay.dir);
    LdsSpline<double> spl{};
    FastSpline fsp{};
    if constexpr (MODE == NLO_PREC_FP32) {
        // N records (row 0) of the local-coordinate FP32 table, 16 B each,
        // between a constant record below the grid (N = 2^-126) and one above
        // it (N = 1), so that out-of-range lookups need no selects
        float4* rec4 = reinterpret_cast<float4*>(srec);
        const float4* src = reinterpret_cast<const float4*>(v.t32 + 2 + v.R);
        const bool aligned = ((reinterpret_cast<uintptr_t>(v.t32) + (2 + v.R) * 4) & 15) == 0;
        for (int i = threadIdx.x; i < v.R - 1; i += blockDim.x) {
            if (aligned) {
                rec4[i + 1] = src[i];
            } else {
                const float* f = v.t32 + 2 + v.R + 4 * i;
                rec4[i + 1] = make_float4(f[0], f[1], f[2], f[3]);
            }
        }
        if (threadIdx.x == 0) {
            rec4[0] = make_float4(-126.0f, 0.0f, 0.0f, 0.0f);
            rec4[v.R] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        const double x0 = v.t64[2], xl = v.t64[2 + v.R - 1];
        const double inv_h = (double)(v.R - 1) / (xl - x0);
        fsp.rec = rec4;
        fsp.inv_h2 = (float)(inv_h * kLn2);  // the integrand's logarithms are base 2
        fsp.off = (float)(1.0 - x0 * inv_h);
        fsp.umax = (float)v.R;
    } else {
        // FP64 reference formula reads the S spline (row 1); the stable form reads N (row 0).
        stage_spline<double>(v.t64, v.R, MODE == NLO_PREC_FP64 ? 1 : 0, reinterpret_cast<double*>(sx),
                             reinterpret_cast<double*>(srec), spl);
    }
    const double r0 = v.r[0], rmax = v.r[v.R - 1];
    // nlo_vegas_bounds (S/bk/derivatives.py:424-431)
    const double lo_r = log(0.5 * r0);
    const double w_r = log(2.0 * rmax) - lo_r;
    const int ne = 4 * (a.bins + 1);
    const double* ep = v.edges + (size_t)p * ne;
    if constexpr (MODE == NLO_PREC_FP32) {
        // Per bin, the absolute left end lo + w left and width w (right - left)
        // of the mapped interval, formed in FP64 and rounded once: the sample
        // is then x = X_b + W_b frac with one more rounding, where forming
        // lo + w (left + width frac) in FP32 cancels (lo = ln(r_0 / 2) against
        // the scaled edge) and leaves several ulp in ln z.  The two log-r
        // dimensions are tabulated in base 2 (times log2 e), so the sample is
        // t = log2 z and z = 2^t exactly up to the hardware exp2; the Jacobian
        // then carries (log2 e)^2, which the volume factor takes back.  Wave 0
        // alone builds the table, so that only one wave of the block evaluates
        // the two FP64 logarithms of the bounds.
        float2* E2 = reinterpret_cast<float2*>(smem + lay.edges);
        if (threadIdx.x < kWave) {
            for (int i = threadIdx.x; i < 4 * a.bins; i += kWave) {
                const int d = i / a.bins, b = i - d * a.bins;
                const double lo_d = d < 2 ? lo_r : 0.0, w_d = d < 2 ? w_r : 2.0 * kPi;
                const double unit = d < 2 ? kLog2e : 1.0;
                const double left = ep[d * (a.bins + 1) + b], right = ep[d * (a.bins + 1) + b + 1];
                E2[i] = make_float2((float)((lo_d + w_d * left) * unit), (float)(w_d * (right - left) * unit));
            }
        }
    } else {
        for (int i = threadIdx.x; i < ne; i += blockDim.x) E[i] = ep[i];
    }
    for (int i = threadIdx.x; i < 120; i += blockDim.x) dir[i] = a.dir[i / 30][i % 30];
    unsigned long long* hh = nullptr;
    unsigned int* hc = nullptr;
    double* st_w = nullptr;
    uint32_t* st_b = nullptr;
    if (HIST) {
        hh = reinterpret_cast<unsigned long long*>(smem + lay.hist);
        hc = reinterpret_cast<unsigned int*>(smem + lay.cnt);
        st_w = reinterpret_cast<double*>(smem + lay.absw);
        st_b = reinterpret_cast<uint32_t*>(smem + lay.bins);
        for (int k = threadIdx.x; k < 4 * a.bins; k += blockDim.x) {
            hh[k] = 0ull;
            hc[k] = 0u;
        }
    }
    __syncthreads();

    const double rp = v.r[p];
    const T lo[4] = {(T)lo_r, (T)lo_r, T(0), T(0)};
    const T wd[4] = {(T)w_r, (T)w_r, (T)(2.0 * kPi), (T)(2.0 * kPi)};
    T vol = T(1);
    if constexpr (MODE == NLO_PREC_FP32) {
        // the widths W_b already carry the volume; (ln 2)^2 undoes the base-2
        // units of the two log-r widths (one rounding of B^4 (ln 2)^2)
        const double bb = (double)a.bins;
        vol = (T)((bb * bb) * (bb * bb) * (kLn2 * kLn2));
    } else {
        for (int d = 0; d < 4; ++d) vol = vol * wd[d];
    }
    const T r_t = (T)rp;
    // ln r_p as the spline table holds it (the same FP64 log of r_p)
    const double log_rp = v.t64[2 + p];
    // the FP32 integrand takes log2 r
    const T log_r_t = MODE == NLO_PREC_FP32 ? (T)(log_rp * kLog2e) : (T)log_rp;
    NloCoupling<float> cpl32;
    NloCoupling<double> cpl64;
    if constexpr (ANYC) {
        cpl32 = nlo_coupling<float>(rp, a.c);
        cpl64 = nlo_coupling<double>(rp, a.c);
    } else {
        cpl32 = NloCoupling<float>{0, (float)a.c.b0, (float)a.c.log_mu0, (float)a.c.log_scale0, 0.0f, 0.0f};
        cpl64 = NloCoupling<double>{0, a.c.b0, a.c.log_mu0, a.c.log_scale0, 0.0, 0.0};
    }
    const double r_min64 = v.t64[0], r_max64 = v.t64[1];

    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const uint32_t glo0 = (uint32_t)(lane ^ (lane >> 1));
    uint32_t vlo[4][2];
    for (int d = 0; d < 4; ++d) {
        uint32_t v0 = 0, v1;
        for (int b = 0; b < 6; ++b)
            if ((glo0 >> b) & 1u) v0 ^= dir[d * 30 + b];
        v1 = v0 ^ dir[d * 30 + 5];
        vlo[d][0] = v0 ^ a.shift[d];
        vlo[d][1] = v1 ^ a.shift[d];
    }

    // compensated per-lane sums (value, error): w, w^2, K2 and Kf parts
    double s1 = 0.0, s2 = 0.0, sk2 = 0.0, skf = 0.0;
    double c1 = 0.0, c2 = 0.0, ck2 = 0.0, ckf = 0.0;
    // Each wave walks spt consecutive 64-point groups hi; the wave-uniform
    // high part of the gray code (direction bits 6..29) is built once and
    // then updated by one XOR per group: gray(h) ^ gray(h - 1) = lowest set
    // bit of h.
    const int64_t hi0 = (a.first >> 6) + ((int64_t)blk * 4 + wave) * a.spt;
    // vh[d]: the wave-uniform part of the state, direction bits 6..29 of
    // gray(hi) plus bit 5 of the parity of hi (the one lane-independent bit of
    // the low part), read from the kernel arguments so it stays scalar
    uint32_t vhi[4] = {0u, 0u, 0u, 0u};
    {
        uint64_t g = (uint64_t)hi0 ^ ((uint64_t)hi0 >> 1);
        for (int b = 6; g && b < 30; ++b, g >>= 1)
            if (g & 1u)
                for (int d = 0; d < 4; ++d) vhi[d] ^= a.dir[d][b];
    }
    // One 64-point group per call.  FIRST: the group holding Sobol index 0
    // (hi == 0), whose lane 0 takes the float64 first point and may carry
    // q = 2^30 (the clamped bin); only hi0 == 0, j == 0 is that group, so the
    // steady-state loop carries no first-point test.
    auto group = [&](int j, auto first_tag) {
        constexpr bool FIRST = decltype(first_tag)::value;
        const int64_t hi = hi0 + j;
        if (j > 0) {
            const int b = 6 + __builtin_ctzll((unsigned long long)hi);
            if (b < 30)
                for (int d = 0; d < 4; ++d) vhi[d] ^= a.dir[d][b];
        }
        const int64_t sample = (hi << 6) + lane - a.first;
        const bool valid = sample < a.n;
        const bool odd = (hi & 1) != 0;
        uint32_t q[4];
        for (int d = 0; d < 4; ++d) q[d] = vlo[d][0] ^ (odd ? (vhi[d] ^ a.dir[d][5]) : vhi[d]);
        if (FIRST && lane == 0)
            for (int d = 0; d < 4; ++d) q[d] = a.first_q[d];
        T x[4];
        T jac = T(1);
        uint32_t packed = 0u;
        for (int d = 0; d < 4; ++d) {
            // bin = min(floor(u B), B - 1), frac = u B - bin with u = q 2^-30:
            // for q < 2^30, 4 q B = bin 2^32 + frac 2^32 (high and low words of
            // one 32 x 32 product); q = 2^30 (only a rounded first point) is
            // the clamped case bin = B - 1, frac = 1.
            const uint32_t q4 = q[d] << 2;
            uint32_t b = __umulhi(q4, (uint32_t)a.bins);
            T frac = (T)(q4 * (uint32_t)a.bins) * (T)0x1p-32;
            if (FIRST && q[d] >= (1u << 30)) {
                b = (uint32_t)(a.bins - 1);
                frac = T(1);
            }
            if constexpr (MODE == NLO_PREC_FP32) {
                const float2 e = reinterpret_cast<const float2*>(E)[d * a.bins + b];
                x[d] = fmaf(e.y, frac, e.x);
                jac = jac * e.y;
            } else {
                const T left = E[d * (a.bins + 1) + b], right = E[d * (a.bins + 1) + b + 1];
                const T w = right - left;
                x[d] = lo[d] + wd[d] * (left + w * frac);
                jac = jac * ((T)a.bins * w);
            }
            packed |= b << (8 * d);
        }
        jac = vol * jac;
        double wv = 0.0, k2 = 0.0, kf = 0.0;
        if (valid) {
            if constexpr (MODE == NLO_PREC_FP64) {
                const NloValue64 f = nlo_integrand64<ANYC>(rp, x[0], x[1], x[2], x[3], spl, r_min64,
                                                           r_max64, a.c);
                wv = f.total * (double)jac;
                k2 = f.k2 * (double)jac;
                kf = f.kf * (double)jac;
            } else if constexpr (MODE == NLO_PREC_FP32) {
                const NloStable<float> f = nlo_integrand_fp32<ANYC>(
                    r_t, log_r_t, x[0], x[1], x[2], x[3], fsp, cpl32, (float)(a.c.nf / a.c.nc),
                    (float)a.c.nc, (float)a.c.norm, a.c.symmetrize != 0);
                wv = (double)(f.total * jac);
                k2 = (double)(f.k2 * jac);
                kf = (double)(f.kf * jac);
            } else {
                const NloStable<double> f = nlo_integrand_stable<double, ANYC>(
                    rp, log_rp, x[0], x[1], x[2], x[3], spl, cpl64, a.c.nf / a.c.nc, a.c.nc,
                    a.c.norm, a.c.symmetrize != 0);
                wv = f.total * jac;
                k2 = f.k2 * jac;
                kf = f.kf * jac;
            }
            comp_add(s1, c1, wv);
            if constexpr (FULL) {
                comp_add(s2, c2, wv * wv);
                comp_add(sk2, ck2, k2);
                comp_add(skf, ckf, kf);
            }
        }
        if (HIST) {
            st_w[threadIdx.x + j * 256] = valid ? fabs(wv) : 0.0;
            st_b[threadIdx.x + j * 256] = valid ? packed : 0xffffffffu;
        }
    };
    int j0 = 0;
    if (hi0 == 0) {
        group(0, std::true_type{});
        j0 = 1;
    }
    for (int j = j0; j < a.spt; ++j) group(j, std::false_type{});
    // the block's partials, each one rounding of its compensated pair
    if constexpr (FULL) {
        double t[4] = {s1, s2, sk2, skf}, e[4] = {c1, c2, ck2, ckf};
        block_comp_sums<4>(t, e, red, v.part + pidx * 4);
    } else {  // the other three read 0
        double t[1] = {s1}, e[1] = {c1};
        block_comp_sums<1>(t, e, red, v.part + pidx * 4);
        if (threadIdx.x >= 1 && threadIdx.x < 4) v.part[pidx * 4 + threadIdx.x] = 0.0;
    }
    if (HIST) {
        // block max of |w| -> fixed-point scale; integer LDS sums are exact
        double m = 0.0;
        for (int j = 0; j < a.spt; ++j) m = fmax(m, st_w[threadIdx.x + j * 256]);
        m = wave_max(m);
        if (lane == 0) wmax[wave] = m;
        __syncthreads();
        m = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
        const int e = m > 0.0 ? ilogb(m) : 0;
        const int sc = 50 - e;  // m 2^sc < 2^51, <= 2^12 terms stay below 2^63
        for (int j = 0; j < a.spt; ++j) {
            const uint32_t pk = st_b[threadIdx.x + j * 256];
            if (pk == 0xffffffffu) continue;
            const unsigned long long v =
                (unsigned long long)rint(ldexp(st_w[threadIdx.x + j * 256], sc));
            for (int d = 0; d < 4; ++d) {
                const int b = (pk >> (8 * d)) & 0xff;
                atomicAdd(&hh[d * a.bins + b], v);
                atomicAdd(&hc[d * a.bins + b], 1u);
            }
        }
        __syncthreads();
        for (int k = threadIdx.x; k < 4 * a.bins; k += blockDim.x) {
            a.part_hist[pidx * 4 * a.bins + k] = ldexp((double)hh[k], -sc);
            a.part_cnt[pidx * 4 * a.bins + k] = (double)hc[k];
        }
    }
}

// The evolver's FP32 instance (sum of w only) fits 72 VGPRs without spills:
// 7 waves per SIMD instead of the 6 that its natural 76 would allow.
template <int MODE, bool FULL, bool ANYC>
constexpr int kNloWaves = (MODE == NLO_PREC_FP32 && !FULL && !ANYC) ? 7 : 1;

template <int MODE, bool HIST, bool FULL, bool ANYC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kNloWaves<MODE, FULL, ANYC>))) nlo_kernel(NloArgs a) {
    const RunView v{a.R, a.r, a.N, a.t64, a.t32, a.edges, a.part};
    nlo_body<MODE, HIST, FULL, ANYC>(a, v, blockIdx.y + 1, blockIdx.x);
}

// Compensated fixed-order sums of the bpp partials (4 sums each) at
// part[base ..]: each thread adds its strided partials in order, then a fixed
// butterfly over lanes and a fixed sum over the (at most four) waves, all as
// (value, error) pairs rounded once into out[0..3].  Block of <= 256 threads.
__device__ __forceinline__ void reduce_partials4(const double* __restrict__ part, size_t base, int bpp,
                                                 double (*red)[4][2], double* __restrict__ out) {
    double t[4] = {0.0, 0.0, 0.0, 0.0}, e[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < bpp; b += blockDim.x) {
        const double4 v = *reinterpret_cast<const double4*>(part + (base + b) * 4);
        comp_add(t[0], e[0], v.x);
        comp_add(t[1], e[1], v.y);
        comp_add(t[2], e[2], v.z);
        comp_add(t[3], e[3], v.w);
    }
    block_comp_sums<4>(t, e, red, out);
}

// Fixed-order sum of the per-block partials of each parent (reduce_partials4).
__global__ void __launch_bounds__(256) nlo_reduce_kernel(int R, int bpp, int bins, int hist,
                                                         const double* __restrict__ part,
                                                         const double* __restrict__ ph,
                                                         const double* __restrict__ pc,
                                                         double* __restrict__ sums,
                                                         double* __restrict__ hsum,
                                                         double* __restrict__ csum) {
    __shared__ double red[4][4][2];
    const int p = blockIdx.x;  // 0 .. R-1
    if (p == 0) {
        if (threadIdx.x < 4) sums[threadIdx.x] = 0.0;
        if (hist)
            for (int k = threadIdx.x; k < 4 * bins; k += blockDim.x) {
                hsum[k] = 0.0;
                csum[k] = 0.0;
            }
        return;
    }
    const size_t base = (size_t)(p - 1) * bpp;
    reduce_partials4(part, base, bpp, red, sums + (size_t)p * 4);
    if (hist) {
        const size_t stride = (size_t)4 * bins;
        for (int k = threadIdx.x; k < 4 * bins; k += blockDim.x) {
            const double* qh = ph + base * stride + k;
            const double* qc = pc + base * stride + k;
            double th = 0.0, tc = 0.0;
            int b = 0;
            for (; b + 8 <= bpp; b += 8) {  // eight loads in flight, added in block order
                double vh[8], vc[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    vh[i] = qh[(size_t)(b + i) * stride];
                    vc[i] = qc[(size_t)(b + i) * stride];
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    th += vh[i];
                    tc += vc[i];
                }
            }
            for (; b < bpp; ++b) {
                th += qh[(size_t)b * stride];
                tc += qc[(size_t)b * stride];
            }
            hsum[(size_t)p * stride + k] = th;
            csum[(size_t)p * stride + k] = tc;
        }
    }
}

// counter (optional) accumulates the number of active parents: integer
// adds, so the total is exact whatever the order.
__global__ void rhs_finalize_kernel(int R, const double* __restrict__ N, const double* __restrict__ k1,
                                    const double* __restrict__ sums, double inv_n,
                                    double* __restrict__ dndy, double* __restrict__ k2kf,
                                    unsigned long long* __restrict__ counter) {
    __shared__ unsigned int block_active;
    if (threadIdx.x == 0) block_active = 0u;
    __syncthreads();
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = false;
    if (p < R) {
        active = p >= 1 && N[p] <= 0.99999;
        const double nlo = active ? sums[(size_t)p * 4] * inv_n : 0.0;
        double v = active ? k1[p] + nlo : 0.0;
        dndy[p] = isfinite(v) ? v : 0.0;
        if (k2kf) k2kf[p] = nlo;
    }
    if (counter) {
        if (active) atomicAdd(&block_active, 1u);
        __syncthreads();
        if (threadIdx.x == 0 && block_active) atomicAdd(counter, (unsigned long long)block_active);
    }
}

// out = N + h (a0 K0 + a1 K1 + a2 K2)
__global__ void rk_stage_kernel(int R, const double* __restrict__ N, const double* __restrict__ K0,
                                const double* __restrict__ K1, const double* __restrict__ K2,
                                double a0, double a1, double a2, double h, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    double acc = a0 * K0[i];
    if (K1) acc += a1 * K1[i];
    if (K2) acc += a2 * K2[i];
    out[i] = N[i] + h * acc;
}

// N <- N + h sum_s b_s K_s, clamp to [0, 1] when forced, non-finite flag
// (S/bk/evolution.py:150-153).
__global__ void rk_final_kernel(int R, double* __restrict__ N, const double* __restrict__ K0,
                                const double* __restrict__ K1, const double* __restrict__ K2,
                                const double* __restrict__ K3, double b0, double b1, double b2,
                                double b3, double h, int force_positive, int32_t* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    double acc = b0 * K0[i];
    if (K1) acc += b1 * K1[i];
    if (K2) acc += b2 * K2[i];
    if (K3) acc += b3 * K3[i];
    double v = N[i] + h * acc;
    if (force_positive) v = fmin(fmax(v, 0.0), 1.0);
    if (!isfinite(v)) flag[0] = 1;
    N[i] = v;
}

// ------------------------------------------------------------ batches

__global__ void __launch_bounds__(kSplineThreads) spline_build_batch_kernel(const BatchRun* __restrict__ runs,
                                                                const double* __restrict__ r,
                                                                const double* __restrict__ N,
                                                                int force_positive, double* __restrict__ t64,
                                                                float* __restrict__ t32) {
    const BatchRun run = runs[blockIdx.x];
    spline_build_body(run.R, r + run.node_off, N + run.node_off, force_positive, t64 + run.spline_off,
                      t32 ? t32 + run.spline_off : nullptr);
}

// One block per node of every run (node 0 of a run writes 0).
template <bool ANY>
__global__ void __launch_bounds__(kK1Threads) k1_batch_kernel(const BatchRun* __restrict__ runs,
                                                              const int32_t* __restrict__ node_run,
                                                              const double* __restrict__ r,
                                                              const double* __restrict__ N,
                                                              const double* __restrict__ t64, BkConsts c,
                                                              GaussRule g, double* __restrict__ k1) {
    const BatchRun run = runs[node_run[blockIdx.x]];
    const K1History H{};
    k1_body<ANY>(run.R, r + run.node_off, N + run.node_off, t64 + run.spline_off, c, g, H,
                 blockIdx.x - run.node_off, k1 + run.node_off);
}

// The adaptive rule, one block per node of every run.
template <bool ANY>
__global__ void __launch_bounds__(kK1aWaves * 64)
    k1_adaptive_batch_kernel(const BatchRun* __restrict__ runs, const int32_t* __restrict__ node_run,
                             const double* __restrict__ r, const double* __restrict__ N,
                             const double* __restrict__ t64, BkConsts c, int rint, int tint, int cap,
                             double epsrel, double* __restrict__ k1) {
    const BatchRun run = runs[node_run[blockIdx.x]];
    const K1History H{};
    k1_adaptive_body<ANY>(run.R, r + run.node_off, N + run.node_off, t64 + run.spline_off, c, H, rint,
                          tint, cap, epsrel, blockIdx.x - run.node_off, k1 + run.node_off, nullptr,
                          nullptr);
}

struct BatchNlo {
    const BatchRun* runs;
    const int32_t* node_run;
    const int32_t* parent_node;
    const double* r;
    const double* N;
    const double* t64;
    const float* t32;
    const double* edges;
    double* part;  // parents x bpp x 4
    unsigned long long* counter;
    int64_t items;  // parents x bpp
};

// Persistent K2 + Kf kernel: every block takes items (parent, chunk) from a
// device counter and runs the same per-block body as nlo_kernel on the
// run's slices; each item writes its own partial sums, so the result does not
// depend on which block took it.
template <int MODE, bool ANYC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kNloWaves<MODE, false, ANYC>))) nlo_batch_kernel(NloArgs a, BatchNlo w) {
    __shared__ int64_t item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = (int64_t)atomicAdd(w.counter, 1ull);
        __syncthreads();
        const int64_t item = item_s;
        if (item >= w.items) break;  // block-uniform
        const int gp = (int)(item / a.bpp);
        const int blk = (int)(item - (int64_t)gp * a.bpp);
        const int node = w.parent_node[gp];
        const BatchRun run = w.runs[w.node_run[node]];
        const int p = node - run.node_off;
        const RunView v{run.R,
                        w.r + run.node_off,
                        w.N + run.node_off,
                        w.t64 + run.spline_off,
                        w.t32 + run.spline_off,
                        w.edges + (size_t)run.node_off * 4 * (a.bins + 1),
                        w.part + (size_t)run.parent_off * a.bpp * 4};
        nlo_body<MODE, false, false, ANYC>(a, v, p, blk);
        __syncthreads();  // LDS and item_s are reused by the next item
    }
}

// Fixed-order sums of each parent's partials (the order of nlo_reduce_kernel);
// one block per node, node 0 of a run writes 0.
__global__ void __launch_bounds__(256) nlo_reduce_batch_kernel(const BatchRun* __restrict__ runs,
                                                               const int32_t* __restrict__ node_run, int bpp,
                                                               const double* __restrict__ part,
                                                               double* __restrict__ sums) {
    __shared__ double red[4][4][2];
    const BatchRun run = runs[node_run[blockIdx.x]];
    const int p = blockIdx.x - run.node_off;
    if (p == 0) {
        if (threadIdx.x < 4) sums[(size_t)blockIdx.x * 4 + threadIdx.x] = 0.0;
        return;
    }
    const size_t base = (size_t)(run.parent_off + p - 1) * bpp;
    reduce_partials4(part, base, bpp, red, sums + (size_t)blockIdx.x * 4);
}

__global__ void rhs_finalize_batch_kernel(int nodes, const BatchRun* __restrict__ runs,
                                          const int32_t* __restrict__ node_run, const double* __restrict__ N,
                                          const double* __restrict__ k1, const double* __restrict__ sums,
                                          double inv_n, double* __restrict__ dndy,
                                          unsigned long long* __restrict__ counter) {
    __shared__ unsigned int block_active;
    if (threadIdx.x == 0) block_active = 0u;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = false;
    if (i < nodes) {
        const int p = i - runs[node_run[i]].node_off;
        active = p >= 1 && N[i] <= 0.99999;
        const double nlo = active ? sums[(size_t)i * 4] * inv_n : 0.0;
        const double v = active ? k1[i] + nlo : 0.0;
        dndy[i] = isfinite(v) ? v : 0.0;
    }
    if (counter) {
        if (active) atomicAdd(&block_active, 1u);
        __syncthreads();
        if (threadIdx.x == 0 && block_active) atomicAdd(counter, (unsigned long long)block_active);
    }
}

__global__ void fill_i32_kernel(int32_t* p, int32_t v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_u64_kernel(unsigned long long* p, unsigned long long v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

// ------------------------------------------------------ internal launches

int init_kernel_attributes() {
    static std::once_flag once;
    static int status = NLO_OK;
    std::call_once(once, [] {
        int dev = 0, max_lds = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess) {
            (void)hipGetLastError();
            return;
        }
        const int dyn = max_lds - 2048;  // room for the kernels' static LDS
        if (dyn <= 65536) return;
        const void* fns[] = {
            (const void*)spline_build_kernel,
            (const void*)k1_kernel<false>,
            (const void*)k1_kernel<true>,
            (const void*)k1_adaptive_kernel<false>,
            (const void*)k1_adaptive_kernel<true>,
            (const void*)spline_build_batch_kernel,
            (const void*)k1_batch_kernel<false>,
            (const void*)k1_batch_kernel<true>,
            (const void*)k1_adaptive_batch_kernel<false>,
            (const void*)k1_adaptive_batch_kernel<true>,
            (const void*)nlo_batch_kernel<NLO_PREC_FP32, false>,
            (const void*)nlo_batch_kernel<NLO_PREC_FP32, true>,
            (const void*)nlo_batch_kernel<NLO_PREC_FP64, false>,
            (const void*)nlo_batch_kernel<NLO_PREC_FP64, true>,
            (const void*)nlo_batch_kernel<NLO_PREC_FP64_STABLE, false>,
            (const void*)nlo_batch_kernel<NLO_PREC_FP64_STABLE, true>,
#define NLO_INSTANCES(ANYC)                                                   \
    (const void*)nlo_kernel<NLO_PREC_FP64, false, true, ANYC>,                \
        (const void*)nlo_kernel<NLO_PREC_FP64, false, false, ANYC>,           \
        (const void*)nlo_kernel<NLO_PREC_FP64, true, true, ANYC>,             \
        (const void*)nlo_kernel<NLO_PREC_FP32, false, true, ANYC>,            \
        (const void*)nlo_kernel<NLO_PREC_FP32, false, false, ANYC>,           \
        (const void*)nlo_kernel<NLO_PREC_FP32, true, true, ANYC>,             \
        (const void*)nlo_kernel<NLO_PREC_FP64_STABLE, false, true, ANYC>,     \
        (const void*)nlo_kernel<NLO_PREC_FP64_STABLE, false, false, ANYC>,    \
        (const void*)nlo_kernel<NLO_PREC_FP64_STABLE, true, true, ANYC>
            NLO_INSTANCES(false),
            NLO_INSTANCES(true),
#undef NLO_INSTANCES
        };
        for (const void* f : fns) {
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, dyn) != hipSuccess)
                (void)hipGetLastError();
        }
    });
    return status;
}

int launch_spline_build(int R, const double* r, const double* N, int force_positive, double* t64,
                        float* t32, hipStream_t stream) {
    const size_t smem = (size_t)8 * R * sizeof(double);
    int rc = init_kernel_attributes();
    if (rc != NLO_OK) return rc;
    hipLaunchKernelGGL(spline_build_kernel, dim3(1), dim3(kSplineThreads), smem, stream, R, r, N,
                       force_positive, t64, t32);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int launch_k1(int R, const double* r, const double* N, const double* t64, const BkConsts& c, int nr,
              int nt, double* k1, hipStream_t stream, const K1History* hist) {
    GaussRule g{};
    g.nr = nr;
    g.nt = nt;
    if (nlo_gauss_legendre(nr, g.xr, g.wr) != NLO_OK) return NLO_ERR_INVALID;
    if (nlo_gauss_legendre(nt, g.xt, g.wt) != NLO_OK) return NLO_ERR_INVALID;
    const size_t smem = ((size_t)R + 4 * (size_t)(R - 1)) * sizeof(double);
    int rc = init_kernel_attributes();
    if (rc != NLO_OK) return rc;
    K1History H{};
    if (hist) H = *hist;
    if (c.production && !hist)
        hipLaunchKernelGGL(k1_kernel<false>, dim3(R), dim3(kK1Threads), smem, stream, R, r, N, t64, c,
                           g, H, k1);
    else
        hipLaunchKernelGGL(k1_kernel<true>, dim3(R), dim3(kK1Threads), smem, stream, R, r, N, t64, c, g,
                           H, k1);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int launch_k1_adaptive(int R, const double* r, const double* N, const double* t64, const BkConsts& c,
                       int rint, int tint, double epsrel, double* k1, double* k1_err,
                       long long* k1_evals, hipStream_t stream, const K1History* hist) {
    const int cap = rint > tint ? rint : tint;
    const size_t smem = k1a_lds_doubles(R, cap) * sizeof(double);
    int rc = init_kernel_attributes();
    if (rc != NLO_OK) return rc;
    int dev = 0, max_lds = 0;
    NLO_HIP_CHECK(hipGetDevice(&dev));
    NLO_HIP_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    NLO_REQUIRE(smem + 64 <= (size_t)max_lds,
                "adaptive K1: R=%d with %d intervals needs %zu bytes of LDS (device: %d)", R, cap, smem,
                max_lds);
    K1History H{};
    if (hist) H = *hist;
    if (c.production && !hist)
        hipLaunchKernelGGL(k1_adaptive_kernel<false>, dim3(R), dim3(kK1aWaves * 64), smem, stream, R, r,
                           N, t64, c, H, rint, tint, cap, epsrel, k1, k1_err, k1_evals);
    else
        hipLaunchKernelGGL(k1_adaptive_kernel<true>, dim3(R), dim3(kK1aWaves * 64), smem, stream, R, r,
                           N, t64, c, H, rint, tint, cap, epsrel, k1, k1_err, k1_evals);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

NloPlan nlo_plan(int R, int bins, int64_t n, bool hist) {
    NloPlan pl{};
    // Enough blocks to fill 256 CUs several times over; each lane evaluates
    // spt samples (at most 8 when the histogram stages them in LDS).
    const int64_t parents = R - 1;
    int spt = 16;
    while (spt > 1 && parents * ((n + 256 * spt - 1) / (256 * spt)) < 8192) spt /= 2;
    if (hist && spt > 8) spt = 8;
    pl.spt = spt;
    pl.bpp = (int)((n + 256 * spt - 1) / (256 * spt));
    pl.part_bytes = (size_t)parents * pl.bpp * 4 * sizeof(double);
    pl.hist_bytes = hist ? (size_t)parents * pl.bpp * 4 * bins * sizeof(double) : 0;
    pl.workspace = align_up(pl.part_bytes, 256) + 2 * align_up(pl.hist_bytes, 256);
    return pl;
}

int launch_nlo(int R, const double* r, const double* N, const double* t64, const float* t32,
               const double* edges, int bins, const nlo_sobol_table& sobol, int64_t first, int64_t n,
               int precision, const BkConsts& c, double* sums, double* hist, double* count,
               void* ws, size_t ws_bytes, hipStream_t stream, bool components) {
    const bool want_hist = hist != nullptr;
    const NloPlan pl = nlo_plan(R, bins, n, want_hist);
    if (ws_bytes < pl.workspace || (!ws && pl.workspace)) {
        set_error("nlo_bk_nlo: workspace %zu bytes < %zu", ws_bytes, pl.workspace);
        return NLO_ERR_WORKSPACE;
    }
    NloArgs a{};
    a.R = R;
    a.bins = bins;
    a.spt = pl.spt;
    a.bpp = pl.bpp;
    a.n = n;
    a.first = first;
    a.r = r;
    a.N = N;
    a.t64 = t64;
    a.t32 = t32;
    a.edges = edges;
    char* w = static_cast<char*>(ws);
    a.part = reinterpret_cast<double*>(w);
    a.part_hist = want_hist ? reinterpret_cast<double*>(w + align_up(pl.part_bytes, 256)) : nullptr;
    a.part_cnt = want_hist ? reinterpret_cast<double*>(w + align_up(pl.part_bytes, 256) +
                                                       align_up(pl.hist_bytes, 256))
                           : nullptr;
    a.c = c;
    for (int d = 0; d < 4; ++d) {
        for (int b = 0; b < 30; ++b) a.dir[d][b] = sobol.direction[d][b];
        a.shift[d] = sobol.shift[d];
        a.first_q[d] = sobol.first[d];
    }
    const size_t tsize = precision == NLO_PREC_FP32 ? sizeof(float) : sizeof(double);
    const size_t smem = nlo_layout(R, bins, pl.spt, want_hist, tsize).total;
    const dim3 grid(pl.bpp, R - 1);
    int rc = init_kernel_attributes();
    if (rc != NLO_OK) return rc;
#define NLO_PICK2(MODE, ANYC)                                                                        \
    if (want_hist)                                                                                   \
        hipLaunchKernelGGL((nlo_kernel<MODE, true, true, ANYC>), grid, dim3(256), smem, stream, a);  \
    else if (components)                                                                             \
        hipLaunchKernelGGL((nlo_kernel<MODE, false, true, ANYC>), grid, dim3(256), smem, stream, a); \
    else                                                                                             \
        hipLaunchKernelGGL((nlo_kernel<MODE, false, false, ANYC>), grid, dim3(256), smem, stream, a)
#define NLO_PICK(MODE)             \
    if (production_coupling) {     \
        NLO_PICK2(MODE, false);    \
    } else {                       \
        NLO_PICK2(MODE, true);     \
    }
    // the production coupling: smooth (NF <= 3), not fixed, at the smallest distance
    const bool production_coupling = !c.fixed && !c.thresholds && c.rc_nlo == NLO_RC_NLO_SMALLEST;
    if (precision == NLO_PREC_FP64) {
        NLO_PICK(NLO_PREC_FP64);
    } else if (precision == NLO_PREC_FP32) {
        NLO_PICK(NLO_PREC_FP32);
    } else {
        NLO_PICK(NLO_PREC_FP64_STABLE);
    }
#undef NLO_PICK
#undef NLO_PICK2
    NLO_LAUNCH_CHECK();
    hipLaunchKernelGGL(nlo_reduce_kernel, dim3(R), dim3(256), 0, stream, R, pl.bpp, bins,
                       want_hist ? 1 : 0, a.part, a.part_hist, a.part_cnt, sums, hist, count);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int launch_rhs_finalize(int R, const double* N, const double* k1, const double* sums, int64_t n,
                        double* dndy, double* k2kf, unsigned long long* counter, hipStream_t stream) {
    hipLaunchKernelGGL(rhs_finalize_kernel, dim3((R + 255) / 256), dim3(256), 0, stream, R, N, k1,
                       sums, 1.0 / (double)n, dndy, k2kf, counter);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int launch_rk_stage(int R, const double* N, const double* K0, const double* K1, const double* K2,
                    double a0, double a1, double a2, double h, double* out, hipStream_t stream) {
    hipLaunchKernelGGL(rk_stage_kernel, dim3((R + 255) / 256), dim3(256), 0, stream, R, N, K0, K1, K2,
                       a0, a1, a2, h, out);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int launch_rk_final(int R, double* N, const double* K0, const double* K1, const double* K2,
                    const double* K3, double b0, double b1, double b2, double b3, double h,
                    int force_positive, int32_t* flag, hipStream_t stream) {
    hipLaunchKernelGGL(rk_final_kernel, dim3((R + 255) / 256), dim3(256), 0, stream, R, N, K0, K1, K2,
                       K3, b0, b1, b2, b3, h, force_positive, flag);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int launch_spline_build_batch(const BatchView& b, const double* r, const double* N, int force_positive,
                              double* t64, float* t32, hipStream_t stream) {
    const size_t smem = (size_t)8 * b.Rmax * sizeof(double);
    int rc = init_kernel_attributes();
    if (rc != NLO_OK) return rc;
    hipLaunchKernelGGL(spline_build_batch_kernel, dim3(b.runs), dim3(kSplineThreads), smem, stream, b.runs_d,
                       r, N, force_positive, t64, t32);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int launch_k1_batch(const BatchView& b, const double* r, const double* N, const double* t64,
                    const BkConsts& c, int nr, int nt, double* k1, hipStream_t stream) {
    GaussRule g{};
    g.nr = nr;
    g.nt = nt;
    if (nlo_gauss_legendre(nr, g.xr, g.wr) != NLO_OK) return NLO_ERR_INVALID;
    if (nlo_gauss_legendre(nt, g.xt, g.wt) != NLO_OK) return NLO_ERR_INVALID;
    const size_t smem = ((size_t)b.Rmax + 4 * (size_t)(b.Rmax - 1)) * sizeof(double);
    int rc = init_kernel_attributes();
    if (rc != NLO_OK) return rc;
    if (c.production)
        hipLaunchKernelGGL(k1_batch_kernel<false>, dim3(b.nodes), dim3(kK1Threads), smem, stream, b.runs_d,
                           b.node_run, r, N, t64, c, g, k1);
    else
        hipLaunchKernelGGL(k1_batch_kernel<true>, dim3(b.nodes), dim3(kK1Threads), smem, stream, b.runs_d,
                           b.node_run, r, N, t64, c, g, k1);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int launch_k1_adaptive_batch(const BatchView& b, const double* r, const double* N, const double* t64,
                             const BkConsts& c, int rint, int tint, double epsrel, double* k1,
                             hipStream_t stream) {
    const int cap = rint > tint ? rint : tint;
    const size_t smem = k1a_lds_doubles(b.Rmax, cap) * sizeof(double);
    int rc = init_kernel_attributes();
    if (rc != NLO_OK) return rc;
    int dev = 0, max_lds = 0;
    NLO_HIP_CHECK(hipGetDevice(&dev));
    NLO_HIP_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    NLO_REQUIRE(smem + 64 <= (size_t)max_lds,
                "adaptive K1: R=%d with %d intervals needs %zu bytes of LDS (device: %d)", b.Rmax, cap,
                smem, max_lds);
    if (c.production)
        hipLaunchKernelGGL(k1_adaptive_batch_kernel<false>, dim3(b.nodes), dim3(kK1aWaves * 64), smem,
                           stream, b.runs_d, b.node_run, r, N, t64, c, rint, tint, cap, epsrel, k1);
    else
        hipLaunchKernelGGL(k1_adaptive_batch_kernel<true>, dim3(b.nodes), dim3(kK1aWaves * 64), smem,
                           stream, b.runs_d, b.node_run, r, N, t64, c, rint, tint, cap, epsrel, k1);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

size_t nlo_batch_partials(const BatchView& b, int64_t n, int spt) {
    const int64_t bpp = (n + 256 * spt - 1) / (256 * spt);
    return (size_t)b.parents * (size_t)bpp * 4 * sizeof(double);
}

namespace {

template <int MODE, bool ANYC>
const void* nlo_batch_fn() {
    return reinterpret_cast<const void*>(nlo_batch_kernel<MODE, ANYC>);
}

const void* nlo_batch_pick(int precision, bool production_coupling) {
    if (precision == NLO_PREC_FP32)
        return production_coupling ? nlo_batch_fn<NLO_PREC_FP32, false>() : nlo_batch_fn<NLO_PREC_FP32, true>();
    if (precision == NLO_PREC_FP64)
        return production_coupling ? nlo_batch_fn<NLO_PREC_FP64, false>() : nlo_batch_fn<NLO_PREC_FP64, true>();
    return production_coupling ? nlo_batch_fn<NLO_PREC_FP64_STABLE, false>()
                               : nlo_batch_fn<NLO_PREC_FP64_STABLE, true>();
}

size_t nlo_batch_lds(int precision, int Rmax, int bins, int spt) {
    return nlo_layout(Rmax, bins, spt, false, precision == NLO_PREC_FP32 ? sizeof(float) : sizeof(double))
        .total;
}

bool production_nlo_coupling(const BkConsts& c) {
    return !c.fixed && !c.thresholds && c.rc_nlo == NLO_RC_NLO_SMALLEST;
}

}  // namespace

int nlo_batch_resident_blocks(int precision, const BkConsts& c, int Rmax, int bins, int spt) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    const void* fn = nlo_batch_pick(precision, production_nlo_coupling(c));
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, nlo_batch_lds(precision, Rmax, bins, spt)) !=
        hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return cus * (per_cu > 0 ? per_cu : 1);
}

int launch_nlo_batch(const BatchView& b, const double* r, const double* N, const double* t64,
                     const float* t32, const double* edges, int bins, const nlo_sobol_table& sobol,
                     int64_t first, int64_t n, int precision, const BkConsts& c, int spt, int grid,
                     unsigned long long* counter, double* part, double* sums, hipStream_t stream) {
    NloArgs a{};
    a.R = b.Rmax;
    a.bins = bins;
    a.spt = spt;
    a.bpp = (int)((n + 256 * spt - 1) / (256 * spt));
    a.n = n;
    a.first = first;
    a.c = c;
    for (int d = 0; d < 4; ++d) {
        for (int k = 0; k < 30; ++k) a.dir[d][k] = sobol.direction[d][k];
        a.shift[d] = sobol.shift[d];
        a.first_q[d] = sobol.first[d];
    }
    BatchNlo w{b.runs_d, b.node_run, b.parent_node, r, N, t64, t32, edges, part, counter,
               (int64_t)b.parents * a.bpp};
    const size_t smem = nlo_batch_lds(precision, b.Rmax, bins, spt);
    int rc = init_kernel_attributes();
    if (rc != NLO_OK) return rc;
    rc = launch_fill_u64(counter, 0ull, 1, stream);  // the item counter
    if (rc != NLO_OK) return rc;
    const bool prod = production_nlo_coupling(c);
#define NLO_BATCH(MODE)                                                                                    \
    if (prod)                                                                                              \
        hipLaunchKernelGGL((nlo_batch_kernel<MODE, false>), dim3(grid), dim3(256), smem, stream, a, w);   \
    else                                                                                                   \
        hipLaunchKernelGGL((nlo_batch_kernel<MODE, true>), dim3(grid), dim3(256), smem, stream, a, w)
    if (precision == NLO_PREC_FP32) {
        NLO_BATCH(NLO_PREC_FP32);
    } else if (precision == NLO_PREC_FP64) {
        NLO_BATCH(NLO_PREC_FP64);
    } else {
        NLO_BATCH(NLO_PREC_FP64_STABLE);
    }
#undef NLO_BATCH
    NLO_LAUNCH_CHECK();
    hipLaunchKernelGGL(nlo_reduce_batch_kernel, dim3(b.nodes), dim3(256), 0, stream, b.runs_d, b.node_run,
                       a.bpp, part, sums);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int launch_rhs_finalize_batch(const BatchView& b, const double* N, const double* k1, const double* sums,
                              int64_t n, double* dndy, unsigned long long* counter, hipStream_t stream) {
    hipLaunchKernelGGL(rhs_finalize_batch_kernel, dim3((b.nodes + 255) / 256), dim3(256), 0, stream, b.nodes,
                       b.runs_d, b.node_run, N, k1, sums, 1.0 / (double)n, dndy, counter);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

// ------------------------------------------- no-allocation K2 + Kf entry

// Scratch of nlo_bk_nlo_noalloc: the per-block partial sums of the K2 + Kf
// launch, (R - 1) x blocks-per-parent x 4 doubles, every element written by
// its block before the reduce reads it (nothing has to start at zero).
// File-scope device memory: each GPU the library runs on holds its own copy
// until the process ends.  Host code never hands `scratch` itself to a
// kernel or a copy: it uses the device address scratch_address() looked up
// for the current device.
static __device__ __align__(256) unsigned char scratch[NLO_BK_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex scratch_mutex;
void* scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

// The current device's address of `scratch`: looked up once per device, with
// that device current, on its first use, and kept.
int scratch_address(void** out) {
    int dev = -1;
    NLO_HIP_CHECK(hipGetDevice(&dev));
    NLO_REQUIRE(dev >= 0 && dev < kMaxScratchDevices,
                "nlo_bk_nlo_noalloc: device %d beyond the %d supported", dev, kMaxScratchDevices);
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        NLO_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        NLO_REQUIRE(p != nullptr, "nlo_bk_nlo_noalloc: no device address for the scratch array");
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return NLO_OK;
}

int launch_fill_i32(int32_t* p, int32_t v, int n, hipStream_t stream) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, v, n);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int launch_fill_u64(unsigned long long* p, unsigned long long v, int n, hipStream_t stream) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, v, n);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

}  // namespace nlo

// ------------------------------------------------------------------- C ABI

extern "C" {

int nlo_bk_spline_build(int32_t R, const double* d_r, const double* d_N, int32_t force_positive,
                        double* d_table64, float* d_table32, hipStream_t stream) {
    NLO_REQUIRE(R >= 3 && R <= nlo::kMaxR, "nlo_bk_spline_build: R=%d outside [3, %d]", R, nlo::kMaxR);
    NLO_REQUIRE(d_r && d_N && d_table64, "nlo_bk_spline_build: null pointer");
    const int rc = nlo::launch_spline_build(R, d_r, d_N, force_positive, d_table64, d_table32, stream);
    if (rc == NLO_OK) nlo::clear_error();
    return rc;
}

int nlo_bk_k1(int32_t R, const double* d_r, const double* d_N, const double* d_table64,
              const nlo_bk_physics* physics, int32_t nr, int32_t nt, double* d_k1,
              hipStream_t stream) {
    NLO_REQUIRE(R >= 3 && R <= nlo::kMaxR, "nlo_bk_k1: R=%d outside [3, %d]", R, nlo::kMaxR);
    NLO_REQUIRE(nr >= 1 && nr <= nlo::kMaxGaussR && nt >= 1 && nt <= nlo::kMaxGaussT,
                "nlo_bk_k1: GL orders (%d, %d) outside [1, %d] x [1, %d]", nr, nt, nlo::kMaxGaussR,
                nlo::kMaxGaussT);
    NLO_REQUIRE(d_r && d_N && d_table64 && d_k1, "nlo_bk_k1: null pointer");
    int rc = nlo::validate_physics(physics);
    if (rc != NLO_OK) return rc;
    rc = nlo::launch_k1(R, d_r, d_N, d_table64, nlo::make_consts(*physics), nr, nt, d_k1, stream);
    if (rc == NLO_OK) nlo::clear_error();
    return rc;
}

int nlo_bk_k1_history(int32_t R, const double* d_r, const double* d_N, const double* d_table64,
                      const nlo_bk_physics* physics, int32_t nr, int32_t nt,
                      const double* d_hist_y, const double* d_hist_N, const int32_t* d_rows,
                      int32_t capacity, double* d_k1, hipStream_t stream) {
    NLO_REQUIRE(R >= 3 && R <= nlo::kMaxR, "nlo_bk_k1_history: R=%d outside [3, %d]", R, nlo::kMaxR);
    NLO_REQUIRE(nr >= 1 && nr <= nlo::kMaxGaussR && nt >= 1 && nt <= nlo::kMaxGaussT,
                "nlo_bk_k1_history: GL orders (%d, %d) outside [1, %d] x [1, %d]", nr, nt,
                nlo::kMaxGaussR, nlo::kMaxGaussT);
    NLO_REQUIRE(d_r && d_N && d_table64 && d_k1 && d_hist_y && d_hist_N && d_rows,
                "nlo_bk_k1_history: null pointer");
    NLO_REQUIRE(capacity >= 1, "nlo_bk_k1_history: capacity must be positive");
    int rc = nlo::validate_physics(physics);
    if (rc != NLO_OK) return rc;
    nlo::K1History H{d_hist_y, d_hist_N, d_r, d_rows, R, capacity};
    rc = nlo::launch_k1(R, d_r, d_N, d_table64, nlo::make_consts(*physics), nr, nt, d_k1, stream, &H);
    if (rc == NLO_OK) nlo::clear_error();
    return rc;
}

int nlo_bk_k1_adaptive(int32_t R, const double* d_r, const double* d_N, const double* d_table64,
                       const nlo_bk_physics* physics, int32_t rint, int32_t tint, double epsrel,
                       const double* d_hist_y, const double* d_hist_N, const int32_t* d_rows,
                       int32_t capacity, double* d_k1, double* d_err, int64_t* d_evals,
                       hipStream_t stream) {
    NLO_REQUIRE(R >= 3 && R <= nlo::kMaxR, "nlo_bk_k1_adaptive: R=%d outside [3, %d]", R, nlo::kMaxR);
    NLO_REQUIRE(rint >= 1 && rint <= nlo::kMaxGkIntervals && tint >= 1 && tint <= nlo::kMaxGkIntervals,
                "nlo_bk_k1_adaptive: interval limits (%d, %d) outside [1, %d]", rint, tint,
                nlo::kMaxGkIntervals);
    NLO_REQUIRE(epsrel > 0.0, "nlo_bk_k1_adaptive: epsrel must be positive");
    NLO_REQUIRE(d_r && d_N && d_table64 && d_k1, "nlo_bk_k1_adaptive: null pointer");
    const bool kin = d_hist_y || d_hist_N || d_rows;
    NLO_REQUIRE(!kin || (d_hist_y && d_hist_N && d_rows && capacity >= 1),
                "nlo_bk_k1_adaptive: the history needs y, N, rows and a positive capacity");
    int rc = nlo::validate_physics(physics);
    if (rc != NLO_OK) return rc;
    nlo::K1History H{d_hist_y, d_hist_N, d_r, d_rows, R, capacity};
    rc = nlo::launch_k1_adaptive(R, d_r, d_N, d_table64, nlo::make_consts(*physics), rint, tint, epsrel,
                                 d_k1, d_err, reinterpret_cast<long long*>(d_evals), stream,
                                 kin ? &H : nullptr);
    if (rc == NLO_OK) nlo::clear_error();
    return rc;
}

size_t nlo_bk_nlo_workspace(int32_t R, int32_t bins, int64_t n, int32_t want_hist) {
    if (R < 3 || bins < 2 || n < 1) return 0;
    return nlo::nlo_plan(R, bins, n, want_hist != 0).workspace;
}

}  // extern "C"

namespace nlo {
namespace {

// Argument checks shared by nlo_bk_nlo and nlo_bk_nlo_noalloc (fn names the
// entry point in the messages); precision arrives without NLO_BK_SUM_ONLY.
int check_nlo_args(const char* fn, int32_t R, const double* d_r, const double* d_N,
                   const double* d_table64, const float* d_table32, const double* d_edges,
                   int32_t bins, const nlo_sobol_table* sobol, int64_t first_index, int64_t n,
                   int32_t precision, const nlo_bk_physics* physics, const double* d_sums) {
    NLO_REQUIRE(R >= 3 && R <= kMaxR, "%s: R=%d outside [3, %d]", fn, R, kMaxR);
    NLO_REQUIRE(bins >= 2 && bins <= 255, "%s: bins=%d outside [2, 255]", fn, bins);
    NLO_REQUIRE(sobol && sobol->dim == 4, "%s: a 4-dimensional Sobol table is required", fn);
    NLO_REQUIRE(n >= 1 && first_index >= 0 && first_index % 64 == 0 &&
                    first_index + n <= (int64_t(1) << 30),
                "%s: sample range [%lld, %lld) invalid (first must be a multiple of 64, "
                "end <= 2^30)",
                fn, (long long)first_index, (long long)(first_index + n));
    NLO_REQUIRE(precision == NLO_PREC_FP64 || precision == NLO_PREC_FP32 ||
                    precision == NLO_PREC_FP64_STABLE,
                "%s: unknown precision %d", fn, precision);
    NLO_REQUIRE(d_r && d_N && d_table64 && d_edges && d_sums, "%s: null pointer", fn);
    NLO_REQUIRE(precision != NLO_PREC_FP32 || d_table32, "%s: FP32 mode needs d_table32", fn);
    return validate_physics(physics);
}

}  // namespace
}  // namespace nlo

extern "C" {

int nlo_bk_nlo(int32_t R, const double* d_r, const double* d_N, const double* d_table64,
               const float* d_table32, const double* d_edges, int32_t bins,
               const nlo_sobol_table* sobol, int64_t first_index, int64_t n, int32_t precision,
               const nlo_bk_physics* physics, double* d_sums, double* d_hist, double* d_count,
               void* d_workspace, size_t workspace_bytes, hipStream_t stream) {
    const bool sum_only = (precision & NLO_BK_SUM_ONLY) != 0;
    precision &= ~NLO_BK_SUM_ONLY;
    int rc = nlo::check_nlo_args("nlo_bk_nlo", R, d_r, d_N, d_table64, d_table32, d_edges, bins, sobol,
                                 first_index, n, precision, physics, d_sums);
    if (rc != NLO_OK) return rc;
    NLO_REQUIRE(!(sum_only && d_hist), "nlo_bk_nlo: NLO_BK_SUM_ONLY with histograms");
    NLO_REQUIRE((d_hist == nullptr) == (d_count == nullptr),
                "nlo_bk_nlo: d_hist and d_count must be given together");
    rc = nlo::launch_nlo(R, d_r, d_N, d_table64, d_table32, d_edges, bins, *sobol, first_index, n,
                         precision, nlo::make_consts(*physics), d_sums, d_hist, d_count, d_workspace,
                         workspace_bytes, stream, /*components=*/!sum_only);
    if (rc == NLO_OK) nlo::clear_error();
    return rc;
}

int nlo_bk_nlo_noalloc(int32_t R, const double* d_r, const double* d_N, const double* d_table64,
                       const float* d_table32, const double* d_edges, int32_t bins,
                       const nlo_sobol_table* sobol, int64_t first_index, int64_t n, int32_t precision,
                       const nlo_bk_physics* physics, double* d_sums, hipStream_t stream) {
    const bool sum_only = (precision & NLO_BK_SUM_ONLY) != 0;
    precision &= ~NLO_BK_SUM_ONLY;
    int rc = nlo::check_nlo_args("nlo_bk_nlo_noalloc", R, d_r, d_N, d_table64, d_table32, d_edges, bins,
                                 sobol, first_index, n, precision, physics, d_sums);
    if (rc != NLO_OK) return rc;
    const size_t need = nlo::nlo_plan(R, bins, n, false).workspace;
    if (need > (size_t)NLO_BK_NOALLOC_SCRATCH_BYTES) {
        nlo::set_error("nlo_bk_nlo_noalloc: R=%d with n=%lld needs %zu bytes of partial sums, more than the "
                       "%zu-byte static scratch; use nlo_bk_nlo with a workspace",
                       R, (long long)n, need, (size_t)NLO_BK_NOALLOC_SCRATCH_BYTES);
        return NLO_ERR_WORKSPACE;
    }
    void* ws = nullptr;
    rc = nlo::scratch_address(&ws);
    if (rc != NLO_OK) return rc;
    rc = nlo::launch_nlo(R, d_r, d_N, d_table64, d_table32, d_edges, bins, *sobol, first_index, n,
                         precision, nlo::make_consts(*physics), d_sums, nullptr, nullptr, ws,
                         (size_t)NLO_BK_NOALLOC_SCRATCH_BYTES, stream, /*components=*/!sum_only);
    if (rc == NLO_OK) nlo::clear_error();
    return rc;
}

int nlo_bk_rhs_finalize(int32_t R, const double* d_N, const double* d_k1, const double* d_sums,
                        int64_t n, double* d_dndy, double* d_k2kf, hipStream_t stream) {
    NLO_REQUIRE(R >= 1 && n >= 1, "nlo_bk_rhs_finalize: R=%d n=%lld invalid", R, (long long)n);
    NLO_REQUIRE(d_N && d_k1 && d_sums && d_dndy, "nlo_bk_rhs_finalize: null pointer");
    const int rc = nlo::launch_rhs_finalize(R, d_N, d_k1, d_sums, n, d_dndy, d_k2kf, nullptr, stream);
    if (rc == NLO_OK) nlo::clear_error();
    return rc;
}

}  // extern "C"
