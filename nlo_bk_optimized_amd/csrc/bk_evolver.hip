// Rapidity driver: one RK step = spline build, K1, K2 + Kf, finalize and the
// stage update per stage, captured once into a hipGraph and replayed per step
// (S/bk/evolution.py:85-159, :190-326).  Fixed-step methods need no host
// synchronisation; the adaptive RK2 controller reads one error ratio per
// attempt, as the reference does.  No allocation beyond a 16-byte pinned
// host buffer: every device buffer lives in the caller's workspace.
#include <cfloat>
#include <cmath>
#include <cstring>

#include "bk_internal.hpp"

struct nlo_bk_evolver {
    nlo_bk_evolver_desc d;
    nlo::BkConsts c;
    const double* r;
    const double* edges;
    double* N;      // state
    double* Nst;    // stage input
    double* K[4];   // stage derivatives
    double* F;      // adaptive: one full step
    double* Hs;     // adaptive: state after the first half step
    double* T;      // adaptive: two half steps
    double* ctl;    // adaptive: [0] step, [1] error ratio
    double* k1;
    double* k2kf;
    double* sums;
    double* t64;
    float* t32;
    void* nlo_ws;
    size_t nlo_ws_bytes;
    int32_t* flag;
    int32_t* rows;               // kinematical constraint: history rows in use
    unsigned long long* active;  // active-parent evaluations since creation
    double* host_ctl;            // pinned: step out, ratio in
    hipStream_t capture;
    hipGraph_t graph;            // one fixed step, or one adaptive attempt
    hipGraphExec_t exec;
    hipGraph_t graph_k0;         // adaptive: K[0] = f(N)
    hipGraphExec_t exec_k0;
    // controller state (host)
    double y, h;
    int64_t accepted, rejected;
    bool k0_valid;
    int rows_host;
};

namespace nlo {
namespace {

struct EvolverLayout {
    size_t N, Nst, K, F, Hs, T, ctl, k1, k2kf, sums, t64, t32, flag, rows, active, nlo_ws,
        nlo_ws_bytes, total;
};

EvolverLayout evolver_layout(const nlo_bk_evolver_desc& d) {
    EvolverLayout L{};
    const size_t R = (size_t)d.R;
    size_t o = 0;
    auto take = [&o](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes, 256);
        return at;
    };
    L.N = take(R * sizeof(double));
    L.Nst = take(R * sizeof(double));
    L.K = take(4 * R * sizeof(double));
    L.F = take(R * sizeof(double));
    L.Hs = take(R * sizeof(double));
    L.T = take(R * sizeof(double));
    L.ctl = take(4 * sizeof(double));
    L.k1 = take(R * sizeof(double));
    L.k2kf = take(R * sizeof(double));
    L.sums = take(4 * R * sizeof(double));
    L.t64 = take((size_t)NLO_SPLINE_WORDS(d.R) * sizeof(double));
    L.t32 = take((size_t)NLO_SPLINE_WORDS(d.R) * sizeof(float));
    L.flag = take(sizeof(int32_t));
    L.rows = take(sizeof(int32_t));
    L.active = take(sizeof(unsigned long long));
    L.nlo_ws_bytes = d.nlo ? nlo_plan(d.R, d.bins, d.n_samples, false).workspace : 0;
    L.nlo_ws = take(L.nlo_ws_bytes > 0 ? L.nlo_ws_bytes : 1);
    L.total = o;
    return L;
}

// Butcher rows: stage s input = N + h sum_j a[s][j] K_j; result weights b.
struct Tableau {
    int stages;
    double a[4][3];
    double b[4];
};

Tableau tableau(int method) {
    Tableau t{};
    switch (method) {
        case NLO_RK_EULER:
            t.stages = 1;
            t.b[0] = 1.0;
            break;
        case NLO_RK_MIDPOINT:  // S/bk/evolution.py:265-276
            t.stages = 2;
            t.a[1][0] = 0.5;
            t.b[1] = 1.0;
            break;
        case NLO_RK_KUTTA3:  // S/bk/evolution.py:204-236
            t.stages = 3;
            t.a[1][0] = 0.5;
            t.a[2][0] = -1.0;
            t.a[2][1] = 2.0;
            t.b[0] = 1.0 / 6.0;
            t.b[1] = 4.0 / 6.0;
            t.b[2] = 1.0 / 6.0;
            break;
        default:  // classical RK4
            t.stages = 4;
            t.a[1][0] = 0.5;
            t.a[2][1] = 0.5;
            t.a[3][2] = 1.0;
            t.b[0] = 1.0 / 6.0;
            t.b[1] = 2.0 / 6.0;
            t.b[2] = 2.0 / 6.0;
            t.b[3] = 1.0 / 6.0;
            break;
    }
    return t;
}

#pragma clang fp contract(off)
// out = base + (coef * ctl[0]) K, rounded as the reference's
// `base + coef * step * K` (product, then sum; no fused multiply-add).
__global__ void axpy_step_kernel(int R, const double* __restrict__ base, const double* __restrict__ K,
                                 double coef, const double* __restrict__ ctl,
                                 double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    const double cs = coef * ctl[0];
    const double t = cs * K[i];
    out[i] = base[i] + t;
}

// ctl[1] = max_i (|T - F| / 3) / (abserr + relerr |T|), NaN if any term is
// NaN (torch.max propagates it).  One block.
__global__ void error_ratio_kernel(int R, const double* __restrict__ T, const double* __restrict__ F,
                                   double abserr, double relerr, double* __restrict__ ctl) {
    __shared__ double red[1024 / 64];
    __shared__ int any_nan;
    if (threadIdx.x == 0) any_nan = 0;
    __syncthreads();
    double m = 0.0;
    int nan_here = 0;
    for (int i = threadIdx.x; i < R; i += blockDim.x) {
        const double e = fabs(T[i] - F[i]) / 3.0;
        const double rel = relerr * fabs(T[i]);
        const double tol = abserr + rel;
        const double v = e / tol;
        if (v != v) nan_here = 1;
        else m = fmax(m, v);
    }
    if (nan_here) any_nan = 1;
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = fmax(t, red[w]);
        ctl[1] = any_nan ? NAN : t;
    }
}
#pragma clang fp contract(fast)

__global__ void copy_f64_kernel(int R, const double* __restrict__ src, double* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < R) dst[i] = src[i];
}

// Clamp to [0, 1] when forced and flag non-finite values (S/bk/evolution.py:150-153).
__global__ void clamp_flag_kernel(int R, double* __restrict__ N, int force_positive,
                                  int32_t* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    double v = N[i];
    if (force_positive) v = fmin(fmax(v, 0.0), 1.0);
    if (!isfinite(v)) flag[0] = 1;
    N[i] = v;
}

// Kinematical constraint: append N and y_prev + step as history row *rows
// (S/bk/evolution.py:154-156), then count it.  One block; rows beyond the
// capacity are not written (the host refuses such steps before launching).
__global__ void history_append_kernel(int R, const double* __restrict__ N, double step, int capacity,
                                      double* __restrict__ hy, double* __restrict__ hN,
                                      int32_t* __restrict__ rows) {
    const int k = *rows;
    if (k < 1 || k >= capacity) return;
    for (int i = threadIdx.x; i < R; i += blockDim.x) hN[(size_t)k * R + i] = N[i];
    if (threadIdx.x == 0) hy[k] = hy[k - 1] + step;
    __syncthreads();
    if (threadIdx.x == 0) *rows = k + 1;
}

// dN/dy at `in` into `out`: spline build, K1, K2 + Kf, finalize.
int enqueue_rhs(nlo_bk_evolver* ev, const double* in, double* out, hipStream_t s) {
    const nlo_bk_evolver_desc& d = ev->d;
    const int R = d.R;
    int rc = launch_spline_build(R, ev->r, in, d.force_positive, ev->t64, ev->t32, s);
    if (rc != NLO_OK) return rc;
    K1History H{d.d_history_y, d.d_history_N, ev->r, ev->rows, R, d.history_capacity};
    if (d.k1_rule == NLO_K1_ADAPTIVE)
        rc = launch_k1_adaptive(R, ev->r, in, ev->t64, ev->c, d.k1_rint, d.k1_tint, d.k1_epsrel, ev->k1,
                                nullptr, nullptr, s, d.kinematical ? &H : nullptr);
    else
        rc = launch_k1(R, ev->r, in, ev->t64, ev->c, d.nr, d.nt, ev->k1, s, d.kinematical ? &H : nullptr);
    if (rc != NLO_OK) return rc;
    if (d.nlo) {
        rc = launch_nlo(R, ev->r, in, ev->t64, ev->t32, ev->edges, d.bins, d.sobol, d.first_index,
                        d.n_samples, d.precision, ev->c, ev->sums, nullptr, nullptr, ev->nlo_ws,
                        ev->nlo_ws_bytes, s, /*components=*/false);
        if (rc != NLO_OK) return rc;
        return launch_rhs_finalize(R, in, ev->k1, ev->sums, d.n_samples, out, ev->k2kf, ev->active, s);
    }
    // LO-only: K2 + Kf sums stay zero
    return launch_rhs_finalize(R, in, ev->k1, ev->sums, 1, out, ev->k2kf, ev->active, s);
}

// Enqueues one fixed RK step on `s` (captured into the graph).
int enqueue_step(nlo_bk_evolver* ev, hipStream_t s) {
    const nlo_bk_evolver_desc& d = ev->d;
    const Tableau tb = tableau(d.rk_method);
    const int R = d.R;
    int rc;
    for (int st = 0; st < tb.stages; ++st) {
        const double* in = ev->N;
        if (st > 0) {
            rc = launch_rk_stage(R, ev->N, ev->K[0], st > 1 ? ev->K[1] : nullptr,
                                 st > 2 ? ev->K[2] : nullptr, tb.a[st][0], tb.a[st][1], tb.a[st][2],
                                 d.step, ev->Nst, s);
            if (rc != NLO_OK) return rc;
            in = ev->Nst;
        }
        rc = enqueue_rhs(ev, in, ev->K[st], s);
        if (rc != NLO_OK) return rc;
    }
    rc = launch_rk_final(R, ev->N, ev->K[0], tb.stages > 1 ? ev->K[1] : nullptr,
                         tb.stages > 2 ? ev->K[2] : nullptr, tb.stages > 3 ? ev->K[3] : nullptr, tb.b[0],
                         tb.b[1], tb.b[2], tb.b[3], d.step, d.force_positive, ev->flag, s);
    if (rc != NLO_OK || !d.kinematical) return rc;
    hipLaunchKernelGGL(history_append_kernel, dim3(1), dim3(1024), 0, s, R, ev->N, d.step,
                       d.history_capacity, d.d_history_y, d.d_history_N, ev->rows);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int launch_axpy(nlo_bk_evolver* ev, const double* base, const double* K, double coef, double* out,
                hipStream_t s) {
    const int R = ev->d.R;
    hipLaunchKernelGGL(axpy_step_kernel, dim3((R + 255) / 256), dim3(256), 0, s, R, base, K, coef,
                       ev->ctl, out);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

// One attempt of _adaptive_rk2_to (S/bk/evolution.py:253-314) with the step
// in ctl[0] and K[0] = f(N): the full midpoint step F, two half steps T and
// the error ratio into ctl[1].
int enqueue_attempt(nlo_bk_evolver* ev, hipStream_t s) {
    const int R = ev->d.R;
    int rc;
    if ((rc = launch_axpy(ev, ev->N, ev->K[0], 0.5, ev->Nst, s)) != NLO_OK) return rc;
    if ((rc = enqueue_rhs(ev, ev->Nst, ev->K[1], s)) != NLO_OK) return rc;
    if ((rc = launch_axpy(ev, ev->N, ev->K[1], 1.0, ev->F, s)) != NLO_OK) return rc;
    if ((rc = launch_axpy(ev, ev->N, ev->K[0], 0.25, ev->Nst, s)) != NLO_OK) return rc;
    if ((rc = enqueue_rhs(ev, ev->Nst, ev->K[2], s)) != NLO_OK) return rc;
    if ((rc = launch_axpy(ev, ev->N, ev->K[2], 0.5, ev->Hs, s)) != NLO_OK) return rc;
    if ((rc = enqueue_rhs(ev, ev->Hs, ev->K[3], s)) != NLO_OK) return rc;
    if ((rc = launch_axpy(ev, ev->Hs, ev->K[3], 0.25, ev->Nst, s)) != NLO_OK) return rc;
    if ((rc = enqueue_rhs(ev, ev->Nst, ev->K[1], s)) != NLO_OK) return rc;
    if ((rc = launch_axpy(ev, ev->Hs, ev->K[1], 0.5, ev->T, s)) != NLO_OK) return rc;
    hipLaunchKernelGGL(error_ratio_kernel, dim3(1), dim3(1024), 0, s, R, ev->T, ev->F, ev->d.abserr,
                       ev->d.relerr, ev->ctl);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

__global__ void zero_f64_kernel(double* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Captures `enqueue` on the evolver's capture stream into an executable graph.
template <typename F>
int capture_graph(nlo_bk_evolver* ev, F&& enqueue, hipGraph_t* graph, hipGraphExec_t* exec) {
    hipError_t e = hipStreamBeginCapture(ev->capture, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        set_error("hipStreamBeginCapture: %s", hipGetErrorString(e));
        return NLO_ERR_HIP;
    }
    const int rc = enqueue(ev, ev->capture);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(ev->capture, &g);
    if (rc != NLO_OK || e != hipSuccess) {
        if (rc == NLO_OK) set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
        if (g) (void)hipGraphDestroy(g);
        return rc != NLO_OK ? rc : NLO_ERR_HIP;
    }
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
        (void)hipGraphDestroy(g);
        return NLO_ERR_HIP;
    }
    *graph = g;
    return NLO_OK;
}

void release(nlo_bk_evolver* ev) {
    if (ev->exec) (void)hipGraphExecDestroy(ev->exec);
    if (ev->graph) (void)hipGraphDestroy(ev->graph);
    if (ev->exec_k0) (void)hipGraphExecDestroy(ev->exec_k0);
    if (ev->graph_k0) (void)hipGraphDestroy(ev->graph_k0);
    if (ev->capture) (void)hipStreamDestroy(ev->capture);
    if (ev->host_ctl) (void)hipHostFree(ev->host_ctl);
    delete ev;
}

// _adaptive_rk2_to (S/bk/evolution.py:239-326) from ev->y to ev->y + step.
int adaptive_output_step(nlo_bk_evolver* ev, hipStream_t stream) {
    const nlo_bk_evolver_desc& d = ev->d;
    const double nexty = ev->y + d.step;
    while (ev->y < nexty) {
        const double step = std::fmin(ev->h, nexty - ev->y);
        if (!ev->k0_valid) {
            NLO_HIP_CHECK(hipGraphLaunch(ev->exec_k0, stream));
            ev->k0_valid = true;
        }
        ev->host_ctl[0] = step;
        NLO_HIP_CHECK(hipMemcpyAsync(ev->ctl, ev->host_ctl, sizeof(double), hipMemcpyHostToDevice, stream));
        NLO_HIP_CHECK(hipGraphLaunch(ev->exec, stream));
        NLO_HIP_CHECK(hipMemcpyAsync(ev->host_ctl + 1, ev->ctl + 1, sizeof(double), hipMemcpyDeviceToHost,
                                     stream));
        NLO_HIP_CHECK(hipStreamSynchronize(stream));
        const double ratio = ev->host_ctl[1];
        if (ratio <= 1.0) {
            hipLaunchKernelGGL(copy_f64_kernel, dim3((d.R + 255) / 256), dim3(256), 0, stream, d.R, ev->T,
                               ev->N);
            NLO_LAUNCH_CHECK();
            ev->k0_valid = false;
            ev->y += step;
            const double factor = ratio == 0.0 ? 5.0 : std::fmin(5.0, 0.9 * std::pow(ratio, -1.0 / 3.0));
            ev->h = step * factor;
            ++ev->accepted;
        } else {
            // fmax ignores a NaN ratio: max(0.1, nan) = 0.1 in the reference's Python
            ev->h = step * std::fmax(0.1, 0.9 * std::pow(ratio, -1.0 / 3.0));
            ++ev->rejected;
            if (ev->h <= d.eps * std::fmax(1.0, std::fabs(ev->y))) {
                set_error("BK RK2 step underflow at y=%.17g", ev->y);
                return NLO_ERR_INVALID;
            }
        }
    }
    hipLaunchKernelGGL(clamp_flag_kernel, dim3((d.R + 255) / 256), dim3(256), 0, stream, d.R, ev->N,
                       d.force_positive, ev->flag);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

}  // namespace
}  // namespace nlo

extern "C" {

size_t nlo_bk_evolver_workspace(const nlo_bk_evolver_desc* desc) {
    if (!desc || desc->R < 3 || desc->R > nlo::kMaxR || desc->bins < 2 || desc->n_samples < 1) return 0;
    return nlo::evolver_layout(*desc).total;
}

int nlo_bk_evolver_create(const nlo_bk_evolver_desc* desc, const double* d_r, const double* d_edges,
                          void* d_workspace, size_t workspace_bytes, hipStream_t stream,
                          nlo_bk_evolver** out) {
    NLO_REQUIRE(desc && out, "nlo_bk_evolver_create: null argument");
    *out = nullptr;
    const nlo_bk_evolver_desc& d = *desc;
    NLO_REQUIRE(d.R >= 3 && d.R <= nlo::kMaxR, "nlo_bk_evolver_create: R=%d outside [3, %d]", d.R,
                nlo::kMaxR);
    NLO_REQUIRE(d.bins >= 2 && d.bins <= 255, "nlo_bk_evolver_create: bins=%d outside [2, 255]", d.bins);
    NLO_REQUIRE(d.n_samples >= 1 && d.first_index >= 0 && d.first_index % 64 == 0 &&
                    d.first_index + d.n_samples <= (int64_t(1) << 30),
                "nlo_bk_evolver_create: invalid Sobol range");
    NLO_REQUIRE(d.precision == NLO_PREC_FP64 || d.precision == NLO_PREC_FP32 ||
                    d.precision == NLO_PREC_FP64_STABLE,
                "nlo_bk_evolver_create: unknown precision %d", d.precision);
    NLO_REQUIRE(d.rk_method >= NLO_RK_EULER && d.rk_method <= NLO_RK_ADAPTIVE2,
                "nlo_bk_evolver_create: unknown rk_method %d", d.rk_method);
    NLO_REQUIRE(d.step > 0.0, "nlo_bk_evolver_create: step must be positive");
    NLO_REQUIRE(d.nr >= 1 && d.nr <= nlo::kMaxGaussR && d.nt >= 1 && d.nt <= nlo::kMaxGaussT,
                "nlo_bk_evolver_create: GL orders invalid");
    NLO_REQUIRE(d.sobol.dim == 4, "nlo_bk_evolver_create: 4-dimensional Sobol table required");
    NLO_REQUIRE(d.k1_rule == NLO_K1_FIXED || d.k1_rule == NLO_K1_ADAPTIVE,
                "nlo_bk_evolver_create: unknown k1_rule %d", d.k1_rule);
    if (d.k1_rule == NLO_K1_ADAPTIVE)
        NLO_REQUIRE(d.k1_rint >= 1 && d.k1_rint <= nlo::kMaxGkIntervals && d.k1_tint >= 1 &&
                        d.k1_tint <= nlo::kMaxGkIntervals && d.k1_epsrel > 0.0,
                    "nlo_bk_evolver_create: adaptive K1 needs interval limits in [1, %d] and epsrel > 0",
                    nlo::kMaxGkIntervals);
    NLO_REQUIRE(d_r && d_edges && d_workspace, "nlo_bk_evolver_create: null pointer");
    if (d.rk_method == NLO_RK_ADAPTIVE2)
        NLO_REQUIRE(d.abserr >= 0.0 && d.relerr >= 0.0 && d.abserr + d.relerr > 0.0 && d.eps > 0.0,
                    "nlo_bk_evolver_create: adaptive RK2 needs positive tolerances and eps");
    if (d.kinematical) {
        NLO_REQUIRE(d.rk_method == NLO_RK_EULER,
                    "nlo_bk_evolver_create: the kinematical constraint requires the Euler method");
        NLO_REQUIRE(d.history_capacity >= 1 && d.d_history_y && d.d_history_N,
                    "nlo_bk_evolver_create: the kinematical constraint needs history buffers");
    }
    int rc = nlo::validate_physics(&d.physics);
    if (rc != NLO_OK) return rc;
    const nlo::EvolverLayout L = nlo::evolver_layout(d);
    if (workspace_bytes < L.total) {
        nlo::set_error("nlo_bk_evolver_create: workspace %zu bytes < %zu", workspace_bytes, L.total);
        return NLO_ERR_WORKSPACE;
    }
    rc = nlo::init_kernel_attributes();
    if (rc != NLO_OK) return rc;

    nlo_bk_evolver* ev = new nlo_bk_evolver{};
    ev->d = d;
    ev->c = nlo::make_consts(d.physics);
    ev->r = d_r;
    ev->edges = d_edges;
    char* w = static_cast<char*>(d_workspace);
    ev->N = reinterpret_cast<double*>(w + L.N);
    ev->Nst = reinterpret_cast<double*>(w + L.Nst);
    for (int s = 0; s < 4; ++s) ev->K[s] = reinterpret_cast<double*>(w + L.K) + (size_t)s * d.R;
    ev->F = reinterpret_cast<double*>(w + L.F);
    ev->Hs = reinterpret_cast<double*>(w + L.Hs);
    ev->T = reinterpret_cast<double*>(w + L.T);
    ev->ctl = reinterpret_cast<double*>(w + L.ctl);
    ev->k1 = reinterpret_cast<double*>(w + L.k1);
    ev->k2kf = reinterpret_cast<double*>(w + L.k2kf);
    ev->sums = reinterpret_cast<double*>(w + L.sums);
    ev->t64 = reinterpret_cast<double*>(w + L.t64);
    ev->t32 = reinterpret_cast<float*>(w + L.t32);
    ev->flag = reinterpret_cast<int32_t*>(w + L.flag);
    ev->rows = reinterpret_cast<int32_t*>(w + L.rows);
    ev->active = reinterpret_cast<unsigned long long*>(w + L.active);
    ev->nlo_ws = w + L.nlo_ws;
    ev->nlo_ws_bytes = L.nlo_ws_bytes;
    ev->y = 0.0;
    ev->h = d.step;
    ev->rows_host = 1;

    // Zero the state that an LO-only step reads, the non-finite flag and the
    // controller slots; one history row in use.  Fill kernels on the caller's
    // stream (no memset nodes anywhere).
    hipLaunchKernelGGL(nlo::zero_f64_kernel, dim3((4 * d.R + 255) / 256), dim3(256), 0, stream,
                       ev->sums, 4 * d.R);
    hipLaunchKernelGGL(nlo::zero_f64_kernel, dim3(1), dim3(64), 0, stream, ev->ctl, 4);
    rc = nlo::launch_fill_i32(ev->flag, 0, 1, stream);
    if (rc == NLO_OK) rc = nlo::launch_fill_i32(ev->rows, 1, 1, stream);
    if (rc == NLO_OK) rc = nlo::launch_fill_u64(ev->active, 0ull, 1, stream);
    if (rc != NLO_OK) {
        nlo::release(ev);
        return rc;
    }
    if (d.rk_method == NLO_RK_ADAPTIVE2) {
        hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&ev->host_ctl), 2 * sizeof(double),
                                     hipHostMallocDefault);
        if (e != hipSuccess) {
            nlo::set_error("hipHostMalloc: %s", hipGetErrorString(e));
            ev->host_ctl = nullptr;
            nlo::release(ev);
            return NLO_ERR_HIP;
        }
    }

    hipError_t e = hipStreamCreateWithFlags(&ev->capture, hipStreamNonBlocking);
    if (e != hipSuccess) {
        nlo::set_error("hipStreamCreateWithFlags: %s", hipGetErrorString(e));
        ev->capture = nullptr;
        nlo::release(ev);
        return NLO_ERR_HIP;
    }
    if (d.rk_method == NLO_RK_ADAPTIVE2) {
        rc = nlo::capture_graph(
            ev, [](nlo_bk_evolver* v, hipStream_t s) { return nlo::enqueue_attempt(v, s); }, &ev->graph,
            &ev->exec);
        if (rc == NLO_OK)
            rc = nlo::capture_graph(
                ev, [](nlo_bk_evolver* v, hipStream_t s) { return nlo::enqueue_rhs(v, v->N, v->K[0], s); },
                &ev->graph_k0, &ev->exec_k0);
    } else {
        rc = nlo::capture_graph(
            ev, [](nlo_bk_evolver* v, hipStream_t s) { return nlo::enqueue_step(v, s); }, &ev->graph,
            &ev->exec);
    }
    if (rc != NLO_OK) {
        nlo::release(ev);
        return rc;
    }
    *out = ev;
    nlo::clear_error();
    return NLO_OK;
}

double* nlo_bk_evolver_state(nlo_bk_evolver* ev) { return ev ? ev->N : nullptr; }

int32_t* nlo_bk_evolver_nonfinite_flag(nlo_bk_evolver* ev) { return ev ? ev->flag : nullptr; }

const double* nlo_bk_evolver_k1(nlo_bk_evolver* ev) { return ev ? ev->k1 : nullptr; }

uint64_t* nlo_bk_evolver_active_evaluations(nlo_bk_evolver* ev) {
    return ev ? reinterpret_cast<uint64_t*>(ev->active) : nullptr;
}

const double* nlo_bk_evolver_k2kf(nlo_bk_evolver* ev) { return ev ? ev->k2kf : nullptr; }

int nlo_bk_evolver_step(nlo_bk_evolver* ev, hipStream_t stream) {
    NLO_REQUIRE(ev, "nlo_bk_evolver_step: null evolver");
    if (ev->d.rk_method == NLO_RK_ADAPTIVE2) {
        const int rc = nlo::adaptive_output_step(ev, stream);
        if (rc == NLO_OK) nlo::clear_error();
        return rc;
    }
    if (ev->d.kinematical) {
        NLO_REQUIRE(ev->rows_host < ev->d.history_capacity,
                    "nlo_bk_evolver_step: the history holds %d rows, all in use", ev->d.history_capacity);
        ++ev->rows_host;
    }
    NLO_HIP_CHECK(hipGraphLaunch(ev->exec, stream));
    ev->y += ev->d.step;
    ++ev->accepted;
    nlo::clear_error();
    return NLO_OK;
}

int nlo_bk_evolver_adaptive_state(const nlo_bk_evolver* ev, double* y, double* h, int64_t* accepted,
                                  int64_t* rejected) {
    NLO_REQUIRE(ev, "nlo_bk_evolver_adaptive_state: null evolver");
    if (y) *y = ev->y;
    if (h) *h = ev->d.rk_method == NLO_RK_ADAPTIVE2 ? ev->h : ev->d.step;
    if (accepted) *accepted = ev->accepted;
    if (rejected) *rejected = ev->rejected;
    return NLO_OK;
}

int nlo_bk_evolver_destroy(nlo_bk_evolver* ev) {
    if (!ev) return NLO_OK;
    nlo::release(ev);
    return NLO_OK;
}

}  // extern "C"
