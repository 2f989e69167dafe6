// Rapidity driver: one RK step = spline build, K1, K2 + Kf, finalize and the
// stage update per stage, captured once into a hipGraph and replayed per step
// (S/bk/evolution.py:85-159, :190-326).  No host synchronisation, no
// allocation: every buffer lives in the caller's workspace.
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
    double* k1;
    double* k2kf;
    double* sums;
    double* t64;
    float* t32;
    void* nlo_ws;
    size_t nlo_ws_bytes;
    int32_t* flag;
    unsigned long long* active;  // active-parent evaluations since creation
    hipStream_t capture;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

namespace nlo {
namespace {

struct EvolverLayout {
    size_t N, Nst, K, k1, k2kf, sums, t64, t32, flag, active, nlo_ws, nlo_ws_bytes, total;
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
    L.k1 = take(R * sizeof(double));
    L.k2kf = take(R * sizeof(double));
    L.sums = take(4 * R * sizeof(double));
    L.t64 = take((size_t)NLO_SPLINE_WORDS(d.R) * sizeof(double));
    L.t32 = take((size_t)NLO_SPLINE_WORDS(d.R) * sizeof(float));
    L.flag = take(sizeof(int32_t));
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

// Enqueues one RK step on `s` (captured into the graph).
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
        rc = launch_spline_build(R, ev->r, in, d.force_positive, ev->t64, ev->t32, s);
        if (rc != NLO_OK) return rc;
        rc = launch_k1(R, ev->r, in, ev->t64, ev->c, d.nr, d.nt, ev->k1, s);
        if (rc != NLO_OK) return rc;
        if (d.nlo) {
            rc = launch_nlo(R, ev->r, in, ev->t64, ev->t32, ev->edges, d.bins, d.sobol, d.first_index,
                            d.n_samples, d.precision, ev->c, ev->sums, nullptr, nullptr, ev->nlo_ws,
                            ev->nlo_ws_bytes, s, /*components=*/false);
            if (rc != NLO_OK) return rc;
            rc = launch_rhs_finalize(R, in, ev->k1, ev->sums, d.n_samples, ev->K[st], ev->k2kf,
                                     ev->active, s);
        } else {
            // LO-only: K2 + Kf sums stay zero
            rc = launch_rhs_finalize(R, in, ev->k1, ev->sums, 1, ev->K[st], ev->k2kf, ev->active, s);
        }
        if (rc != NLO_OK) return rc;
    }
    return launch_rk_final(R, ev->N, ev->K[0], tb.stages > 1 ? ev->K[1] : nullptr,
                           tb.stages > 2 ? ev->K[2] : nullptr, tb.stages > 3 ? ev->K[3] : nullptr,
                           tb.b[0], tb.b[1], tb.b[2], tb.b[3], d.step, d.force_positive, ev->flag, s);
}

__global__ void zero_f64_kernel(double* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
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
    NLO_REQUIRE(d.rk_method >= NLO_RK_EULER && d.rk_method <= NLO_RK_CLASSIC4,
                "nlo_bk_evolver_create: unknown rk_method %d", d.rk_method);
    NLO_REQUIRE(d.step > 0.0, "nlo_bk_evolver_create: step must be positive");
    NLO_REQUIRE(d.nr >= 1 && d.nr <= nlo::kMaxGaussR && d.nt >= 1 && d.nt <= nlo::kMaxGaussT,
                "nlo_bk_evolver_create: GL orders invalid");
    NLO_REQUIRE(d.sobol.dim == 4, "nlo_bk_evolver_create: 4-dimensional Sobol table required");
    NLO_REQUIRE(d_r && d_edges && d_workspace, "nlo_bk_evolver_create: null pointer");
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
    ev->k1 = reinterpret_cast<double*>(w + L.k1);
    ev->k2kf = reinterpret_cast<double*>(w + L.k2kf);
    ev->sums = reinterpret_cast<double*>(w + L.sums);
    ev->t64 = reinterpret_cast<double*>(w + L.t64);
    ev->t32 = reinterpret_cast<float*>(w + L.t32);
    ev->flag = reinterpret_cast<int32_t*>(w + L.flag);
    ev->active = reinterpret_cast<unsigned long long*>(w + L.active);
    ev->nlo_ws = w + L.nlo_ws;
    ev->nlo_ws_bytes = L.nlo_ws_bytes;

    // Zero the state that an LO-only step reads, and the non-finite flag,
    // with fill kernels on the caller's stream (no memset nodes anywhere).
    hipLaunchKernelGGL(nlo::zero_f64_kernel, dim3((4 * d.R + 255) / 256), dim3(256), 0, stream,
                       ev->sums, 4 * d.R);
    rc = nlo::launch_fill_i32(ev->flag, 0, 1, stream);
    if (rc == NLO_OK) rc = nlo::launch_fill_u64(ev->active, 0ull, 1, stream);
    if (rc != NLO_OK) {
        delete ev;
        return rc;
    }

    hipError_t e = hipStreamCreateWithFlags(&ev->capture, hipStreamNonBlocking);
    if (e != hipSuccess) {
        nlo::set_error("hipStreamCreateWithFlags: %s", hipGetErrorString(e));
        delete ev;
        return NLO_ERR_HIP;
    }
    e = hipStreamBeginCapture(ev->capture, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        nlo::set_error("hipStreamBeginCapture: %s", hipGetErrorString(e));
        (void)hipStreamDestroy(ev->capture);
        delete ev;
        return NLO_ERR_HIP;
    }
    rc = nlo::enqueue_step(ev, ev->capture);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(ev->capture, &g);
    if (rc != NLO_OK || e != hipSuccess) {
        if (rc == NLO_OK) nlo::set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
        if (g) (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(ev->capture);
        delete ev;
        return rc != NLO_OK ? rc : NLO_ERR_HIP;
    }
    ev->graph = g;
    e = hipGraphInstantiate(&ev->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        nlo::set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
        (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(ev->capture);
        delete ev;
        return NLO_ERR_HIP;
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
    NLO_HIP_CHECK(hipGraphLaunch(ev->exec, stream));
    nlo::clear_error();
    return NLO_OK;
}

int nlo_bk_evolver_destroy(nlo_bk_evolver* ev) {
    if (!ev) return NLO_OK;
    if (ev->exec) (void)hipGraphExecDestroy(ev->exec);
    if (ev->graph) (void)hipGraphDestroy(ev->graph);
    if (ev->capture) (void)hipStreamDestroy(ev->capture);
    delete ev;
    return NLO_OK;
}

}  // extern "C"
