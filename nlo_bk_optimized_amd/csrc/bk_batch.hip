// Many BK evolutions of uneven grid sizes advanced together: per RK stage one
// spline-build launch (a block per run), one K1 launch (a block per node of
// every run), one persistent K2 + Kf launch over a device work-list of
// (parent, chunk) items, one reduce and one finalize, all captured into one
// hipGraph per step.  The K2 + Kf launch shape (samples per lane, persistent
// grid size) is chosen on the device by timing candidates on the first step.
// Same right-hand side as the single-run driver (S/bk/evolution.py:85-159),
// per run.
#include <algorithm>
#include <cstring>
#include <vector>

#include "bk_internal.hpp"

struct nlo_bk_batch {
    nlo_bk_batch_desc d;
    nlo::BkConsts c;
    nlo::BatchView b;
    const double* r;
    const double* edges;
    double* N;
    double* Nst;
    double* K[4];
    double* k1;
    double* sums;
    double* t64;
    float* t32;
    double* part;
    int32_t* flag;
    unsigned long long* active;
    unsigned long long* counter;
    int spt, grid;
    double tuned_ms;
    bool captured;
    hipStream_t capture;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

namespace nlo {
namespace {

constexpr int kSptCandidates[] = {4, 8, 16};
constexpr int kGridFactors[] = {1, 2, 4};

struct BatchLayout {
    size_t N, Nst, K, k1, sums, t64, t32, part, flag, active, counter, runs, node_run, parent_node, total;
};

struct BatchShape {
    int runs = 0, nodes = 0, parents = 0, Rmax = 0, words = 0;
};

BatchShape batch_shape(const nlo_bk_batch_desc& d, const int32_t* R) {
    BatchShape s;
    s.runs = d.runs;
    for (int i = 0; i < d.runs; ++i) {
        s.nodes += R[i];
        s.parents += R[i] - 1;
        s.Rmax = std::max(s.Rmax, (int)R[i]);
        s.words += NLO_SPLINE_WORDS(R[i]);
    }
    return s;
}

BatchLayout batch_layout(const nlo_bk_batch_desc& d, const BatchShape& s) {
    BatchLayout L{};
    size_t o = 0;
    auto take = [&o](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes, 256);
        return at;
    };
    const size_t n = (size_t)s.nodes;
    L.N = take(n * sizeof(double));
    L.Nst = take(n * sizeof(double));
    L.K = take(4 * n * sizeof(double));
    L.k1 = take(n * sizeof(double));
    L.sums = take(4 * n * sizeof(double));
    L.t64 = take((size_t)s.words * sizeof(double));
    L.t32 = take((size_t)s.words * sizeof(float));
    // partials for the smallest samples-per-lane candidate (most chunks)
    const int64_t bpp = (d.n_samples + 256 * kSptCandidates[0] - 1) / (256 * kSptCandidates[0]);
    L.part = take(d.nlo ? (size_t)s.parents * (size_t)bpp * 4 * sizeof(double) : 1);
    L.flag = take(sizeof(int32_t));
    L.active = take(sizeof(unsigned long long));
    L.counter = take(sizeof(unsigned long long));
    L.runs = take((size_t)s.runs * sizeof(BatchRun));
    L.node_run = take(n * sizeof(int32_t));
    L.parent_node = take((size_t)s.parents * sizeof(int32_t));
    L.total = o;
    return L;
}

struct Tableau {
    int stages;
    double a[4][3];
    double b[4];
};

Tableau batch_tableau(int method) {
    Tableau t{};
    switch (method) {
        case NLO_RK_EULER:
            t.stages = 1;
            t.b[0] = 1.0;
            break;
        case NLO_RK_MIDPOINT:
            t.stages = 2;
            t.a[1][0] = 0.5;
            t.b[1] = 1.0;
            break;
        case NLO_RK_KUTTA3:
            t.stages = 3;
            t.a[1][0] = 0.5;
            t.a[2][0] = -1.0;
            t.a[2][1] = 2.0;
            t.b[0] = 1.0 / 6.0;
            t.b[1] = 4.0 / 6.0;
            t.b[2] = 1.0 / 6.0;
            break;
        default:
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

int enqueue_batch_rhs(nlo_bk_batch* bt, const double* in, double* out, hipStream_t s) {
    const nlo_bk_batch_desc& d = bt->d;
    int rc = launch_spline_build_batch(bt->b, bt->r, in, d.force_positive, bt->t64, bt->t32, s);
    if (rc != NLO_OK) return rc;
    if (d.k1_rule == NLO_K1_ADAPTIVE)
        rc = launch_k1_adaptive_batch(bt->b, bt->r, in, bt->t64, bt->c, d.k1_rint, d.k1_tint, d.k1_epsrel,
                                      bt->k1, s);
    else
        rc = launch_k1_batch(bt->b, bt->r, in, bt->t64, bt->c, d.nr, d.nt, bt->k1, s);
    if (rc != NLO_OK) return rc;
    if (d.nlo) {
        rc = launch_nlo_batch(bt->b, bt->r, in, bt->t64, bt->t32, bt->edges, d.bins, d.sobol, d.first_index,
                              d.n_samples, d.precision, bt->c, bt->spt, bt->grid, bt->counter, bt->part,
                              bt->sums, s);
        if (rc != NLO_OK) return rc;
        return launch_rhs_finalize_batch(bt->b, in, bt->k1, bt->sums, d.n_samples, out, bt->active, s);
    }
    return launch_rhs_finalize_batch(bt->b, in, bt->k1, bt->sums, 1, out, bt->active, s);
}

int enqueue_batch_step(nlo_bk_batch* bt, hipStream_t s) {
    const nlo_bk_batch_desc& d = bt->d;
    const Tableau tb = batch_tableau(d.rk_method);
    const int n = bt->b.nodes;
    int rc;
    for (int st = 0; st < tb.stages; ++st) {
        const double* in = bt->N;
        if (st > 0) {
            rc = launch_rk_stage(n, bt->N, bt->K[0], st > 1 ? bt->K[1] : nullptr, st > 2 ? bt->K[2] : nullptr,
                                 tb.a[st][0], tb.a[st][1], tb.a[st][2], d.step, bt->Nst, s);
            if (rc != NLO_OK) return rc;
            in = bt->Nst;
        }
        rc = enqueue_batch_rhs(bt, in, bt->K[st], s);
        if (rc != NLO_OK) return rc;
    }
    return launch_rk_final(n, bt->N, bt->K[0], tb.stages > 1 ? bt->K[1] : nullptr,
                           tb.stages > 2 ? bt->K[2] : nullptr, tb.stages > 3 ? bt->K[3] : nullptr, tb.b[0],
                           tb.b[1], tb.b[2], tb.b[3], d.step, d.force_positive, bt->flag, s);
}

// Times the K2 + Kf launch for every (samples per lane, grid) candidate on the
// current state and keeps the fastest.  One warm launch, then two timed ones.
int autotune(nlo_bk_batch* bt, hipStream_t s) {
    const nlo_bk_batch_desc& d = bt->d;
    if (!d.nlo) {
        bt->spt = d.spt > 0 ? d.spt : 16;
        bt->grid = d.grid_blocks > 0 ? d.grid_blocks : 1;
        return NLO_OK;
    }
    int rc = launch_spline_build_batch(bt->b, bt->r, bt->N, d.force_positive, bt->t64, bt->t32, s);
    if (rc != NLO_OK) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    NLO_HIP_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error("hipEventCreate failed");
        return NLO_ERR_HIP;
    }
    double best = -1.0;
    int best_spt = 16, best_grid = 1;
    for (int spt : kSptCandidates) {
        if (d.spt > 0 && spt != d.spt) continue;
        const int resident = nlo_batch_resident_blocks(d.precision, bt->c, bt->b.Rmax, d.bins, spt);
        for (int f : kGridFactors) {
            if (d.grid_blocks > 0 && f != kGridFactors[0]) continue;
            const int grid = d.grid_blocks > 0 ? d.grid_blocks : std::max(1, resident * f);
            for (int k = 0; k < 3 && rc == NLO_OK; ++k) {
                if (k == 1 && hipEventRecord(e0, s) != hipSuccess) rc = NLO_ERR_HIP;
                if (rc == NLO_OK)
                    rc = launch_nlo_batch(bt->b, bt->r, bt->N, bt->t64, bt->t32, bt->edges, d.bins, d.sobol,
                                          d.first_index, d.n_samples, d.precision, bt->c, spt, grid, bt->counter,
                                          bt->part, bt->sums, s);
            }
            float ms = 0.0f;
            if (rc == NLO_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                 hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = NLO_ERR_HIP;
            if (rc != NLO_OK) break;
            if (best < 0.0 || ms / 2.0 < best) {
                best = ms / 2.0;
                best_spt = spt;
                best_grid = grid;
            }
        }
        if (rc != NLO_OK) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != NLO_OK) {
        if (rc == NLO_ERR_HIP) set_error("autotuning the batched K2 + Kf launch failed");
        return rc;
    }
    bt->spt = best_spt;
    bt->grid = best_grid;
    bt->tuned_ms = best;
    return NLO_OK;
}

int capture_batch(nlo_bk_batch* bt) {
    hipError_t e = hipStreamBeginCapture(bt->capture, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        set_error("hipStreamBeginCapture: %s", hipGetErrorString(e));
        return NLO_ERR_HIP;
    }
    const int rc = enqueue_batch_step(bt, bt->capture);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(bt->capture, &g);
    if (rc != NLO_OK || e != hipSuccess) {
        if (rc == NLO_OK) set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
        if (g) (void)hipGraphDestroy(g);
        return rc != NLO_OK ? rc : NLO_ERR_HIP;
    }
    e = hipGraphInstantiate(&bt->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
        (void)hipGraphDestroy(g);
        return NLO_ERR_HIP;
    }
    bt->graph = g;
    bt->captured = true;
    return NLO_OK;
}

__global__ void zero_batch_kernel(double* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

void release_batch(nlo_bk_batch* bt) {
    if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
    if (bt->graph) (void)hipGraphDestroy(bt->graph);
    if (bt->capture) (void)hipStreamDestroy(bt->capture);
    delete bt;
}

}  // namespace
}  // namespace nlo

extern "C" {

size_t nlo_bk_batch_workspace(const nlo_bk_batch_desc* desc, const int32_t* h_R) {
    if (!desc || !h_R || desc->runs < 1 || desc->bins < 2 || desc->n_samples < 1) return 0;
    for (int i = 0; i < desc->runs; ++i)
        if (h_R[i] < 3 || h_R[i] > nlo::kMaxR) return 0;
    return nlo::batch_layout(*desc, nlo::batch_shape(*desc, h_R)).total;
}

int nlo_bk_batch_create(const nlo_bk_batch_desc* desc, const int32_t* h_R, const double* d_r,
                        const double* d_edges, void* d_workspace, size_t workspace_bytes, hipStream_t stream,
                        nlo_bk_batch** out) {
    NLO_REQUIRE(desc && h_R && out, "nlo_bk_batch_create: null argument");
    *out = nullptr;
    const nlo_bk_batch_desc& d = *desc;
    NLO_REQUIRE(d.runs >= 1, "nlo_bk_batch_create: at least one run");
    for (int i = 0; i < d.runs; ++i)
        NLO_REQUIRE(h_R[i] >= 3 && h_R[i] <= nlo::kMaxR, "nlo_bk_batch_create: run %d has R=%d outside [3, %d]",
                    i, h_R[i], nlo::kMaxR);
    NLO_REQUIRE(d.bins >= 2 && d.bins <= 255, "nlo_bk_batch_create: bins=%d outside [2, 255]", d.bins);
    NLO_REQUIRE(d.n_samples >= 1 && d.first_index >= 0 && d.first_index % 64 == 0 &&
                    d.first_index + d.n_samples <= (int64_t(1) << 30),
                "nlo_bk_batch_create: invalid Sobol range");
    NLO_REQUIRE(d.precision == NLO_PREC_FP64 || d.precision == NLO_PREC_FP32 ||
                    d.precision == NLO_PREC_FP64_STABLE,
                "nlo_bk_batch_create: unknown precision %d", d.precision);
    NLO_REQUIRE(d.rk_method >= NLO_RK_EULER && d.rk_method <= NLO_RK_CLASSIC4,
                "nlo_bk_batch_create: fixed-step methods only (rk_method %d)", d.rk_method);
    NLO_REQUIRE(d.step > 0.0, "nlo_bk_batch_create: step must be positive");
    NLO_REQUIRE(d.nr >= 1 && d.nr <= nlo::kMaxGaussR && d.nt >= 1 && d.nt <= nlo::kMaxGaussT,
                "nlo_bk_batch_create: GL orders invalid");
    NLO_REQUIRE(d.spt == 0 || d.spt == 4 || d.spt == 8 || d.spt == 16,
                "nlo_bk_batch_create: samples per lane must be 0 (tune), 4, 8 or 16");
    NLO_REQUIRE(d.grid_blocks >= 0, "nlo_bk_batch_create: grid_blocks must be >= 0");
    NLO_REQUIRE(d.sobol.dim == 4, "nlo_bk_batch_create: 4-dimensional Sobol table required");
    NLO_REQUIRE(d.k1_rule == NLO_K1_FIXED || d.k1_rule == NLO_K1_ADAPTIVE,
                "nlo_bk_batch_create: unknown k1_rule %d", d.k1_rule);
    if (d.k1_rule == NLO_K1_ADAPTIVE)
        NLO_REQUIRE(d.k1_rint >= 1 && d.k1_rint <= nlo::kMaxGkIntervals && d.k1_tint >= 1 &&
                        d.k1_tint <= nlo::kMaxGkIntervals && d.k1_epsrel > 0.0,
                    "nlo_bk_batch_create: adaptive K1 needs interval limits in [1, %d] and epsrel > 0",
                    nlo::kMaxGkIntervals);
    NLO_REQUIRE(d_r && d_edges && d_workspace, "nlo_bk_batch_create: null pointer");
    int rc = nlo::validate_physics(&d.physics);
    if (rc != NLO_OK) return rc;
    const nlo::BatchShape sh = nlo::batch_shape(d, h_R);
    const nlo::BatchLayout L = nlo::batch_layout(d, sh);
    if (workspace_bytes < L.total) {
        nlo::set_error("nlo_bk_batch_create: workspace %zu bytes < %zu", workspace_bytes, L.total);
        return NLO_ERR_WORKSPACE;
    }
    rc = nlo::init_kernel_attributes();
    if (rc != NLO_OK) return rc;

    // host metadata: runs, the run of every node, the node of every parent
    std::vector<nlo::BatchRun> runs(d.runs);
    std::vector<int32_t> node_run(sh.nodes), parent_node(sh.parents);
    int node = 0, words = 0, parent = 0;
    for (int i = 0; i < d.runs; ++i) {
        runs[i] = nlo::BatchRun{h_R[i], node, words, parent};
        for (int p = 0; p < h_R[i]; ++p) {
            node_run[node + p] = i;
            if (p >= 1) parent_node[parent++] = node + p;
        }
        node += h_R[i];
        words += NLO_SPLINE_WORDS(h_R[i]);
    }

    nlo_bk_batch* bt = new nlo_bk_batch{};
    bt->d = d;
    bt->c = nlo::make_consts(d.physics);
    bt->r = d_r;
    bt->edges = d_edges;
    char* w = static_cast<char*>(d_workspace);
    bt->N = reinterpret_cast<double*>(w + L.N);
    bt->Nst = reinterpret_cast<double*>(w + L.Nst);
    for (int s = 0; s < 4; ++s) bt->K[s] = reinterpret_cast<double*>(w + L.K) + (size_t)s * sh.nodes;
    bt->k1 = reinterpret_cast<double*>(w + L.k1);
    bt->sums = reinterpret_cast<double*>(w + L.sums);
    bt->t64 = reinterpret_cast<double*>(w + L.t64);
    bt->t32 = reinterpret_cast<float*>(w + L.t32);
    bt->part = reinterpret_cast<double*>(w + L.part);
    bt->flag = reinterpret_cast<int32_t*>(w + L.flag);
    bt->active = reinterpret_cast<unsigned long long*>(w + L.active);
    bt->counter = reinterpret_cast<unsigned long long*>(w + L.counter);
    nlo::BatchRun* runs_d = reinterpret_cast<nlo::BatchRun*>(w + L.runs);
    int32_t* node_run_d = reinterpret_cast<int32_t*>(w + L.node_run);
    int32_t* parent_node_d = reinterpret_cast<int32_t*>(w + L.parent_node);
    bt->b = nlo::BatchView{sh.runs, sh.nodes, sh.parents, sh.Rmax, runs_d, node_run_d, parent_node_d};

    hipError_t e = hipMemcpyAsync(runs_d, runs.data(), runs.size() * sizeof(nlo::BatchRun),
                                  hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(node_run_d, node_run.data(), node_run.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                           stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(parent_node_d, parent_node.data(), parent_node.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);  // the host vectors go out of scope
    if (e != hipSuccess) {
        nlo::set_error("nlo_bk_batch_create: metadata upload: %s", hipGetErrorString(e));
        nlo::release_batch(bt);
        return NLO_ERR_HIP;
    }
    hipLaunchKernelGGL(nlo::zero_batch_kernel, dim3((4 * sh.nodes + 255) / 256), dim3(256), 0, stream, bt->sums,
                       4 * sh.nodes);
    rc = nlo::launch_fill_i32(bt->flag, 0, 1, stream);
    if (rc == NLO_OK) rc = nlo::launch_fill_u64(bt->active, 0ull, 1, stream);
    if (rc == NLO_OK && hipStreamCreateWithFlags(&bt->capture, hipStreamNonBlocking) != hipSuccess) {
        bt->capture = nullptr;
        nlo::set_error("hipStreamCreateWithFlags failed");
        rc = NLO_ERR_HIP;
    }
    if (rc != NLO_OK) {
        nlo::release_batch(bt);
        return rc;
    }
    *out = bt;
    nlo::clear_error();
    return NLO_OK;
}

double* nlo_bk_batch_state(nlo_bk_batch* bt) { return bt ? bt->N : nullptr; }

int32_t* nlo_bk_batch_nonfinite_flag(nlo_bk_batch* bt) { return bt ? bt->flag : nullptr; }

uint64_t* nlo_bk_batch_active_evaluations(nlo_bk_batch* bt) {
    return bt ? reinterpret_cast<uint64_t*>(bt->active) : nullptr;
}

int nlo_bk_batch_step(nlo_bk_batch* bt, hipStream_t stream) {
    NLO_REQUIRE(bt, "nlo_bk_batch_step: null batch");
    if (!bt->captured) {
        // first step: tune the K2 + Kf launch on the state the caller set,
        // then capture the step with the chosen shape
        int rc = nlo::autotune(bt, stream);
        if (rc == NLO_OK) rc = nlo::capture_batch(bt);
        if (rc != NLO_OK) return rc;
    }
    NLO_HIP_CHECK(hipGraphLaunch(bt->exec, stream));
    nlo::clear_error();
    return NLO_OK;
}

int nlo_bk_batch_tuning(const nlo_bk_batch* bt, int32_t* spt, int32_t* grid_blocks, double* ms_per_launch) {
    NLO_REQUIRE(bt, "nlo_bk_batch_tuning: null batch");
    NLO_REQUIRE(bt->captured, "nlo_bk_batch_tuning: the launch shape is chosen on the first step");
    if (spt) *spt = bt->spt;
    if (grid_blocks) *grid_blocks = bt->grid;
    if (ms_per_launch) *ms_per_launch = bt->tuned_ms;
    return NLO_OK;
}

int nlo_bk_batch_destroy(nlo_bk_batch* bt) {
    if (!bt) return NLO_OK;
    nlo::release_batch(bt);
    return NLO_OK;
}

}  // extern "C"
