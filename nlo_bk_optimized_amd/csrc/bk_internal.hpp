// Internal launch helpers shared by the C ABI entry points and the rapidity
// driver (which captures the same launches into a hipGraph).
#pragma once

#include "bk_device.hpp"

namespace nlo {

constexpr int kMaxR = 2048;       // r nodes (LDS: FP64 spline + Thomas scratch)
constexpr int kMaxGaussR = 32;    // K1 radial GL order
constexpr int kMaxGaussT = 64;    // K1 angular GL order
constexpr int kMaxGkIntervals = 100;  // adaptive K1: intervals per list (RINTPOINTS, THETAINTPOINTS)

struct NloPlan {
    int spt;              // samples per lane
    int bpp;              // blocks per parent
    size_t part_bytes;    // per-block sums
    size_t hist_bytes;    // per-block histograms (0 without histogram)
    size_t workspace;     // total workspace
};

BkConsts make_consts(const nlo_bk_physics& p);
int validate_physics(const nlo_bk_physics* p);
// Raises the dynamic-LDS limit of the kernels that may need more than 64 KiB
// (once per process; must run before any stream capture).
int init_kernel_attributes();

NloPlan nlo_plan(int R, int bins, int64_t n, bool hist);

int launch_spline_build(int R, const double* r, const double* N, int force_positive, double* t64,
                        float* t32, hipStream_t stream);
int launch_k1(int R, const double* r, const double* N, const double* t64, const BkConsts& c, int nr,
              int nt, double* k1, hipStream_t stream, const K1History* hist = nullptr);
int launch_k1_adaptive(int R, const double* r, const double* N, const double* t64, const BkConsts& c,
                       int rint, int tint, double epsrel, double* k1, double* k1_err,
                       long long* k1_evals, hipStream_t stream, const K1History* hist = nullptr);
int launch_nlo(int R, const double* r, const double* N, const double* t64, const float* t32,
               const double* edges, int bins, const nlo_sobol_table& sobol, int64_t first, int64_t n,
               int precision, const BkConsts& c, double* sums, double* hist, double* count, void* ws,
               size_t ws_bytes, hipStream_t stream, bool components = true);
int launch_rhs_finalize(int R, const double* N, const double* k1, const double* sums, int64_t n,
                        double* dndy, double* k2kf, unsigned long long* counter, hipStream_t stream);
int launch_rk_stage(int R, const double* N, const double* K0, const double* K1, const double* K2,
                    double a0, double a1, double a2, double h, double* out, hipStream_t stream);
int launch_rk_final(int R, double* N, const double* K0, const double* K1, const double* K2,
                    const double* K3, double b0, double b1, double b2, double b3, double h,
                    int force_positive, int32_t* flag, hipStream_t stream);
int launch_fill_i32(int32_t* p, int32_t v, int n, hipStream_t stream);

// ------------------------------------------------------------ batches
// Many evolutions in one set of launches: their r, N, K1, dN/dy arrays are
// concatenated (node_off), their spline tables too (spline_off), and their
// parents p >= 1 form one list (parent_off + p - 1).  Edges: 4 (bins + 1)
// doubles per node; NLO partials: bpp x 4 doubles per parent.
struct BatchRun {
    int R;
    int node_off;
    int spline_off;
    int parent_off;
};

struct BatchView {
    int runs, nodes, parents, Rmax;
    const BatchRun* runs_d;   // device: runs
    const int32_t* node_run;  // device: run of each node
    const int32_t* parent_node;  // device: global node of each parent
};

int launch_spline_build_batch(const BatchView& b, const double* r, const double* N, int force_positive,
                              double* t64, float* t32, hipStream_t stream);
int launch_k1_adaptive_batch(const BatchView& b, const double* r, const double* N, const double* t64,
                             const BkConsts& c, int rint, int tint, double epsrel, double* k1,
                             hipStream_t stream);
int launch_k1_batch(const BatchView& b, const double* r, const double* N, const double* t64,
                    const BkConsts& c, int nr, int nt, double* k1, hipStream_t stream);
// Persistent K2 + Kf launch: `grid` blocks take (parent, chunk) items from
// *counter (reset here by a fill kernel) until parents x bpp items are done.
size_t nlo_batch_partials(const BatchView& b, int64_t n, int spt);
int launch_nlo_batch(const BatchView& b, const double* r, const double* N, const double* t64,
                     const float* t32, const double* edges, int bins, const nlo_sobol_table& sobol,
                     int64_t first, int64_t n, int precision, const BkConsts& c, int spt, int grid,
                     unsigned long long* counter, double* part, double* sums, hipStream_t stream);
int launch_rhs_finalize_batch(const BatchView& b, const double* N, const double* k1, const double* sums,
                              int64_t n, double* dndy, unsigned long long* counter, hipStream_t stream);
// Blocks of the persistent NLO kernel resident on the device at once.
int nlo_batch_resident_blocks(int precision, const BkConsts& c, int Rmax, int bins, int spt);
int launch_fill_u64(unsigned long long* p, unsigned long long v, int n, hipStream_t stream);

}  // namespace nlo
