// The GBW fixed-Sobol qqbar-g estimator (gbw_qg_fixed_sobol,
// S/dis/fixed_sobol.py:274-368) and its paired-endpoint Sobol map
// (dis_sobol_endpoint_map, K/dis/sampling/kernel.cu:15-81), batched over any
// number of estimators (point, polarisation, quark).
//
// One estimator: `rounds` training point sets learn one frozen 5-D grid from
// the normalised magnitudes of I1, I2 and I3 (learn_triple_importance_grid,
// :161-210); `replicates` independently scrambled point sets through that grid
// give I1, I2, I3 and their correlated sum as replicate means and standard
// errors (integrate_triple_fixed_sobol, :213-271).  Every point is mapped
// through the grid and split into the two nearest-endpoint sectors of the
// gluon plane (paired_endpoint_inputs, :116-158); the integrands are
// evaluated at both and combined with the sector weights.
//
// Device layout.  A pass is one launch per integrand group over a work-list of
// (estimator, point set, block): I1 (FP64 on both paths: its UV-subtracted sum
// cancels to O(z2)) and the nested I2 / I3 pair (one pass over the inner rule
// for both) are separate kernels, each with its own register budget, and both
// recompute the cheap Sobol point, grid map and endpoint split.  A wave holds
// 32 points, the two sectors of point l in lanes 2l and 2l + 1, which combine
// through one lane exchange.  Sums are compensated FP64 in a fixed order; the
// training histograms are summed per block in point order (no atomics).
// Reference prefixes: S/ = torch_nlo_clean/src/nlo_torch/, K/ = S/custom_kernels/.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "dis_device.hpp"
#include "dis_integrands.hpp"
#include "vegas_grid.hpp"
#include "nlo_gfx950.h"

namespace nlo {
namespace dis {
namespace {

constexpr int kFsDim = 5;
constexpr int kFsGroupBits = 5;  // 32 points (64 lanes) per wave and group
constexpr int kFsPart = 3;       // sums per block: I1, I2, I3

// -------------------------------------------------- map and endpoint split

// A unit-cube point mapped through the frozen grid and split into the two
// endpoint sectors, in the arithmetic T of the point set's dtype.
template <typename T>
struct EndpointPair {
    T src[kFsDim];  // the mapped point
    T inv;          // inverse density prod_d B (e[b + 1] - e[b])
    T f3, f4, w1;   // first sector: radial and angular coordinates, weight
    T s3, s4, w2;   // second sector
    uint64_t bins;  // bin d in byte d
};

template <typename T>
__device__ __forceinline__ T fs_cos(T x) { return cos(x); }
template <typename T>
__device__ __forceinline__ T fs_sin(T x) { return sin(x); }
template <typename T>
__device__ __forceinline__ T fs_atan2(T y, T x) { return atan2(y, x); }
template <>
__device__ __forceinline__ float fs_cos<float>(float x) { return cosf(x); }
template <>
__device__ __forceinline__ float fs_sin<float>(float x) { return sinf(x); }
template <>
__device__ __forceinline__ float fs_atan2<float>(float y, float x) { return atan2f(y, x); }

// map_importance_grid (S/dis/fixed_sobol.py:100-113) and
// paired_endpoint_inputs (:116-158), operation by operation in T (no
// contraction): parent x01 = maxr s2 along the x axis; sector one, the points
// nearer endpoint 0, in polar coordinates about endpoint 0 out to the
// perpendicular bisector or maxr; sector two, the points nearer endpoint 1, in
// polar coordinates about endpoint 1 out to the bisector or the disk
// |x02| <= maxr, mapped back to (|x02|, angle) about endpoint 0 with the
// Jacobian of that change.  With fold_angle the angle covers [0, pi) (the
// integrands depend on cos phi only).  E: the grid, dim x (bins + 1), in T.
template <typename T>
__device__ __forceinline__ EndpointPair<T> endpoint_pair(const T (&u)[kFsDim], const T* E, int bins, T maxr,
                                                         bool fold) {
#pragma clang fp contract(off)
    EndpointPair<T> o;
    o.inv = T(1);
    o.bins = 0;
    const T B = (T)bins;
#pragma unroll
    for (int d = 0; d < kFsDim; ++d) {
        const T scaled = u[d] * B;
        const int b = min((int)scaled, bins - 1);
        const T frac = scaled - (T)b;
        const T left = E[d * (bins + 1) + b];
        const T width = E[d * (bins + 1) + b + 1] - left;
        o.src[d] = left + width * frac;
        o.inv = o.inv * (B * width);
        o.bins |= (uint64_t)b << (8 * d);
    }
    const T two_pi = T(2.0 * kPiD);
    const T tiny = std::is_same<T, float>::value ? (T)FLT_MIN : (T)DBL_MIN;
    const T parent = maxr * o.src[2];
    const T rf = o.src[3];
    const T af = fold ? T(0.5) * o.src[4] : o.src[4];
    const T angle = two_pi * af;
    const T c = fs_cos<T>(angle), s = fs_sin<T>(angle);
    const T closer = c > T(0) ? parent / (T(2) * c) : T(INFINITY);
    const T l1 = fmin(maxr, closer);
    o.f3 = rf * l1 / maxr;
    o.f4 = af;
    o.w1 = l1 / maxr;
    const T disk = parent * c + sqrt(fmax(maxr * maxr - (parent * parent) * (s * s), T(0)));
    const T l2 = fmax(fmin(disk, closer), T(0));
    const T r2 = rf * l2;
    const T gx = parent - r2 * c;
    const T gy = r2 * s;
    const T gr = sqrt(gx * gx + gy * gy);
    T ga = fs_atan2<T>(gy, gx);
    if (ga < T(0)) ga = ga + two_pi;  // remainder(atan2, 2 pi)
    o.s3 = gr / maxr;
    o.s4 = ga / two_pi;
    o.w2 = r2 * l2 / (fmax(gr, tiny) * maxr);
    return o;
}

template <typename T>
__device__ __forceinline__ T sobol_value(uint32_t q);
template <>
__device__ __forceinline__ float sobol_value<float>(uint32_t q) { return (float)q * 0x1p-30f; }
template <>
__device__ __forceinline__ double sobol_value<double>(uint32_t q) { return (double)q * 0x1p-30; }

// ------------------------------------------------------------- estimator

struct FsDev {
    const DisArgs* args;            // per estimator
    const int* set;                 // per estimator: its table set
    const nlo_sobol_table* tables;  // sets x (rounds + replicates)
    double* edges;                  // estimators x 2 buffers x dim x (bins + 1)
    double* part;                   // estimators x part_stride: per-block partial sums
    double* stat;                   // estimators x stat_stride: reduced sums, scores, estimates
    int bins, rounds, reps, fold, spt;
    int blocks_train, blocks_rep;   // per estimator and point set
    int64_t n_train, n_rep, part_stride, stat_stride;
    double i1_weight;
};

__host__ __device__ inline int fs_cells(int bins) { return kFsDim * bins; }
// training partial of one block: sums of |w| (3), histograms of |w| (3 x
// cells), sample counts (cells)
__host__ __device__ inline int64_t fs_train_part(int bins) { return kFsPart + 4 * (int64_t)fs_cells(bins); }
// reduced statistics: the training partials summed (3 + 4 cells), the score
// histogram (cells), the replicate estimates (reps x 3)
__host__ __device__ inline int64_t fs_stat_stride(int bins, int reps) {
    return fs_train_part(bins) + fs_cells(bins) + (int64_t)reps * kFsPart;
}

constexpr int kFsI1 = 0;      // the I1 pass
constexpr int kFsNested = 1;  // the nested I2 / I3 pass

// Dynamic LDS of a pass: grid (T), Sobol directions, and in training the
// staged |w| (one or two per point) and packed bins of the block's points.
template <typename T>
__host__ __device__ inline size_t fs_lds(int bins, int spt, bool train) {
    size_t o = align_up((size_t)kFsDim * (bins + 1) * sizeof(T), 16);
    o = align_up(o + (size_t)kFsDim * NLO_SOBOL_BITS * sizeof(uint32_t), 16);
    if (train) {
        const size_t pts = (size_t)128 * spt;
        o = align_up(o + 2 * pts * sizeof(double), 16);
        o = align_up(o + pts * sizeof(uint64_t), 16);
    }
    return o;
}

// One block of a pass.  Block blk of point set k holds the points
// [128 spt blk, 128 spt (blk + 1)); wave w walks the 32-point groups
// hi = (4 blk + w) spt + j, j < spt: point hi 32 + l in lanes 2l (sector one)
// and 2l + 1 (sector two).  The gray code of point s = 32 hi + l is
// gray(l) ^ (bit 0 of hi) << 4 in its low five bits and gray(hi) above, so the
// high part is wave-uniform and changes by one direction per group.
template <typename T, int PART, bool TRAIN>
__device__ __forceinline__ void fs_pass(const FsDev& f, int item, int k, int blk, int64_t n, const double* grid,
                                        double* part) {
    constexpr int NK = PART == kFsI1 ? 1 : 2;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[4][NK][2];
    __shared__ double ry[kMaxNested], rw[kMaxNested];
    const DisArgs& a = f.args[item];
    RuleView rule{0, ry, rw};
    if constexpr (PART == kFsNested) rule = stage_nested_rule(a.nested_points, ry, rw);
    const int bins = f.bins, spt = f.spt;
    const nlo_sobol_table* tab = f.tables + (size_t)f.set[item] * (f.rounds + f.reps) + k;
    T* E = reinterpret_cast<T*>(smem);
    size_t o = align_up((size_t)kFsDim * (bins + 1) * sizeof(T), 16);
    uint32_t* dir = reinterpret_cast<uint32_t*>(smem + o);
    o = align_up(o + (size_t)kFsDim * NLO_SOBOL_BITS * sizeof(uint32_t), 16);
    const int pts = 128 * spt;
    double* sw = reinterpret_cast<double*>(smem + o);  // [NK][pts]
    uint64_t* sb = reinterpret_cast<uint64_t*>(smem + align_up(o + 2 * (size_t)pts * sizeof(double), 16));
    for (int i = threadIdx.x; i < kFsDim * (bins + 1); i += blockDim.x) E[i] = (T)grid[i];
    for (int i = threadIdx.x; i < kFsDim * NLO_SOBOL_BITS; i += blockDim.x)
        dir[i] = tab->direction[i / NLO_SOBOL_BITS][i % NLO_SOBOL_BITS];
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = uniform(threadIdx.x >> 6);
    const int l = lane >> 1, sector = lane & 1;
    const uint32_t gl = (uint32_t)(l ^ (l >> 1));
    uint32_t vlo[kFsDim], d4[kFsDim], vhi[kFsDim];
#pragma unroll
    for (int d = 0; d < kFsDim; ++d) {
        uint32_t v0 = 0;
        for (int b = 0; b < kFsGroupBits; ++b)
            if ((gl >> b) & 1u) v0 ^= dir[d * NLO_SOBOL_BITS + b];
        vlo[d] = v0 ^ tab->shift[d];
        d4[d] = (uint32_t)uniform((int)dir[d * NLO_SOBOL_BITS + kFsGroupBits - 1]);
        vhi[d] = 0u;
    }
    const int64_t hi0 = ((int64_t)blk * 4 + wave) * spt;
    {
        uint64_t g = (uint64_t)hi0 ^ ((uint64_t)hi0 >> 1);
        for (int b = kFsGroupBits; g && b < NLO_SOBOL_BITS; ++b, g >>= 1)
            if (g & 1u)
#pragma unroll
                for (int d = 0; d < kFsDim; ++d) vhi[d] ^= (uint32_t)uniform((int)dir[d * NLO_SOBOL_BITS + b]);
    }
    const T maxr = (T)a.maxr;
    double t[NK], e[NK];
#pragma unroll
    for (int q = 0; q < NK; ++q) t[q] = e[q] = 0.0;
    for (int j = 0; j < spt; ++j) {
        const int64_t hi = hi0 + j;
        if (j > 0) {
            const int b = kFsGroupBits + __builtin_ctzll((unsigned long long)hi);
            if (b < NLO_SOBOL_BITS)
#pragma unroll
                for (int d = 0; d < kFsDim; ++d) vhi[d] ^= (uint32_t)uniform((int)dir[d * NLO_SOBOL_BITS + b]);
        }
        const int64_t s = (hi << kFsGroupBits) + l;
        const bool valid = s < n;
        T u[kFsDim];
#pragma unroll
        for (int d = 0; d < kFsDim; ++d) {
            uint32_t q = vlo[d] ^ ((hi & 1) ? (vhi[d] ^ d4[d]) : vhi[d]);
            if (s == 0) q = tab->first[d];
            u[d] = sobol_value<T>(q);
        }
        const EndpointPair<T> ep = endpoint_pair<T>(u, E, bins, maxr, f.fold != 0);
        const double ud[kFsDim] = {(double)ep.src[0], (double)ep.src[1], (double)ep.src[2],
                                   (double)(sector ? ep.s3 : ep.f3), (double)(sector ? ep.s4 : ep.f4)};
        const double w = (double)(sector ? ep.w2 : ep.w1);
        double x[NK];
        if (valid) {
            if constexpr (PART == kFsI1) {
                x[0] = (double)dis_value<double, NLO_DIS_QG_I1>(a, ud, rule) * w;
            } else {
                T i2, i3;
                qg_nested_pair<T>(a, ud, rule, i2, i3);
                x[0] = (double)i2 * w;
                x[1] = (double)i3 * w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < NK; ++q) x[q] = 0.0;
        }
        // (first value x first weight + second value x second weight) x inverse density
        const int idx = (wave * spt + j) * 32 + l;
#pragma unroll
        for (int q = 0; q < NK; ++q) {
            const double other = __shfl_xor(x[q], 1, kWave);
            const double wq = (sector ? other + x[q] : x[q] + other) * (double)ep.inv;
            if (sector == 0 && valid) comp_add(t[q], e[q], TRAIN ? fabs(wq) : wq);
            if (TRAIN && sector == 0) sw[q * pts + idx] = valid ? fabs(wq) : 0.0;
        }
        if (TRAIN && sector == 0) sb[idx] = valid ? ep.bins : ~0ull;
    }
    double* out = part;
    block_comp_sums<NK>(t, e, red, out + (PART == kFsI1 ? 0 : 1));
    if constexpr (TRAIN) {
        __syncthreads();
        const int cells = fs_cells(bins);
        for (int c = threadIdx.x; c < cells; c += blockDim.x) {
            const int d = c / bins, b = c - d * bins;
            double h[NK];
#pragma unroll
            for (int q = 0; q < NK; ++q) h[q] = 0.0;
            double cnt = 0.0;
            for (int i = 0; i < pts; ++i) {
                const uint64_t pk = sb[i];
                if (pk == ~0ull || (int)((pk >> (8 * d)) & 0xff) != b) continue;
#pragma unroll
                for (int q = 0; q < NK; ++q) h[q] += sw[q * pts + i];
                cnt += 1.0;
            }
            if constexpr (PART == kFsI1) {
                out[kFsPart + c] = h[0];
                out[kFsPart + 3 * (int64_t)cells + c] = cnt;
            } else {
                out[kFsPart + (int64_t)cells + c] = h[0];
                out[kFsPart + 2 * (int64_t)cells + c] = h[1];
            }
        }
    }
}

// Occupancy: the FP64 I1 pass at two waves per SIMD (256 VGPRs), the nested
// pass at four in FP32 (128 VGPRs), two in FP64.
template <typename T, int PART>
constexpr int kFsWaves = (PART == kFsNested && std::is_same<T, float>::value) ? 4 : 2;

// Training round `round` of every estimator: block (estimator, block).
template <typename T, int PART>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFsWaves<T, PART>)))
fs_train_kernel(FsDev f, int round) {
    const int item = blockIdx.x / f.blocks_train, blk = blockIdx.x - item * f.blocks_train;
    const double* grid = f.edges + (int64_t)item * 2 * kFsDim * (f.bins + 1) + (int64_t)(round & 1) * kFsDim * (f.bins + 1);
    double* part = f.part + (int64_t)item * f.part_stride + (int64_t)blk * fs_train_part(f.bins);
    fs_pass<T, PART, true>(f, item, round, blk, f.n_train, grid, part);
}

// Every replicate of every estimator: block (estimator, replicate, block).
template <typename T, int PART>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFsWaves<T, PART>)))
fs_rep_kernel(FsDev f) {
    const int per = f.reps * f.blocks_rep;
    const int item = blockIdx.x / per, rem = blockIdx.x - item * per;
    const int r = rem / f.blocks_rep, blk = rem - r * f.blocks_rep;
    const double* grid = f.edges + (int64_t)item * 2 * kFsDim * (f.bins + 1) + (int64_t)(f.rounds & 1) * kFsDim * (f.bins + 1);
    double* part = f.part + (int64_t)item * f.part_stride + (int64_t)rem * kFsPart;
    fs_pass<T, PART, false>(f, item, f.rounds + r, blk, f.n_rep, grid, part);
}

// Starting grid of every estimator (buffer 0): uniform, or the caller's frozen
// grid (n_items x dim x (bins + 1)).
__global__ void __launch_bounds__(256) fs_init_kernel(FsDev f, const double* __restrict__ frozen) {
    const int item = blockIdx.x;
    const int E = kFsDim * (f.bins + 1);
    double* g = f.edges + (int64_t)item * 2 * E;
    for (int i = threadIdx.x; i < E; i += blockDim.x)
        g[i] = frozen ? frozen[(int64_t)item * E + i] : (double)(i % (f.bins + 1)) / (double)f.bins;
}

// Training partials of every estimator summed over its blocks: block
// (estimator, slice of 16 slots), thread (slot, group g < 16) sums blocks
// g, g + 16, ... in order, then the 16 group sums of a slot are merged in
// group order (a fixed order whatever the hardware schedule).
constexpr int kFsReduceGroups = 16;
constexpr int kFsReduceSlots = 256 / kFsReduceGroups;

__global__ void __launch_bounds__(256) fs_train_reduce_kernel(FsDev f, int slices) {
    __shared__ double gs[kFsReduceGroups][kFsReduceSlots], gc[kFsReduceGroups][kFsReduceSlots];
    const int item = blockIdx.x / slices, sl = blockIdx.x - item * slices;
    const int64_t L = fs_train_part(f.bins);
    const int j = threadIdx.x % kFsReduceSlots, g = threadIdx.x / kFsReduceSlots;
    const int64_t k = (int64_t)sl * kFsReduceSlots + j;
    double s = 0.0, c = 0.0;
    if (k < L) {
        const double* p = f.part + (int64_t)item * f.part_stride + k;
        for (int b = g; b < f.blocks_train; b += kFsReduceGroups) comp_add(s, c, p[(int64_t)b * L]);
    }
    gs[g][j] = s;
    gc[g][j] = c;
    __syncthreads();
    if (g == 0 && k < L) {
        for (int q = 1; q < kFsReduceGroups; ++q) comp_merge(s, c, gs[q][j], gc[q][j]);
        f.stat[(int64_t)item * f.stat_stride + k] = s + c;
    }
}

// The score histogram sum_k weight_k |w_k| / mean |w_k| per bin
// (learn_triple_importance_grid, S/dis/fixed_sobol.py:193-208; the per-point
// scores summed per bin equal the per-bin |w_k| sums so weighted, the sums
// being linear) and the refined grid: block (estimator, dimension) forms the
// scores and their smoothed importances (vegas_smoothed) in parallel in LDS,
// then one lane refines the edge row from the current buffer into the other
// one with the same operations as vegas_adapt_row.
__global__ void __launch_bounds__(64) fs_adapt_kernel(FsDev f, int round, double eps, double tiny) {
    __shared__ double sc[256], cn[256], ed[256], sm[256];
    const int item = blockIdx.x / kFsDim, d = blockIdx.x - item * kFsDim;
    const int bins = f.bins, cells = fs_cells(bins);
    double* st = f.stat + (int64_t)item * f.stat_stride;
    const double* H = st + kFsPart;
    const double* cnt = H + 3 * (int64_t)cells;
    const double n = (double)f.n_train;
    const double wt[3] = {f.i1_weight, 1.0, 1.0};
    double inv_mean[3];
    for (int q = 0; q < 3; ++q) inv_mean[q] = 1.0 / fmax(st[q] / n, tiny);
    double* g = f.edges + (int64_t)item * 2 * kFsDim * (bins + 1);
    const double* e = g + (int64_t)(round & 1) * kFsDim * (bins + 1) + (int64_t)d * (bins + 1);
    double* o = g + (int64_t)((round + 1) & 1) * kFsDim * (bins + 1) + (int64_t)d * (bins + 1);
    for (int b = threadIdx.x; b < bins; b += blockDim.x) {
        double v = 0.0;
        for (int q = 0; q < 3; ++q) v += wt[q] * H[(int64_t)q * cells + d * bins + b] * inv_mean[q];
        sc[b] = v;
        cn[b] = cnt[(int64_t)d * bins + b];
        st[fs_train_part(bins) + (int64_t)d * bins + b] = v;
    }
    for (int b = threadIdx.x; b <= bins; b += blockDim.x) ed[b] = e[b];
    __syncthreads();
    for (int b = threadIdx.x; b < bins; b += blockDim.x) sm[b] = vegas_smoothed(sc, cn, bins, b, eps);
    __syncthreads();
    if (threadIdx.x == 0) vegas_adapt_row_with(ed, [&](int i) { return sm[i]; }, bins, o);
}

// Replicate estimates: block (estimator, replicate), one wave: the block sums
// of each contribution in block order, divided by the points per replicate.
__global__ void __launch_bounds__(64) fs_rep_reduce_kernel(FsDev f) {
    const int item = blockIdx.x / f.reps, r = blockIdx.x - item * f.reps;
    const double* p = f.part + (int64_t)item * f.part_stride + (int64_t)r * f.blocks_rep * kFsPart;
    double* est = f.stat + (int64_t)item * f.stat_stride + fs_train_part(f.bins) + fs_cells(f.bins) +
                  (int64_t)r * kFsPart;
    for (int q = 0; q < kFsPart; ++q) {
        double s = 0.0, c = 0.0;
        for (int b = threadIdx.x; b < f.blocks_rep; b += blockDim.x) comp_add(s, c, p[(int64_t)b * kFsPart + q]);
        wave_comp_sum(s, c);
        if (threadIdx.x == 0) est[q] = (s + c) / (double)f.n_rep;
    }
}

// Per estimator: I1, I2, I3 and their per-replicate sum as (mean, standard
// deviation / sqrt(replicates), evaluations, error <= epsrel |mean|)
// (integrate_triple_fixed_sobol, S/dis/fixed_sobol.py:256-271).
__global__ void __launch_bounds__(64) fs_stats_kernel(FsDev f, int n_items, double epsrel, double* __restrict__ out) {
    const int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= n_items) return;
    const double* est = f.stat + (int64_t)item * f.stat_stride + fs_train_part(f.bins) + fs_cells(f.bins);
    const double evals = 2.0 * (double)f.rounds * (double)f.n_train + 2.0 * (double)f.reps * (double)f.n_rep;
    const int R = f.reps;
    for (int row = 0; row < 4; ++row) {
        double mean = 0.0;
        for (int r = 0; r < R; ++r) {
            const double* v = est + (int64_t)r * kFsPart;
            mean += row < 3 ? v[row] : (v[0] + v[1]) + v[2];
        }
        mean /= (double)R;
        double ss = 0.0;
        for (int r = 0; r < R; ++r) {
            const double* v = est + (int64_t)r * kFsPart;
            const double dv = (row < 3 ? v[row] : (v[0] + v[1]) + v[2]) - mean;
            ss += dv * dv;
        }
        const double error = sqrt(ss / (double)(R - 1)) / sqrt((double)R);
        double* o = out + ((int64_t)item * 4 + row) * 4;
        o[0] = mean;
        o[1] = error;
        o[2] = evals;
        o[3] = error <= epsrel * fabs(mean) ? 1.0 : 0.0;
    }
}

// The learned grids (the frozen buffer) into a dense estimators x dim x
// (bins + 1) array.
__global__ void __launch_bounds__(256) fs_edges_out_kernel(FsDev f, double* __restrict__ out) {
    const int item = blockIdx.x;
    const int E = kFsDim * (f.bins + 1);
    const double* g = f.edges + (int64_t)item * 2 * E + (int64_t)(f.rounds & 1) * E;
    for (int i = threadIdx.x; i < E; i += blockDim.x) out[(int64_t)item * E + i] = g[i];
}

// The standalone endpoint map (one point per thread).
template <typename T>
__global__ void __launch_bounds__(256) endpoint_map_kernel(int64_t n, const T* __restrict__ points,
                                                           const T* __restrict__ edges, int bins, T maxr, int fold,
                                                           T* __restrict__ first, T* __restrict__ w1,
                                                           T* __restrict__ second, T* __restrict__ w2) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* E = reinterpret_cast<T*>(smem);
    for (int i = threadIdx.x; i < kFsDim * (bins + 1); i += blockDim.x) E[i] = edges[i];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T u[kFsDim];
#pragma unroll
    for (int d = 0; d < kFsDim; ++d) u[d] = points[i * kFsDim + d];
    const EndpointPair<T> ep = endpoint_pair<T>(u, E, bins, maxr, fold != 0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        first[i * kFsDim + d] = ep.src[d];
        second[i * kFsDim + d] = ep.src[d];
    }
    first[i * kFsDim + 3] = ep.f3;
    first[i * kFsDim + 4] = ep.f4;
    second[i * kFsDim + 3] = ep.s3;
    second[i * kFsDim + 4] = ep.s4;
    w1[i] = ep.inv * ep.w1;
    w2[i] = ep.inv * ep.w2;
}

// ------------------------------------------------------------------- host

struct FsPlan {
    std::vector<DisArgs> args;
    std::vector<int> set;
    int n_sets = 0, spt = 1, blocks_train = 0, blocks_rep = 0;
    int64_t part_stride = 0, stat_stride = 0;
    size_t o_args = 0, o_set = 0, o_tables = 0, o_edges = 0, o_part = 0, o_stat = 0, total = 0;
    size_t lds_train = 0, lds_rep = 0;
};

bool power_of_two(int64_t v) { return v >= 2 && (v & (v - 1)) == 0; }

int fs_plan(const nlo_dis_fixed_sobol_desc* d, FsPlan& P, bool build_args) {
    NLO_REQUIRE(d != nullptr, "nlo_dis_fixed_sobol: desc is NULL");
    NLO_REQUIRE(d->n_items >= 1 && d->items, "nlo_dis_fixed_sobol: no items");
    NLO_REQUIRE(d->n_tables >= 1 && d->tables, "nlo_dis_fixed_sobol: no Sobol tables");
    NLO_REQUIRE(d->d_grid ? d->training_rounds == 0
                          : (d->training_rounds >= 1 && d->training_rounds <= NLO_FIXED_SOBOL_MAX_ROUNDS),
                "nlo_dis_fixed_sobol: training_rounds must be 1..%d (0 with frozen grids)",
                NLO_FIXED_SOBOL_MAX_ROUNDS);
    NLO_REQUIRE(d->replicates >= 2 && d->replicates <= NLO_FIXED_SOBOL_MAX_REPLICATES,
                "nlo_dis_fixed_sobol: replicates must be 2..%d", NLO_FIXED_SOBOL_MAX_REPLICATES);
    NLO_REQUIRE(power_of_two(d->training_samples) && power_of_two(d->samples_per_replicate),
                "nlo_dis_fixed_sobol: Sobol sample counts must be powers of two greater than one");
    NLO_REQUIRE(d->training_samples <= (1 << 24) && d->samples_per_replicate <= (1 << 24),
                "nlo_dis_fixed_sobol: at most 2^24 points per set");
    NLO_REQUIRE(d->bins >= 2 && d->bins <= 255, "nlo_dis_fixed_sobol: bins must be 2..255");
    NLO_REQUIRE(d->precision == NLO_PREC_FP32 || d->precision == NLO_PREC_FP64,
                "nlo_dis_fixed_sobol: precision must be FP32 or FP64");
    NLO_REQUIRE(d->epsrel >= 0.0 && std::isfinite(d->i1_weight), "nlo_dis_fixed_sobol: bad epsrel or i1_weight");
    const int per_set = d->training_rounds + d->replicates;
    NLO_REQUIRE(d->n_tables % per_set == 0,
                "nlo_dis_fixed_sobol: tables must come in sets of training_rounds + replicates (%d)", per_set);
    P.n_sets = d->n_tables / per_set;
    for (int t = 0; t < d->n_tables; ++t)
        NLO_REQUIRE(d->tables[t].dim == kFsDim, "nlo_dis_fixed_sobol: table %d must be five-dimensional", t);
    for (int i = 0; i < d->n_items; ++i) {
        const nlo_dis_fixed_sobol_item& it = d->items[i];
        NLO_REQUIRE(it.table_set >= 0 && it.table_set < P.n_sets,
                    "nlo_dis_fixed_sobol: item %d: table set %d out of range", i, it.table_set);
        NLO_REQUIRE(it.desc.dipole.kind == NLO_DIPOLE_GBW, "nlo_dis_fixed_sobol: item %d: GBW dipoles only", i);
        P.set.push_back(it.table_set);
        if (build_args) {
            DisArgs a;
            const int rc = make_args(&it.desc, a);
            if (rc != NLO_OK) return rc;
            P.args.push_back(a);
        }
    }
    // 128 points per block (one 32-point group per wave), whatever the batch:
    // an estimator's sums do not depend on what else runs with it
    P.spt = 1;
    P.blocks_train = (int)((d->training_samples + 128 * P.spt - 1) / (128 * P.spt));
    P.blocks_rep = (int)((d->samples_per_replicate + 128 * P.spt - 1) / (128 * P.spt));
    P.part_stride = std::max<int64_t>((int64_t)P.blocks_train * fs_train_part(d->bins),
                                      (int64_t)d->replicates * P.blocks_rep * kFsPart);
    P.stat_stride = fs_stat_stride(d->bins, d->replicates);
    const size_t lds_t = d->precision == NLO_PREC_FP32 ? fs_lds<float>(d->bins, P.spt, true)
                                                       : fs_lds<double>(d->bins, P.spt, true);
    const size_t lds_r = d->precision == NLO_PREC_FP32 ? fs_lds<float>(d->bins, P.spt, false)
                                                       : fs_lds<double>(d->bins, P.spt, false);
    P.lds_train = lds_t;
    P.lds_rep = lds_r;
    size_t o = 0;
    auto section = [&](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes, 256);
        return at;
    };
    P.o_args = section(sizeof(DisArgs) * (size_t)d->n_items);
    P.o_set = section(sizeof(int) * (size_t)d->n_items);
    P.o_tables = section(sizeof(nlo_sobol_table) * (size_t)d->n_tables);
    P.o_edges = section(sizeof(double) * (size_t)d->n_items * 2 * kFsDim * (d->bins + 1));
    P.o_part = section(sizeof(double) * (size_t)d->n_items * (size_t)P.part_stride);
    P.o_stat = section(sizeof(double) * (size_t)d->n_items * (size_t)P.stat_stride);
    P.total = o;
    return NLO_OK;
}

template <typename T>
int fs_enqueue(const nlo_dis_fixed_sobol_desc* d, const FsPlan& P, const FsDev& f, hipStream_t s) {
    const int n = d->n_items;
    const double eps = std::is_same<T, float>::value ? (double)FLT_EPSILON : DBL_EPSILON;
    const double tiny = std::is_same<T, float>::value ? (double)FLT_MIN : DBL_MIN;
    hipLaunchKernelGGL(fs_init_kernel, dim3(n), dim3(256), 0, s, f, d->d_grid);
    NLO_LAUNCH_CHECK();
    const int slices = (int)((fs_train_part(d->bins) + kFsReduceSlots - 1) / kFsReduceSlots);
    for (int round = 0; round < d->training_rounds; ++round) {
        const dim3 grid((unsigned)((int64_t)n * P.blocks_train));
        hipLaunchKernelGGL((fs_train_kernel<T, kFsI1>), grid, dim3(256), P.lds_train, s, f, round);
        NLO_LAUNCH_CHECK();
        hipLaunchKernelGGL((fs_train_kernel<T, kFsNested>), grid, dim3(256), P.lds_train, s, f, round);
        NLO_LAUNCH_CHECK();
        hipLaunchKernelGGL(fs_train_reduce_kernel, dim3((unsigned)(n * slices)), dim3(256), 0, s, f, slices);
        NLO_LAUNCH_CHECK();
        hipLaunchKernelGGL(fs_adapt_kernel, dim3(n * kFsDim), dim3(64), 0, s, f, round, eps, tiny);
        NLO_LAUNCH_CHECK();
    }
    const dim3 grid((unsigned)((int64_t)n * d->replicates * P.blocks_rep));
    hipLaunchKernelGGL((fs_rep_kernel<T, kFsI1>), grid, dim3(256), P.lds_rep, s, f);
    NLO_LAUNCH_CHECK();
    hipLaunchKernelGGL((fs_rep_kernel<T, kFsNested>), grid, dim3(256), P.lds_rep, s, f);
    NLO_LAUNCH_CHECK();
    hipLaunchKernelGGL(fs_rep_reduce_kernel, dim3((unsigned)(n * d->replicates)), dim3(64), 0, s, f);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

}  // namespace
}  // namespace dis
}  // namespace nlo

using namespace nlo;
using namespace nlo::dis;

extern "C" {

size_t nlo_dis_fixed_sobol_workspace(const nlo_dis_fixed_sobol_desc* desc) {
    clear_error();
    FsPlan P;
    if (fs_plan(desc, P, false) != NLO_OK) return 0;
    return P.total;
}

int nlo_dis_fixed_sobol(const nlo_dis_fixed_sobol_desc* desc, double* d_results, double* d_edges, void* d_workspace,
                        size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    FsPlan P;
    int rc = fs_plan(desc, P, true);
    if (rc != NLO_OK) return rc;
    NLO_REQUIRE(d_results != nullptr, "nlo_dis_fixed_sobol: NULL results");
    if (!d_workspace || workspace_bytes < P.total) {
        set_error("nlo_dis_fixed_sobol: workspace %zu bytes < %zu", workspace_bytes, P.total);
        return NLO_ERR_WORKSPACE;
    }
    NLO_REQUIRE(P.lds_train <= 64 * 1024, "nlo_dis_fixed_sobol: LDS %zu bytes exceeds 64 KiB", P.lds_train);
    unsigned char* w = static_cast<unsigned char*>(d_workspace);
    std::vector<unsigned char> host(P.o_edges, 0);
    std::memcpy(host.data() + P.o_args, P.args.data(), sizeof(DisArgs) * P.args.size());
    std::memcpy(host.data() + P.o_set, P.set.data(), sizeof(int) * P.set.size());
    std::memcpy(host.data() + P.o_tables, desc->tables, sizeof(nlo_sobol_table) * (size_t)desc->n_tables);
    NLO_HIP_CHECK(hipMemcpyAsync(w, host.data(), host.size(), hipMemcpyHostToDevice, stream));
    FsDev f{};
    f.args = reinterpret_cast<const DisArgs*>(w + P.o_args);
    f.set = reinterpret_cast<const int*>(w + P.o_set);
    f.tables = reinterpret_cast<const nlo_sobol_table*>(w + P.o_tables);
    f.edges = reinterpret_cast<double*>(w + P.o_edges);
    f.part = reinterpret_cast<double*>(w + P.o_part);
    f.stat = reinterpret_cast<double*>(w + P.o_stat);
    f.bins = desc->bins;
    f.rounds = desc->training_rounds;
    f.reps = desc->replicates;
    f.fold = desc->fold_angle != 0;
    f.spt = P.spt;
    f.blocks_train = P.blocks_train;
    f.blocks_rep = P.blocks_rep;
    f.n_train = desc->training_samples;
    f.n_rep = desc->samples_per_replicate;
    f.part_stride = P.part_stride;
    f.stat_stride = P.stat_stride;
    f.i1_weight = desc->i1_weight;
    int status = desc->precision == NLO_PREC_FP32 ? fs_enqueue<float>(desc, P, f, stream)
                                                  : fs_enqueue<double>(desc, P, f, stream);
    if (status == NLO_OK) {
        hipLaunchKernelGGL(fs_stats_kernel, dim3((desc->n_items + 63) / 64), dim3(64), 0, stream, f, desc->n_items,
                           desc->epsrel, d_results);
        if (d_edges) hipLaunchKernelGGL(fs_edges_out_kernel, dim3(desc->n_items), dim3(256), 0, stream, f, d_edges);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_error("nlo_dis_fixed_sobol: results: %s", hipGetErrorString(e));
            status = NLO_ERR_HIP;
        }
    }
    // the staging buffer is released once the stream is done
    const hipError_t es = hipStreamSynchronize(stream);
    if (status == NLO_OK && es != hipSuccess) {
        set_error("nlo_dis_fixed_sobol: %s", hipGetErrorString(es));
        status = NLO_ERR_HIP;
    }
    return status;
}

int nlo_dis_endpoint_map(int32_t precision, int64_t n, const void* d_points, const void* d_edges, int32_t bins,
                         double maxr, int32_t fold_angle, void* d_first, void* d_first_weight, void* d_second,
                         void* d_second_weight, hipStream_t stream) {
    clear_error();
    NLO_REQUIRE(precision == NLO_PREC_FP32 || precision == NLO_PREC_FP64,
                "nlo_dis_endpoint_map: precision must be FP32 or FP64");
    NLO_REQUIRE(n >= 0, "nlo_dis_endpoint_map: negative point count");
    NLO_REQUIRE(bins >= 2 && bins <= 255, "nlo_dis_endpoint_map: bins must be 2..255");
    NLO_REQUIRE(maxr > 0.0 && std::isfinite(maxr), "nlo_dis_endpoint_map: maxr must be positive and finite");
    if (n == 0) return NLO_OK;
    NLO_REQUIRE(d_points && d_edges && d_first && d_first_weight && d_second && d_second_weight,
                "nlo_dis_endpoint_map: NULL array");
    const unsigned blocks = (unsigned)((n + 255) / 256);
    NLO_REQUIRE(n <= (int64_t)256 * 0x7fffffff, "nlo_dis_endpoint_map: too many points");
    if (precision == NLO_PREC_FP32) {
        const size_t lds = (size_t)kFsDim * (bins + 1) * sizeof(float);
        hipLaunchKernelGGL(endpoint_map_kernel<float>, dim3(blocks), dim3(256), lds, stream, n,
                           static_cast<const float*>(d_points), static_cast<const float*>(d_edges), (int)bins,
                           (float)maxr, (int)fold_angle, static_cast<float*>(d_first),
                           static_cast<float*>(d_first_weight), static_cast<float*>(d_second),
                           static_cast<float*>(d_second_weight));
    } else {
        const size_t lds = (size_t)kFsDim * (bins + 1) * sizeof(double);
        hipLaunchKernelGGL(endpoint_map_kernel<double>, dim3(blocks), dim3(256), lds, stream, n,
                           static_cast<const double*>(d_points), static_cast<const double*>(d_edges), (int)bins,
                           maxr, (int)fold_angle, static_cast<double*>(d_first),
                           static_cast<double*>(d_first_weight), static_cast<double*>(d_second),
                           static_cast<double*>(d_second_weight));
    }
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

}  // extern "C"
