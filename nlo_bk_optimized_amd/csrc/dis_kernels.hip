// NLO DIS kernels and their C ABI: special functions, dipole amplitudes, the
// BKDipole initialisation, per-point integrands, one VEGAS pass with in-kernel
// scrambled Sobol points and deterministic FP64 sums, and the 2-D tensor
// Gauss-Legendre rule.  Reference prefixes: S/ = torch_nlo_clean/src/nlo_torch/,
// K/ = S/custom_kernels/.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
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

// ------------------------------------------------------------------ kernels

template <typename T>
__global__ void __launch_bounds__(256) special_kernel(int fn, int64_t n, const double* __restrict__ x,
                                                      double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T v = (T)x[i];
    T r;
    if (fn == NLO_FN_K0) r = bessel_k0<T>(v);
    else if (fn == NLO_FN_K1) r = bessel_k1<T>(v);
    else if (fn == NLO_FN_K2) r = bessel_k2<T>(v);
    else r = dilog<T>(v);
    out[i] = (double)r;
}

__global__ void __launch_bounds__(256) dipole_kernel(DisDipole d, int64_t n, const double* __restrict__ r,
                                                     const double* __restrict__ Y, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = dipole_amplitude<double>(r[i], Y[i], d);
}

// One block (64 lanes): N at the (moved) table nodes through the local-spline
// path, a natural spline in r through them, then the maxr_interpolate scan of
// BKDipole._find_maxr_interpolate on that spline (lane 0).
__global__ void __launch_bounds__(64) bk_init_kernel(DisDipole d, double Y, double* __restrict__ coef,
                                                     double* __restrict__ maxr_out) {
    extern __shared__ double sh[];
    const int R = d.R;
    double* yv = sh;          // R
    double* m = sh + R;       // R second derivatives
    double* cp = sh + 2 * R;  // R
    d.initialized = 0;
    for (int i = threadIdx.x; i < R; i += blockDim.x) {
        double rr = d.r[i];
        if (i == 0) rr *= 1.0001;
        if (i == R - 1) rr *= 0.9999;
        yv[i] = bk_amplitude(rr, Y, d);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // natural spline through (r_i, yv_i): Thomas on the interior second derivatives
        const int ni = R - 2;
        double dprev = 0.0;
        for (int i = 0; i < ni; ++i) {
            const double h0 = d.r[i + 1] - d.r[i], h1 = d.r[i + 2] - d.r[i + 1];
            const double rhs = 6.0 * ((yv[i + 2] - yv[i + 1]) / h1 - (yv[i + 1] - yv[i]) / h0);
            const double lower = i > 0 ? h0 : 0.0;
            const double den = 2.0 * (h0 + h1) - lower * (i > 0 ? cp[i - 1] : 0.0);
            cp[i] = h1 / den;
            dprev = (rhs - lower * dprev) / den;
            m[i + 1] = dprev;  // forward value, back-substituted below
        }
        m[0] = 0.0;
        m[R - 1] = 0.0;
        for (int i = ni - 2; i >= 0; --i) m[i + 1] = m[i + 1] - cp[i] * m[i + 2];
        for (int k = 0; k < R - 1; ++k) {
            const double h = d.r[k + 1] - d.r[k];
            coef[4 * k + 0] = yv[k];
            coef[4 * k + 1] = (yv[k + 1] - yv[k]) / h - h * (2.0 * m[k] + m[k + 1]) / 6.0;
            coef[4 * k + 2] = 0.5 * m[k];
            coef[4 * k + 3] = (m[k + 1] - m[k]) / (6.0 * h);
        }
        __threadfence_block();
        // _find_maxr_interpolate (S/dipole/amplitude.py:242-260)
        double step = 2.0, prev = 0.1, rr = 0.01, maxr = -1.0;
        const double rmin = d.r[0], rmax = d.r[R - 1];
        for (int it = 0; it < 41; ++it) {
            if (rr > rmax) break;
            double re = rr < rmin ? rmin * 1.000001 : (rr > rmax ? rmax * 0.999999 : rr);
            int k = search_right(d.r, R, re);
            k = k < 0 ? 0 : (k > R - 2 ? R - 2 : k);
            const double t = re - d.r[k];
            double v = coef[4 * k] + t * (coef[4 * k + 1] + t * (coef[4 * k + 2] + t * coef[4 * k + 3]));
            v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
            if (v >= 0.99999) {
                if (step < 1e-2) {
                    maxr = rr;
                    break;
                }
                step /= 1.5;
                rr = prev;
            }
            prev = rr;
            rr += step;
        }
        *maxr_out = maxr;
    }
}

// The local-spline table: thread (y, i) for row y and r interval i.
__global__ void __launch_bounds__(256) bk_local_kernel(DisDipole d, double* __restrict__ coef) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.T * (d.R - 1)) return;
    const int y = k / (d.R - 1), i = k - y * (d.R - 1);
    int start, pts;
    local_window(i, d.R, start, pts);
    double xs[9], ys[9];
    for (int j = 0; j < pts; ++j) {
        xs[j] = d.r[start + j];
        ys[j] = d.N[(size_t)y * d.R + start + j];
    }
    double c[4];
    local_natural_piece(xs, ys, pts, i - start, c);
    for (int j = 0; j < 4; ++j) coef[4 * (size_t)k + j] = c[j];
}

template <typename T, int K>
__global__ void __launch_bounds__(256) integrand_kernel(DisArgs a, int64_t n, const double* __restrict__ x,
                                                        double* __restrict__ out) {
    constexpr int D = kDims[K];
    __shared__ double ry[kMaxNested], rw[kMaxNested];
    RuleView rule{0, ry, rw};
    if constexpr (K == NLO_DIS_QG_I2_NESTED || K == NLO_DIS_QG_I3_NESTED) rule = stage_nested_rule(a.nested_points, ry, rw);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = x[i * D + d];
    out[i] = (double)dis_value<T, K>(a, u, rule);
}

// The nested I2 and I3 at each point from one pass over the inner rule.
template <typename T>
__global__ void __launch_bounds__(256) i23_kernel(DisArgs a, int64_t n, const double* __restrict__ x,
                                                  double* __restrict__ i2, double* __restrict__ i3) {
    __shared__ double ry[kMaxNested], rw[kMaxNested];
    const RuleView rule = stage_nested_rule(a.nested_points, ry, rw);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[5];
#pragma unroll
    for (int d = 0; d < 5; ++d) u[d] = x[i * 5 + d];
    T v2, v3;
    qg_nested_pair<T>(a, u, rule, v2, v3);
    i2[i] = (double)v2;
    i3[i] = (double)v3;
}

struct VegasArgs {
    int bins, spt, blocks;
    int64_t n, first;
    const double* edges;  // D x (bins + 1)
    double* part;         // blocks x 2
    double* part_hist;    // blocks x D x bins
    double* part_cnt;
    uint32_t dir[NLO_SOBOL_MAXDIM][NLO_SOBOL_BITS];
    uint32_t shift[NLO_SOBOL_MAXDIM];
    uint32_t first_q[NLO_SOBOL_MAXDIM];
};

// LDS of the VEGAS kernel: edges, directions, and (with the histogram) the
// staged |w| and packed bins of the block's samples plus the fixed-point sums.
struct VegasLayout {
    size_t edges, dir, absw, pbin, hist, cnt, total;
};

__host__ __device__ inline VegasLayout vegas_layout(int D, int bins, int spt, bool hist) {
    VegasLayout L{};
    size_t o = 0;
    L.edges = o;
    o = align_up(o + (size_t)D * (bins + 1) * sizeof(double), 16);
    L.dir = o;
    o = align_up(o + (size_t)D * NLO_SOBOL_BITS * sizeof(uint32_t), 16);
    if (hist) {
        L.absw = o;
        o = align_up(o + (size_t)256 * spt * sizeof(double), 16);
        L.pbin = o;
        o = align_up(o + (size_t)256 * spt * sizeof(uint64_t), 16);
        L.hist = o;
        o = align_up(o + (size_t)D * bins * sizeof(unsigned long long), 16);
        L.cnt = o;
        o = align_up(o + (size_t)D * bins * sizeof(unsigned int), 16);
    }
    L.total = o;
    return L;
}

// The sampling side of one VEGAS pass: grid, Sobol scramble and the block's
// partial outputs.  The Sobol directions, shift and integer first point come
// from a table source (the kernel arguments for a single pass, the batch's
// tables in device memory for a batch), read by value.
struct VegasView {
    int bins, spt;
    int64_t n, first;
    const double* edges;  // D x (bins + 1)
    double* part;         // blocks x 2
    double* part_hist;    // blocks x D x bins
    double* part_cnt;
};

struct ArgsSobol {  // a single pass: the table in the kernel arguments
    const VegasArgs& v;
    __device__ uint32_t dir(int d, int b) const { return v.dir[d][b]; }
    __device__ uint32_t shift(int d) const { return v.shift[d]; }
    __device__ uint32_t first(int d) const { return v.first_q[d]; }
};

struct TableSobol {  // a batch item: one of the batch's tables in device memory
    const nlo_sobol_table* t;
    __device__ uint32_t dir(int d, int b) const { return t->direction[d][b]; }
    __device__ uint32_t shift(int d) const { return t->shift[d]; }
    __device__ uint32_t first(int d) const { return t->first[d]; }
};

// One block of a VEGAS pass.  Sobol point index = first + s with
// s = ((blk * 4 + wave) * spt + j) * 64 + lane: each wave walks spt
// consecutive 64-point groups; the high part of the gray code (direction
// bits 6..29) is wave-uniform and updated by one XOR per group.
template <typename T, int K, bool HIST, typename Src>
__device__ __forceinline__ void vegas_pass(const DisArgs& a, const VegasView& v, const Src& src, int blk) {
    constexpr int D = kDims[K];
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[4][2][2];  // [wave][sum][value, error]
    __shared__ double wmax[4];
    __shared__ double ry[kMaxNested], rw[kMaxNested];
    RuleView rule{0, ry, rw};
    if constexpr (K == NLO_DIS_QG_I2_NESTED || K == NLO_DIS_QG_I3_NESTED) rule = stage_nested_rule(a.nested_points, ry, rw);
    const VegasLayout lay = vegas_layout(D, v.bins, v.spt, HIST);
    double* E = reinterpret_cast<double*>(smem + lay.edges);
    uint32_t* dir = reinterpret_cast<uint32_t*>(smem + lay.dir);
    for (int i = threadIdx.x; i < D * (v.bins + 1); i += blockDim.x) E[i] = v.edges[i];
    for (int i = threadIdx.x; i < D * NLO_SOBOL_BITS; i += blockDim.x) dir[i] = src.dir(i / NLO_SOBOL_BITS, i % NLO_SOBOL_BITS);
    unsigned long long* hh = nullptr;
    unsigned int* hc = nullptr;
    double* st_w = nullptr;
    uint64_t* st_b = nullptr;
    if (HIST) {
        hh = reinterpret_cast<unsigned long long*>(smem + lay.hist);
        hc = reinterpret_cast<unsigned int*>(smem + lay.cnt);
        st_w = reinterpret_cast<double*>(smem + lay.absw);
        st_b = reinterpret_cast<uint64_t*>(smem + lay.pbin);
        for (int k = threadIdx.x; k < D * v.bins; k += blockDim.x) {
            hh[k] = 0ull;
            hc[k] = 0u;
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const uint32_t glo0 = (uint32_t)(lane ^ (lane >> 1));
    // per-lane low part of the state (direction bits 0..5 of the lane's gray
    // code) and direction 5, which odd groups add: the wave-uniform parity
    // selects between two uniform values, so no per-lane indexing is needed
    uint32_t vlo[D], d5[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        uint32_t v0 = 0;
        for (int b = 0; b < 6; ++b)
            if ((glo0 >> b) & 1u) v0 ^= dir[d * NLO_SOBOL_BITS + b];
        vlo[d] = v0 ^ src.shift(d);
        d5[d] = (uint32_t)uniform((int)dir[d * NLO_SOBOL_BITS + 5]);
    }
    const int64_t hi0 = (v.first >> 6) + ((int64_t)blk * 4 + wave) * v.spt;
    uint32_t vhi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) vhi[d] = 0u;
    {
        uint64_t g = (uint64_t)hi0 ^ ((uint64_t)hi0 >> 1);
        for (int b = 6; g && b < NLO_SOBOL_BITS; ++b, g >>= 1)
            if (g & 1u)
#pragma unroll
                for (int d = 0; d < D; ++d) vhi[d] ^= (uint32_t)uniform((int)dir[d * NLO_SOBOL_BITS + b]);
    }
    double s1 = 0.0, s2 = 0.0, c1 = 0.0, c2 = 0.0;  // compensated sums of w and w^2
    for (int j = 0; j < v.spt; ++j) {
        const int64_t hi = hi0 + j;
        if (j > 0) {
            const int b = 6 + __builtin_ctzll((unsigned long long)hi);
            if (b < NLO_SOBOL_BITS)
#pragma unroll
                for (int d = 0; d < D; ++d) vhi[d] ^= (uint32_t)uniform((int)dir[d * NLO_SOBOL_BITS + b]);
        }
        const int64_t sample = (hi << 6) + lane - v.first;
        const bool valid = sample < v.n;
        const int par = (int)(hi & 1);
        double u[D];
        double jac = 1.0;
        uint64_t packed = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            uint32_t q = vlo[d] ^ (par ? (vhi[d] ^ d5[d]) : vhi[d]);
            if (hi == 0 && lane == 0) q = src.first(d);
            // bin = min(floor(u B), B - 1), frac = u B - bin, u = q 2^-30
            // (map_importance_grid, S/dis/fixed_sobol.py:100-113)
            const uint32_t q4 = q << 2;
            const bool top = q >= (1u << 30);
            const uint32_t b = top ? (uint32_t)(v.bins - 1) : __umulhi(q4, (uint32_t)v.bins);
            const double frac = top ? 1.0 : (double)(q4 * (uint32_t)v.bins) * 0x1p-32;
            const double left = E[d * (v.bins + 1) + b], right = E[d * (v.bins + 1) + b + 1];
            const double w = right - left;
            u[d] = left + w * frac;
            jac = jac * ((double)v.bins * w);
            packed |= (uint64_t)b << (8 * d);
        }
        double wv = 0.0;
        if (valid) {
            wv = (double)dis_value<T, K>(a, u, rule) * jac;
            comp_add(s1, c1, wv);
            comp_add(s2, c2, wv * wv);
        }
        if (HIST) {
            st_w[threadIdx.x + j * 256] = valid ? fabs(wv) : 0.0;
            st_b[threadIdx.x + j * 256] = valid ? packed : ~0ull;
        }
    }
    {
        double t[2] = {s1, s2}, e[2] = {c1, c2};
        block_comp_sums<2>(t, e, red, v.part + (size_t)blk * 2);
    }
    if (HIST) {
        // block max of |w| sets a fixed-point scale; the integer LDS sums are
        // exact and independent of the order of the atomics
        double mx = 0.0;
        for (int j = 0; j < v.spt; ++j) mx = fmax(mx, st_w[threadIdx.x + j * 256]);
        mx = wave_max(mx);
        if (lane == 0) wmax[wave] = mx;
        __syncthreads();
        mx = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
        const int e = mx > 0.0 ? ilogb(mx) : 0;
        const int sc = 50 - e;  // each term < 2^51; <= 2^12 terms per bin stay below 2^63
        for (int j = 0; j < v.spt; ++j) {
            const uint64_t pk = st_b[threadIdx.x + j * 256];
            if (pk == ~0ull) continue;
            const unsigned long long q = (unsigned long long)rint(ldexp(st_w[threadIdx.x + j * 256], sc));
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int b = (int)((pk >> (8 * d)) & 0xff);
                atomicAdd(&hh[d * v.bins + b], q);
                atomicAdd(&hc[d * v.bins + b], 1u);
            }
        }
        __syncthreads();
        for (int k = threadIdx.x; k < D * v.bins; k += blockDim.x) {
            v.part_hist[(size_t)blk * D * v.bins + k] = ldexp((double)hh[k], -sc);
            v.part_cnt[(size_t)blk * D * v.bins + k] = (double)hc[k];
        }
    }
}

// One VEGAS pass (see nlo_dis_vegas).
template <typename T, int K, bool HIST>
__global__ void __launch_bounds__(256) vegas_kernel(DisArgs a, VegasArgs v) {
    const VegasView view{v.bins, v.spt, v.n, v.first, v.edges, v.part, v.part_hist, v.part_cnt};
    vegas_pass<T, K, HIST>(a, view, ArgsSobol{v}, blockIdx.x);
}

// Histogram partials, first level: block (x, g) sums the per-block cells
// [64 x, 64 x + 64) over the g-th contiguous range of blocks, in block order
// (coalesced across lanes: consecutive cells).
constexpr int kHistGroups = 64;

__global__ void __launch_bounds__(64) vegas_hist_groups_kernel(int blocks, int cells, int per_group,
                                                              const double* __restrict__ ph,
                                                              const double* __restrict__ pc,
                                                              double* __restrict__ gh, double* __restrict__ gc) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int g = blockIdx.y;
    if (k >= cells) return;
    const int b0 = g * per_group, b1 = min(b0 + per_group, blocks);
    double h = 0.0, c = 0.0;
    for (int b = b0; b < b1; ++b) {
        h += ph[(size_t)b * cells + k];
        c += pc[(size_t)b * cells + k];
    }
    gh[(size_t)g * cells + k] = h;
    gc[(size_t)g * cells + k] = c;
}

// Fixed-order sums over the per-block partials: sums (2) by strided thread
// sums, a lane butterfly and a wave sum in block 0; each histogram cell by
// one thread over the group partials in group order.
__global__ void __launch_bounds__(256) vegas_reduce_kernel(int blocks, int cells, int groups,
                                                           const double* __restrict__ part,
                                                           const double* __restrict__ gh,
                                                           const double* __restrict__ gc, double* __restrict__ sums,
                                                           double* __restrict__ hist, double* __restrict__ cnt) {
    __shared__ double red[4][2][2];  // [wave][sum][value, error]
    if (blockIdx.x == 0) {
        double t[2] = {0.0, 0.0}, e[2] = {0.0, 0.0};
        for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
            comp_add(t[0], e[0], part[2 * (size_t)b]);
            comp_add(t[1], e[1], part[2 * (size_t)b + 1]);
        }
        block_comp_sums<2>(t, e, red, sums);
        return;
    }
    if (hist == nullptr) return;
    const int k = (blockIdx.x - 1) * blockDim.x + threadIdx.x;
    if (k >= cells) return;
    double h = 0.0, c = 0.0;
    for (int g = 0; g < groups; ++g) {
        h += gh[(size_t)g * cells + k];
        c += gc[(size_t)g * cells + k];
    }
    hist[k] = h;
    cnt[k] = c;
}

// Tensor Gauss-Legendre over the unit square, `points` nodes per axis in
// block 0 and max(2, points / 2) in block 1, nodes computed into LDS; FP64
// sums in a fixed order.
template <typename T, int K>
__device__ __forceinline__ void gl_pass(const DisArgs& a, int points, int blk, double* __restrict__ out) {
    __shared__ double red[4][1][2];
    __shared__ double gx[128], gw[128];
    const int n = blk == 0 ? points : (points / 2 > 2 ? points / 2 : 2);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double x, w;
        gl_node(n, i, x, w);
        gx[i] = 0.5 * (x + 1.0);
        gw[i] = 0.5 * w;
    }
    __syncthreads();
    const RuleView none{0, gx, gw};
    double s[1] = {0.0}, e[1] = {0.0};  // compensated
    for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
        const int i = idx / n, j = idx - i * n;
        const double u[2] = {gx[i], gx[j]};
        comp_add(s[0], e[0], gw[i] * gw[j] * (double)dis_value<T, K>(a, u, none));
    }
    block_comp_sums<1>(s, e, red, out + blk);
}

template <typename T, int K>
__global__ void __launch_bounds__(256) gl_kernel(DisArgs a, int points, double* __restrict__ out) {
    gl_pass<T, K>(a, points, blockIdx.x, out);
}

// ----------------------------------------------------------------- batches
//
// A batch integrates many DIS contributions at once (points x polarisations x
// quarks x terms): every VEGAS integral of the batch runs its own grid, Sobol
// table and iteration schedule, but each pass over all of them is a handful
// of launches (one per contribution kind), and one iteration (passes,
// histogram reductions, grid refinement, statistics and the convergence test)
// is one captured graph replayed without host synchronisation.  Converged
// integrals skip the remaining iterations on the device.

constexpr int kBatchMaxIterations = 16;
constexpr int kBatchSpt = 8;  // 64-point groups per wave: 2048 samples per block

// One VEGAS integral of a batch: its item, plan and where its data live.
struct VItem {
    int item, K, D, table;
    int blocks, groups, per_group;
    int max_it, min_it;
    int64_t n, stride, first;
    int64_t part;   // doubles into the partial buffer: blocks x 2 sums, then the
                    // per-block histogram and counts (2 x blocks x cells), then
                    // the group histogram and counts (2 x groups x cells)
    int64_t grid;   // doubles into the grid buffer: edges (2 buffers of
                    // D x (bins + 1)), histogram, counts (cells each), sums (2)
};

// Per-integral iteration state, on the device.
struct VState {
    double est[kBatchMaxIterations], var[kBatchMaxIterations];
    double value, error;
    int iters, converged;
};

__host__ __device__ inline int64_t vgrid_doubles(int D, int bins) {
    return 2 * (int64_t)D * (bins + 1) + 2 * (int64_t)D * bins + 2;
}

__host__ __device__ inline int64_t vpart_doubles(int blocks, int groups, int cells) {
    return 2 * (int64_t)blocks + 2 * (int64_t)blocks * cells + 2 * (int64_t)groups * cells;
}

struct BatchDev {
    const DisArgs* args;
    const VItem* vi;
    VState* st;
    const nlo_sobol_table* tables;
    double* part;
    double* grid;
    int bins;
};

__device__ __forceinline__ bool vitem_active(const VItem& it, const VState& st) {
    return !st.converged && st.iters < it.max_it;
}

// One pass of every active integral of kind K: work entry (integral, block).
// Occupancy per instance.  FP64: at most 256 VGPRs, two waves per SIMD (T2
// and the qqbar-g terms hold 270-400 at their natural allocation, one wave;
// held to 256 the qqbar-g terms spill 68-268 B per lane and still run 22-30 %
// faster).  FP32 nested I2 / I3: 128 VGPRs, four waves instead of three (I2
// spills 28 B; 4 % off the FP32 batch).  The others keep their allocation.
template <typename T, int K>
constexpr int kDisWaves = std::is_same<T, double>::value ? 2 : (K >= NLO_DIS_QG_I2_NESTED ? 4 : 1);

template <typename T, int K>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kDisWaves<T, K>))) vegas_batch_kernel(BatchDev b, const int2* __restrict__ work) {
    const int2 w = work[blockIdx.x];
    const VItem& it = b.vi[w.x];
    const VState& st = b.st[w.x];
    if (!vitem_active(it, st)) return;
    const int cells = it.D * b.bins;
    const double* edges = b.grid + it.grid + (int64_t)(st.iters & 1) * it.D * (b.bins + 1);
    double* part = b.part + it.part;
    double* ph = part + 2 * (int64_t)it.blocks;
    const VegasView view{b.bins, kBatchSpt, it.n, it.first + (int64_t)st.iters * it.stride, edges, part, ph,
                         ph + (int64_t)it.blocks * cells};
    vegas_pass<T, K, true>(b.args[it.item], view, TableSobol{b.tables + it.table}, w.y);
}

// Histogram partials, first level, for every active integral: block
// (v, g, c) sums cells [64 c, 64 c + 64) of integral v over its g-th range of
// blocks, in block order (vegas_hist_groups_kernel per integral).
__global__ void __launch_bounds__(64) vegas_batch_groups_kernel(BatchDev b, int cblocks) {
    const int v = blockIdx.x / (kHistGroups * cblocks);
    const int rem = blockIdx.x - v * kHistGroups * cblocks;
    const int g = rem / cblocks, c = rem - g * cblocks;
    const VItem& it = b.vi[v];
    if (!vitem_active(it, b.st[v])) return;
    const int cells = it.D * b.bins;
    const int k = c * 64 + threadIdx.x;
    if (g >= it.groups || k >= cells) return;
    const double* ph = b.part + it.part + 2 * (int64_t)it.blocks;
    const double* pc = ph + (int64_t)it.blocks * cells;
    double* gh = b.part + it.part + 2 * (int64_t)it.blocks + 2 * (int64_t)it.blocks * cells;
    double* gc = gh + (int64_t)it.groups * cells;
    const int b0 = g * it.per_group, b1 = min(b0 + it.per_group, it.blocks);
    double h = 0.0, cn = 0.0;
    for (int bb = b0; bb < b1; ++bb) {
        h += ph[(size_t)bb * cells + k];
        cn += pc[(size_t)bb * cells + k];
    }
    gh[(size_t)g * cells + k] = h;
    gc[(size_t)g * cells + k] = cn;
}

// Fixed-order sums of every active integral (vegas_reduce_kernel per
// integral): block (v, 0) the two sums, blocks (v, r > 0) 256 histogram cells.
__global__ void __launch_bounds__(256) vegas_batch_reduce_kernel(BatchDev b, int rblocks) {
    __shared__ double red[4][2][2];  // [wave][sum][value, error]
    const int v = blockIdx.x / rblocks, r = blockIdx.x - v * rblocks;
    const VItem& it = b.vi[v];
    if (!vitem_active(it, b.st[v])) return;
    const int cells = it.D * b.bins;
    const double* part = b.part + it.part;
    double* g = b.grid + it.grid + 2 * (int64_t)it.D * (b.bins + 1);
    double* hist = g;
    double* cnt = g + cells;
    double* sums = g + 2 * cells;
    if (r == 0) {
        double t[2] = {0.0, 0.0}, e[2] = {0.0, 0.0};
        for (int bb = threadIdx.x; bb < it.blocks; bb += blockDim.x) {
            comp_add(t[0], e[0], part[2 * (size_t)bb]);
            comp_add(t[1], e[1], part[2 * (size_t)bb + 1]);
        }
        block_comp_sums<2>(t, e, red, sums);
        return;
    }
    const int k = (r - 1) * blockDim.x + threadIdx.x;
    if (k >= cells) return;
    const double* gh = part + 2 * (int64_t)it.blocks + 2 * (int64_t)it.blocks * cells;
    const double* gc = gh + (int64_t)it.groups * cells;
    double h = 0.0, cn = 0.0;
    for (int gg = 0; gg < it.groups; ++gg) {
        h += gh[(size_t)gg * cells + k];
        cn += gc[(size_t)gg * cells + k];
    }
    hist[k] = h;
    cnt[k] = cn;
}

// Grid refinement of every active integral, one (integral, dimension) per
// thread: the current edge buffer into the other one.
__global__ void __launch_bounds__(64) vegas_batch_adapt_kernel(BatchDev b, const int2* __restrict__ rows,
                                                               int n_rows, double eps) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_rows) return;
    const int2 rw = rows[k];
    const VItem& it = b.vi[rw.x];
    const VState& st = b.st[rw.x];
    if (!vitem_active(it, st)) return;
    const int E = it.D * (b.bins + 1), cells = it.D * b.bins;
    double* g = b.grid + it.grid;
    const double* e = g + (int64_t)(st.iters & 1) * E + (int64_t)rw.y * (b.bins + 1);
    double* o = g + (int64_t)((st.iters + 1) & 1) * E + (int64_t)rw.y * (b.bins + 1);
    const double* hist = g + 2 * (int64_t)E + (int64_t)rw.y * b.bins;
    vegas_adapt_row(e, hist, hist + cells, b.bins, eps, o);
}

// s + x into the double-double (s, c) (TwoSum; Knuth).
__device__ __forceinline__ void two_sum_add(double& s, double& c, double x) {
#pragma clang fp contract(off)
    const double t = s + x;
    const double bp = t - s;
    const double err = (s - (t - bp)) + (x - bp);
    s = t;
    c += err;
}

// The iteration's estimate and variance of every active integral, the
// inverse-variance combination of its iterations so far and the
// convergence test (vegas(), S/numerics/integration.py:500-589): estimate =
// sum w / n, variance = max((sum w^2 - (sum w)^2 / n) / (n - 1) / n, tiny),
// value = sum(e / v) / sum(1 / v), error = sqrt(1 / sum(1 / v)), with the
// sums carried in double-double (the host driver uses math.fsum), converged
// after min_iterations once error <= max(epsabs, epsrel |value|).
__global__ void __launch_bounds__(64) vegas_batch_stats_kernel(BatchDev b, int n_vi, double tiny, double epsrel,
                                                               double epsabs) {
#pragma clang fp contract(off)
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vi) return;
    const VItem& it = b.vi[v];
    VState& st = b.st[v];
    if (!vitem_active(it, st)) return;
    const double* sums = b.grid + it.grid + 2 * (int64_t)it.D * (b.bins + 1) + 2 * (int64_t)it.D * b.bins;
    const double n = (double)it.n;
    const double total = sums[0], total_sq = sums[1];
    const double estimate = total / n;
    const double variance = fmax((fmax(total_sq - total * total / n, 0.0) / (n - 1.0)) / n, tiny);
    const int k = st.iters;
    st.est[k] = estimate;
    st.var[k] = variance;
    double ws = 0.0, wc = 0.0, es = 0.0, ec = 0.0;
    for (int i = 0; i <= k; ++i) {
        const double w = 1.0 / st.var[i];
        two_sum_add(ws, wc, w);
        two_sum_add(es, ec, w * st.est[i]);
    }
    const double wsum = ws + wc, esum = es + ec;
    st.value = esum / wsum;
    st.error = sqrt(1.0 / wsum);
    st.iters = k + 1;
    st.converged = (k + 1 >= it.min_it && st.error <= fmax(epsabs, epsrel * fabs(st.value))) ? 1 : 0;
}

// Zero iteration state and the starting grid of every integral.
__global__ void __launch_bounds__(256) vegas_batch_init_kernel(BatchDev b, const double* __restrict__ edges0) {
    const int v = blockIdx.x;
    const VItem& it = b.vi[v];
    if (threadIdx.x == 0) {
        VState& st = b.st[v];
        for (int i = 0; i < kBatchMaxIterations; ++i) {
            st.est[i] = 0.0;
            st.var[i] = 0.0;
        }
        st.value = 0.0;
        st.error = 0.0;
        st.iters = 0;
        st.converged = 0;
    }
    double* g = b.grid + it.grid;
    for (int i = threadIdx.x; i < it.D * (b.bins + 1); i += blockDim.x) g[i] = edges0[i % (b.bins + 1)];
}

// Gauss-Legendre items of kind K: block (0 | 1, item) the fine or the coarse rule.
template <typename T, int K>
__global__ void __launch_bounds__(256) gl_batch_kernel(const DisArgs* __restrict__ args,
                                                       const int4* __restrict__ work, double* __restrict__ out) {
    const int4 w = work[blockIdx.y];  // (item, points, output slot, -)
    gl_pass<T, K>(args[w.x], w.y, blockIdx.x, out + 2 * (size_t)w.z);
}

// Per item: value, error, evaluations, converged (0 / 1).  map[i] = (0, v)
// for VEGAS integral v, (1, g) for Gauss-Legendre item g (with its points).
__global__ void __launch_bounds__(256) batch_results_kernel(BatchDev b, const int4* __restrict__ map, int n_items,
                                                            const double* __restrict__ gl_out, double epsrel,
                                                            double epsabs, double* __restrict__ results) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items) return;
    const int4 m = map[i];
    double* o = results + 4 * (size_t)i;
    if (m.x == 0) {
        const VItem& it = b.vi[m.y];
        const VState& st = b.st[m.y];
        o[0] = st.value;
        o[1] = st.error;
        o[2] = (double)st.iters * (double)it.n;
        o[3] = st.converged ? 1.0 : 0.0;
    } else {
        const double fine = gl_out[2 * (size_t)m.y], coarse = gl_out[2 * (size_t)m.y + 1];
        const int p = m.z, pc = p / 2 > 2 ? p / 2 : 2;
        const double err = fabs(fine - coarse);
        o[0] = fine;
        o[1] = err;
        o[2] = (double)p * p + (double)pc * pc;
        o[3] = err <= fmax(epsabs, epsrel * fabs(fine)) ? 1.0 : 0.0;
    }
}

}  // namespace

// -------------------------------------------------------------------- host

DisDipole make_dipole(const nlo_dis_dipole& d) {
    DisDipole o{};
    o.kind = d.kind;
    o.qs0sqr = d.qs0sqr;
    o.lambda = d.lambda;
    o.gamma = d.gamma;
    o.y0_gbw = std::log(1.0 / d.x0);
    o.R = d.R;
    o.T = d.T;
    o.r = d.d_r;
    o.Y = d.d_Y;
    o.N = d.d_N;
    o.initialized = d.initialized;
    o.y_init = d.y_init;
    o.maxr_interpolate = d.maxr_interpolate;
    o.init_coef = d.d_init_coef;
    o.bilinear = d.bilinear;
    o.local_coef = d.d_local_coef;
    return o;
}

int check_dipole(const nlo_dis_dipole* d) {
    NLO_REQUIRE(d != nullptr, "dipole is NULL");
    if (d->kind == NLO_DIPOLE_GBW) {
        NLO_REQUIRE(d->x0 > 0.0 && d->qs0sqr > 0.0, "GBW needs x0 > 0 and Qs0sqr > 0");
        return NLO_OK;
    }
    NLO_REQUIRE(d->kind == NLO_DIPOLE_BK_TABLE, "unknown dipole kind %d", d->kind);
    NLO_REQUIRE(d->R >= 9 && d->T >= 1, "BK table needs R >= 9 r points and T >= 1 rapidities (R=%d, T=%d)",
                d->R, d->T);
    NLO_REQUIRE(d->d_r && d->d_Y && d->d_N, "BK table arrays are NULL");
    NLO_REQUIRE(!d->initialized || d->d_init_coef, "initialised BK dipole without coefficients");
    return NLO_OK;
}

int make_args(const nlo_dis_desc* desc, DisArgs& a, const double* ends) {
    NLO_REQUIRE(desc != nullptr, "desc is NULL");
    int rc = check_dipole(&desc->dipole);
    if (rc != NLO_OK) return rc;
    NLO_REQUIRE(desc->Q2 > 0.0 && desc->xbj > 0.0 && desc->xbj <= 1.0, "Q2 must be positive and xbj in (0, 1]");
    NLO_REQUIRE(desc->n_quarks >= 1 && desc->n_quarks <= NLO_DIS_MAX_QUARKS, "n_quarks must be 1..%d",
                NLO_DIS_MAX_QUARKS);
    NLO_REQUIRE(desc->maxr > 0.0 && desc->c2_alpha > 0.0 && desc->max_alpha > 0.0,
                "maxr, C2_alpha and max_alpha_s_freeze must be positive");
    a = DisArgs{};
    a.Q2 = desc->Q2;
    a.xbj = desc->xbj;
    a.transverse = desc->transverse != 0;
    a.nq = desc->n_quarks;
    for (int q = 0; q < a.nq; ++q) {
        a.mass[q] = desc->mass[q];
        a.charge[q] = desc->charge[q];
    }
    a.maxr = desc->maxr;
    a.finite_nc = desc->finite_nc != 0;
    a.parent_coupling = desc->parent_coupling != 0;
    a.coup.b0 = (11.0 * kNc - 2.0 * desc->nf_alphas) / (12.0 * kPiD);
    a.coup.scale_num = 4.0 * desc->c2_alpha / (0.241 * 0.241);  // LambdaQCD (S/dis/config.py:9)
    a.coup.max_alpha = desc->max_alpha;
    a.coup.smooth = desc->smooth_coupling != 0;
    a.dip = make_dipole(desc->dipole);
    const double x0 = desc->dipole.kind == NLO_DIPOLE_GBW ? desc->dipole.x0 : 1.0;  // dipole.X0()
    a.y_lo = std::fmax(std::log(x0 / desc->xbj), 0.0);
    {
        // evolution_rapidity_dipole (S/dis/observables.py:55-68)
        const double W2 = (1.0 - desc->xbj) * desc->Q2 / desc->xbj;
        const double mf = desc->mass[0];
        const double X = desc->mass_dependent_x ? (desc->Q2 + 4.0 * mf * mf) / (desc->Q2 + W2) : desc->Q2 / (desc->Q2 + W2);
        a.y_dip = std::log(1.0 / X);
    }
    a.z2min = kQ0sqr / (desc->Q2 / desc->xbj);
    if (desc->dipole.kind == NLO_DIPOLE_GBW) {
        a.min_r = 1e-30;
        a.max_r = 1e3;
    } else if (ends) {
        a.min_r = ends[0];
        a.max_r = ends[1];
    } else {
        // r_min / r_max are device values: the table's first and last r
        double e[2];
        NLO_HIP_CHECK(hipMemcpy(&e[0], desc->dipole.d_r, sizeof(double), hipMemcpyDeviceToHost));
        NLO_HIP_CHECK(hipMemcpy(&e[1], desc->dipole.d_r + desc->dipole.R - 1, sizeof(double), hipMemcpyDeviceToHost));
        a.min_r = e[0];
        a.max_r = e[1];
    }
    NLO_REQUIRE(desc->nested_points >= 8 && desc->nested_points <= kMaxNested, "nested_points must be 8..%d",
                kMaxNested);
    NLO_REQUIRE(desc->dipole.kind != NLO_DIPOLE_BK_TABLE || desc->dipole.bilinear || desc->dipole.d_local_coef,
                "a SPLINE_LINEAR BK-table dipole needs its local-spline table (nlo_dis_bk_local_splines) "
                "for the integrands");
    a.nested_points = desc->nested_points;
    return NLO_OK;
}

namespace {

#define DIS_DISPATCH(MACRO)           \
    switch (contribution) {           \
        MACRO(NLO_DIS_LO)             \
        MACRO(NLO_DIS_DIP_T0)         \
        MACRO(NLO_DIS_DIP_T1)         \
        MACRO(NLO_DIS_DIP_T2)         \
        MACRO(NLO_DIS_DIP_L_CONST)    \
        MACRO(NLO_DIS_DIP_L_AB)       \
        MACRO(NLO_DIS_DIP_L_CD)       \
        MACRO(NLO_DIS_QG_I1)          \
        MACRO(NLO_DIS_QG_I2)          \
        MACRO(NLO_DIS_QG_I3)          \
        MACRO(NLO_DIS_QG_I2_NESTED)   \
        MACRO(NLO_DIS_QG_I3_NESTED)   \
        default:                      \
            break;                    \
    }

// The arithmetic a contribution uses on the FP32 path.  FP64 where the
// integrand is a small difference of O(1) terms:
//  * qqbar-g I1 (uvsub + tripole I1) cancels to O(z2) in the soft-gluon region
//    while its Jacobian carries 1 / z2 (measured in FP32: up to 9e-3 relative
//    per 2^22-sample pass through the same grid, 35 standard errors);
//  * the longitudinal "ab" term, 2 K0(b1) - K0(b2) - K0(b3) with b2, b3 -> b1
//    for light quarks, and the transverse "T1" term, V1 / VMS1 differences of
//    K1 and K0 at nearby arguments (the reference evaluates both in FP64
//    inside its FP32 kernel, K/dis/lo_dipole/kernel.cu:205-227, 278-331).
int path_precision(int contribution, int precision) {
    switch (contribution) {
        case NLO_DIS_QG_I1:
        case NLO_DIS_DIP_L_AB:
        case NLO_DIS_DIP_T1:
            return NLO_PREC_FP64;
        default:
            return precision;
    }
}

struct VegasPlan {
    int spt, blocks, groups, per_group;
    size_t part, hist, group, workspace, lds;
};

VegasPlan vegas_plan(int contribution, int bins, int64_t n) {
    const int D = kDims[contribution];
    VegasPlan p{};
    const int64_t groups = (n + 63) / 64;
    int spt = 8;
    while (spt > 1 && (groups + 4 * spt - 1) / (4 * spt) < 2048) spt >>= 1;
    p.spt = spt;
    p.blocks = (int)((groups + 4 * spt - 1) / (4 * spt));
    p.part = align_up((size_t)p.blocks * 2 * sizeof(double), 256);
    p.hist = align_up((size_t)p.blocks * D * bins * sizeof(double), 256);
    p.per_group = (p.blocks + kHistGroups - 1) / kHistGroups;
    p.groups = (p.blocks + p.per_group - 1) / p.per_group;
    p.group = align_up((size_t)p.groups * D * bins * sizeof(double), 256);
    p.workspace = p.part + 2 * p.hist + 2 * p.group;
    p.lds = vegas_layout(D, bins, spt, true).total;
    return p;
}

// Host plan of a batch: the device records and the workspace layout.
struct BatchPlan {
    struct Launch {
        int K, prec, off, count;
        size_t lds;
    };
    std::vector<DisArgs> args;
    std::vector<VItem> vi;
    std::vector<int2> work;  // grouped by launch
    std::vector<Launch> launches;
    std::vector<int2> rows;
    std::vector<int4> gl_work;  // grouped by launch
    std::vector<Launch> gl_launches;
    std::vector<int4> map;
    int max_cells = 0, max_iterations = 0;
    int64_t part_doubles = 0, grid_doubles = 0;
    size_t o_args = 0, o_vi = 0, o_st = 0, o_tables = 0, o_work = 0, o_rows = 0, o_gl_work = 0, o_gl_out = 0,
           o_map = 0, o_part = 0, o_grid = 0, total = 0;
};

// Validates the batch and lays it out; with build_args the per-item DisArgs
// are formed too (which reads each BK table's end points from the device once).
int batch_plan(const nlo_dis_batch_desc* d, BatchPlan& P, bool build_args) {
    NLO_REQUIRE(d != nullptr, "nlo_dis_batch: desc is NULL");
    NLO_REQUIRE(d->n_items >= 1 && d->items, "nlo_dis_batch: no items");
    NLO_REQUIRE(d->n_tables >= 0 && (d->n_tables == 0 || d->tables), "nlo_dis_batch: tables missing");
    NLO_REQUIRE(d->bins >= 2 && d->bins <= 255, "nlo_dis_batch: bins must be 2..255");
    NLO_REQUIRE(d->precision == NLO_PREC_FP64 || d->precision == NLO_PREC_FP32,
                "nlo_dis_batch: precision must be FP64 or FP32");
    NLO_REQUIRE(d->epsrel >= 0.0 && d->epsabs >= 0.0, "nlo_dis_batch: tolerances must be non-negative");
    const int bins = d->bins;
    // work entries and Gauss-Legendre items per launch (kind, arithmetic)
    std::vector<std::vector<int2>> work(NLO_DIS_KINDS * 2);
    std::vector<std::vector<int4>> glw(NLO_DIS_KINDS * 2);
    std::vector<std::pair<const double*, std::array<double, 2>>> ends;
    P.map.resize(d->n_items);
    int n_gl = 0;
    for (int i = 0; i < d->n_items; ++i) {
        const nlo_dis_item& item = d->items[i];
        const int K = item.contribution;
        NLO_REQUIRE(K >= 0 && K < NLO_DIS_KINDS, "nlo_dis_batch: item %d: unknown contribution %d", i, K);
        const int prec = path_precision(K, d->precision);
        const int slot = K * 2 + (prec == NLO_PREC_FP64 ? 0 : 1);
        if (build_args) {
            DisArgs a;
            const double* e = nullptr;
            if (item.desc.dipole.kind == NLO_DIPOLE_BK_TABLE && item.desc.dipole.d_r &&
                item.desc.dipole.R >= 2) {
                for (auto& c : ends)
                    if (c.first == item.desc.dipole.d_r) e = c.second.data();
                if (!e) {
                    std::array<double, 2> v{};
                    NLO_HIP_CHECK(hipMemcpy(&v[0], item.desc.dipole.d_r, sizeof(double), hipMemcpyDeviceToHost));
                    NLO_HIP_CHECK(hipMemcpy(&v[1], item.desc.dipole.d_r + item.desc.dipole.R - 1, sizeof(double),
                                            hipMemcpyDeviceToHost));
                    ends.push_back({item.desc.dipole.d_r, v});
                    e = ends.back().second.data();
                }
            }
            const int rc = make_args(&item.desc, a, e);
            if (rc != NLO_OK) return rc;
            P.args.push_back(a);
        }
        if (item.method == NLO_DIS_METHOD_GAUSS_LEGENDRE) {
            NLO_REQUIRE(kDims[K] == 2, "nlo_dis_batch: item %d: Gauss-Legendre needs a 2-D contribution", i);
            NLO_REQUIRE(item.n >= 4 && item.n <= 128, "nlo_dis_batch: item %d: points must be 4..128", i);
            glw[slot].push_back(make_int4(i, (int)item.n, n_gl, 0));
            P.map[i] = make_int4(1, n_gl, (int)item.n, 0);
            ++n_gl;
            continue;
        }
        NLO_REQUIRE(item.method == NLO_DIS_METHOD_VEGAS, "nlo_dis_batch: item %d: unknown method %d", i, item.method);
        const int D = kDims[K];
        NLO_REQUIRE(item.sobol_table >= 0 && item.sobol_table < d->n_tables,
                    "nlo_dis_batch: item %d: Sobol table %d out of range", i, item.sobol_table);
        NLO_REQUIRE(d->tables[item.sobol_table].dim == D,
                    "nlo_dis_batch: item %d: the Sobol table must have the contribution's dimension %d", i, D);
        NLO_REQUIRE(item.n >= 2, "nlo_dis_batch: item %d: at least two samples per iteration", i);
        NLO_REQUIRE(item.max_iterations >= 1 && item.max_iterations <= kBatchMaxIterations &&
                        item.min_iterations >= 1 && item.min_iterations <= item.max_iterations,
                    "nlo_dis_batch: item %d: iterations (min %d, max %d) outside 1 <= min <= max <= %d", i,
                    item.min_iterations, item.max_iterations, kBatchMaxIterations);
        NLO_REQUIRE(item.first_index >= 0 && item.first_index % 64 == 0,
                    "nlo_dis_batch: item %d: first_index must be a non-negative multiple of 64", i);
        VItem v{};
        v.item = i;
        v.K = K;
        v.D = D;
        v.table = item.sobol_table;
        v.n = item.n;
        v.stride = (item.n + 63) / 64 * 64;
        v.first = item.first_index;
        NLO_REQUIRE(v.first + v.stride * item.max_iterations <= ((int64_t)1 << NLO_SOBOL_BITS),
                    "nlo_dis_batch: item %d: the Sobol sequence (2^30 points) is too short for this budget", i);
        v.max_it = item.max_iterations;
        v.min_it = item.min_iterations;
        const int64_t g64 = (item.n + 63) / 64;
        const int64_t blocks = (g64 + 4 * kBatchSpt - 1) / (4 * kBatchSpt);
        NLO_REQUIRE(blocks <= (1 << 24), "nlo_dis_batch: item %d: too many samples", i);
        v.blocks = (int)blocks;
        v.per_group = (v.blocks + kHistGroups - 1) / kHistGroups;
        v.groups = (v.blocks + v.per_group - 1) / v.per_group;
        const int cells = D * bins;
        v.part = P.part_doubles;
        P.part_doubles += vpart_doubles(v.blocks, v.groups, cells);
        v.grid = P.grid_doubles;
        P.grid_doubles += vgrid_doubles(D, bins);
        const int vi = (int)P.vi.size();
        for (int b = 0; b < v.blocks; ++b) work[slot].push_back(make_int2(vi, b));
        for (int dd = 0; dd < D; ++dd) P.rows.push_back(make_int2(vi, dd));
        P.map[i] = make_int4(0, vi, 0, 0);
        P.max_cells = std::max(P.max_cells, cells);
        P.max_iterations = std::max(P.max_iterations, v.max_it);
        P.vi.push_back(v);
    }
    for (int slot = 0; slot < NLO_DIS_KINDS * 2; ++slot) {
        const int K = slot / 2, prec = (slot & 1) ? NLO_PREC_FP32 : NLO_PREC_FP64;
        if (!work[slot].empty()) {
            const size_t lds = vegas_layout(kDims[K], bins, kBatchSpt, true).total;
            NLO_REQUIRE(lds <= 64 * 1024, "nlo_dis_batch: LDS %zu bytes exceeds 64 KiB (fewer bins)", lds);
            P.launches.push_back({K, prec, (int)P.work.size(), (int)work[slot].size(), lds});
            P.work.insert(P.work.end(), work[slot].begin(), work[slot].end());
        }
        if (!glw[slot].empty()) {
            P.gl_launches.push_back({K, prec, (int)P.gl_work.size(), (int)glw[slot].size(), 0});
            P.gl_work.insert(P.gl_work.end(), glw[slot].begin(), glw[slot].end());
        }
    }
    size_t o = 0;
    auto section = [&](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes, 256);
        return at;
    };
    P.o_args = section(sizeof(DisArgs) * (size_t)d->n_items);
    P.o_vi = section(sizeof(VItem) * P.vi.size());
    P.o_st = section(sizeof(VState) * P.vi.size());
    P.o_tables = section(sizeof(nlo_sobol_table) * (size_t)d->n_tables);
    P.o_work = section(sizeof(int2) * P.work.size());
    P.o_rows = section(sizeof(int2) * P.rows.size());
    P.o_gl_work = section(sizeof(int4) * P.gl_work.size());
    P.o_gl_out = section(sizeof(double) * 2 * (size_t)n_gl);
    P.o_map = section(sizeof(int4) * (size_t)d->n_items);
    P.o_part = section(sizeof(double) * (size_t)P.part_doubles);
    P.o_grid = section(sizeof(double) * (size_t)P.grid_doubles);
    P.total = o;
    return NLO_OK;
}

template <typename T, int K>
void launch_vegas_batch(const BatchPlan::Launch& L, const BatchDev& bd, const int2* work, hipStream_t s) {
    hipLaunchKernelGGL((vegas_batch_kernel<T, K>), dim3(L.count), dim3(256), L.lds, s, bd, work + L.off);
}

template <typename T, int K>
void launch_gl_batch(const BatchPlan::Launch& L, const DisArgs* args, const int4* work, double* out,
                     hipStream_t s) {
    hipLaunchKernelGGL((gl_batch_kernel<T, K>), dim3(2, L.count), dim3(256), 0, s, args, work + L.off, out);
}

// One iteration of every active VEGAS integral of the batch on stream s.
int enqueue_batch_iteration(const BatchPlan& P, const BatchDev& bd, const int2* work, const int2* rows, double eps,
                            double tiny, double epsrel, double epsabs, hipStream_t s) {
    for (const auto& L : P.launches) {
        const int contribution = L.K;
#define DIS_VB(K)                                                             \
    case K:                                                                   \
        if (L.prec == NLO_PREC_FP64) launch_vegas_batch<double, K>(L, bd, work, s); \
        else launch_vegas_batch<float, K>(L, bd, work, s);                    \
        break;
        DIS_DISPATCH(DIS_VB)
#undef DIS_VB
        NLO_LAUNCH_CHECK();
    }
    const int nv = (int)P.vi.size();
    const int cblocks = (P.max_cells + 63) / 64;
    hipLaunchKernelGGL(vegas_batch_groups_kernel, dim3((unsigned)((int64_t)nv * kHistGroups * cblocks)), dim3(64), 0,
                       s, bd, cblocks);
    NLO_LAUNCH_CHECK();
    const int rblocks = 1 + (P.max_cells + 255) / 256;
    hipLaunchKernelGGL(vegas_batch_reduce_kernel, dim3((unsigned)((int64_t)nv * rblocks)), dim3(256), 0, s, bd,
                       rblocks);
    NLO_LAUNCH_CHECK();
    const int nr = (int)P.rows.size();
    hipLaunchKernelGGL(vegas_batch_adapt_kernel, dim3((nr + 63) / 64), dim3(64), 0, s, bd, rows, nr, eps);
    NLO_LAUNCH_CHECK();
    hipLaunchKernelGGL(vegas_batch_stats_kernel, dim3((nv + 63) / 64), dim3(64), 0, s, bd, nv, tiny, epsrel, epsabs);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

}  // namespace
}  // namespace dis
}  // namespace nlo

using namespace nlo;
using namespace nlo::dis;

extern "C" {

int nlo_dis_dimensions(int32_t contribution) {
    return (contribution >= 0 && contribution < NLO_DIS_KINDS) ? kDims[contribution] : -1;
}

int nlo_dis_special(int32_t fn, int64_t n, const double* d_x, int32_t precision, double* d_out,
                    hipStream_t stream) {
    clear_error();
    NLO_REQUIRE(fn >= NLO_FN_K0 && fn <= NLO_FN_LI2, "unknown special function %d", fn);
    NLO_REQUIRE(n >= 0, "n must be >= 0");
    NLO_REQUIRE(precision == NLO_PREC_FP64 || precision == NLO_PREC_FP32, "precision must be FP64 or FP32");
    if (n == 0) return NLO_OK;
    NLO_REQUIRE(d_x && d_out, "NULL array");
    const dim3 grid((unsigned)((n + 255) / 256));
    if (precision == NLO_PREC_FP64)
        hipLaunchKernelGGL(special_kernel<double>, grid, dim3(256), 0, stream, fn, n, d_x, d_out);
    else
        hipLaunchKernelGGL(special_kernel<float>, grid, dim3(256), 0, stream, fn, n, d_x, d_out);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int nlo_dis_dipole_amplitude(const nlo_dis_dipole* dipole, int64_t n, const double* d_r, const double* d_Y,
                             double* d_out, hipStream_t stream) {
    clear_error();
    int rc = check_dipole(dipole);
    if (rc != NLO_OK) return rc;
    NLO_REQUIRE(n >= 0, "n must be >= 0");
    if (n == 0) return NLO_OK;
    NLO_REQUIRE(d_r && d_Y && d_out, "NULL array");
    hipLaunchKernelGGL(dipole_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, make_dipole(*dipole),
                       n, d_r, d_Y, d_out);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int nlo_dis_bk_initialize(const nlo_dis_dipole* dipole, double Y, double* d_coef, double* d_maxr,
                          hipStream_t stream) {
    clear_error();
    int rc = check_dipole(dipole);
    if (rc != NLO_OK) return rc;
    NLO_REQUIRE(dipole->kind == NLO_DIPOLE_BK_TABLE, "nlo_dis_bk_initialize needs a BK table dipole");
    NLO_REQUIRE(Y >= 0.0, "cannot initialise the interpolation at negative Y=%g", Y);
    NLO_REQUIRE(d_coef && d_maxr, "NULL output");
    const size_t lds = (size_t)3 * dipole->R * sizeof(double);
    hipLaunchKernelGGL(bk_init_kernel, dim3(1), dim3(64), lds, stream, make_dipole(*dipole), Y, d_coef, d_maxr);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int nlo_dis_bk_local_splines(const nlo_dis_dipole* dipole, double* d_coef, hipStream_t stream) {
    clear_error();
    int rc = check_dipole(dipole);
    if (rc != NLO_OK) return rc;
    NLO_REQUIRE(dipole->kind == NLO_DIPOLE_BK_TABLE, "nlo_dis_bk_local_splines needs a BK table dipole");
    NLO_REQUIRE(d_coef != nullptr, "NULL output");
    DisDipole d = make_dipole(*dipole);
    d.local_coef = nullptr;
    const int n = d.T * (d.R - 1);
    hipLaunchKernelGGL(bk_local_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, d, d_coef);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int nlo_dis_integrand(const nlo_dis_desc* desc, int32_t contribution, int64_t n, const double* d_x,
                      int32_t precision, double* d_out, hipStream_t stream) {
    clear_error();
    NLO_REQUIRE(contribution >= 0 && contribution < NLO_DIS_KINDS, "unknown contribution %d", contribution);
    NLO_REQUIRE(precision == NLO_PREC_FP64 || precision == NLO_PREC_FP32, "precision must be FP64 or FP32");
    precision = path_precision(contribution, precision);
    NLO_REQUIRE(n >= 0, "n must be >= 0");
    DisArgs a;
    int rc = make_args(desc, a);
    if (rc != NLO_OK) return rc;
    if (n == 0) return NLO_OK;
    NLO_REQUIRE(d_x && d_out, "NULL array");
    const dim3 grid((unsigned)((n + 255) / 256));
#define DIS_POINT(K)                                                                                     \
    case K:                                                                                              \
        if (precision == NLO_PREC_FP64)                                                                  \
            hipLaunchKernelGGL((integrand_kernel<double, K>), grid, dim3(256), 0, stream, a, n, d_x, d_out); \
        else                                                                                             \
            hipLaunchKernelGGL((integrand_kernel<float, K>), grid, dim3(256), 0, stream, a, n, d_x, d_out);  \
        break;
    DIS_DISPATCH(DIS_POINT)
#undef DIS_POINT
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int nlo_dis_qg_i23(const nlo_dis_desc* desc, int64_t n, const double* d_x, int32_t precision, double* d_i2,
                   double* d_i3, hipStream_t stream) {
    clear_error();
    NLO_REQUIRE(precision == NLO_PREC_FP64 || precision == NLO_PREC_FP32, "precision must be FP64 or FP32");
    NLO_REQUIRE(n >= 0, "n must be >= 0");
    DisArgs a;
    int rc = make_args(desc, a);
    if (rc != NLO_OK) return rc;
    if (n == 0) return NLO_OK;
    NLO_REQUIRE(d_x && d_i2 && d_i3, "NULL array");
    const dim3 grid((unsigned)((n + 255) / 256));
    if (precision == NLO_PREC_FP64)
        hipLaunchKernelGGL(i23_kernel<double>, grid, dim3(256), 0, stream, a, n, d_x, d_i2, d_i3);
    else
        hipLaunchKernelGGL(i23_kernel<float>, grid, dim3(256), 0, stream, a, n, d_x, d_i2, d_i3);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

size_t nlo_dis_vegas_workspace(int32_t contribution, int32_t bins, int64_t n) {
    if (contribution < 0 || contribution >= NLO_DIS_KINDS || bins < 1 || n < 1) return 0;
    return vegas_plan(contribution, bins, n).workspace;
}

int nlo_dis_vegas(const nlo_dis_desc* desc, int32_t contribution, int32_t bins, const double* d_edges,
                  const nlo_sobol_table* sobol, int64_t first_index, int64_t n, int32_t precision, double* d_sums,
                  double* d_hist, double* d_count, void* d_workspace, size_t workspace_bytes, hipStream_t stream) {
    clear_error();
    NLO_REQUIRE(contribution >= 0 && contribution < NLO_DIS_KINDS, "unknown contribution %d", contribution);
    NLO_REQUIRE(precision == NLO_PREC_FP64 || precision == NLO_PREC_FP32, "precision must be FP64 or FP32");
    precision = path_precision(contribution, precision);
    const int D = kDims[contribution];
    NLO_REQUIRE(bins >= 2 && bins <= 255, "bins must be 2..255");
    NLO_REQUIRE(n >= 1, "n must be >= 1");
    NLO_REQUIRE(sobol && sobol->dim == D, "the Sobol table must have the contribution's dimension %d", D);
    NLO_REQUIRE(first_index >= 0 && first_index % 64 == 0, "first_index must be a non-negative multiple of 64");
    NLO_REQUIRE(first_index + n <= (int64_t)1 << NLO_SOBOL_BITS, "Sobol indices beyond 2^30");
    NLO_REQUIRE(d_edges && d_sums, "NULL array");
    NLO_REQUIRE((d_hist == nullptr) == (d_count == nullptr), "d_hist and d_count must both be given or both NULL");
    DisArgs a;
    int rc = make_args(desc, a);
    if (rc != NLO_OK) return rc;
    const VegasPlan pl = vegas_plan(contribution, bins, n);
    if (workspace_bytes < pl.workspace) {
        set_error("nlo_dis_vegas: workspace %zu bytes < %zu", workspace_bytes, pl.workspace);
        return NLO_ERR_WORKSPACE;
    }
    NLO_REQUIRE(d_workspace != nullptr, "NULL workspace");
    VegasArgs v{};
    v.bins = bins;
    v.spt = pl.spt;
    v.blocks = pl.blocks;
    v.n = n;
    v.first = first_index;
    v.edges = d_edges;
    unsigned char* ws = static_cast<unsigned char*>(d_workspace);
    v.part = reinterpret_cast<double*>(ws);
    v.part_hist = reinterpret_cast<double*>(ws + pl.part);
    v.part_cnt = reinterpret_cast<double*>(ws + pl.part + pl.hist);
    for (int d = 0; d < D; ++d) {
        for (int b = 0; b < NLO_SOBOL_BITS; ++b) v.dir[d][b] = sobol->direction[d][b];
        v.shift[d] = sobol->shift[d];
        v.first_q[d] = sobol->first[d];
    }
    const bool hist = d_hist != nullptr;
    const size_t lds = vegas_layout(D, bins, pl.spt, hist).total;
    NLO_REQUIRE(lds <= 64 * 1024, "LDS %zu bytes exceeds 64 KiB", lds);
#define DIS_VEGAS(K)                                                                                            \
    case K:                                                                                                     \
        if (precision == NLO_PREC_FP64) {                                                                       \
            if (hist) hipLaunchKernelGGL((vegas_kernel<double, K, true>), dim3(pl.blocks), dim3(256), lds, stream, a, v); \
            else hipLaunchKernelGGL((vegas_kernel<double, K, false>), dim3(pl.blocks), dim3(256), lds, stream, a, v);     \
        } else {                                                                                                \
            if (hist) hipLaunchKernelGGL((vegas_kernel<float, K, true>), dim3(pl.blocks), dim3(256), lds, stream, a, v);  \
            else hipLaunchKernelGGL((vegas_kernel<float, K, false>), dim3(pl.blocks), dim3(256), lds, stream, a, v);      \
        }                                                                                                       \
        break;
    DIS_DISPATCH(DIS_VEGAS)
#undef DIS_VEGAS
    NLO_LAUNCH_CHECK();
    const int cells = D * bins;
    double* gh = reinterpret_cast<double*>(ws + pl.part + 2 * pl.hist);
    double* gc = reinterpret_cast<double*>(ws + pl.part + 2 * pl.hist + pl.group);
    if (hist) {
        hipLaunchKernelGGL(vegas_hist_groups_kernel, dim3((cells + 63) / 64, pl.groups), dim3(64), 0, stream,
                           pl.blocks, cells, pl.per_group, v.part_hist, v.part_cnt, gh, gc);
        NLO_LAUNCH_CHECK();
    }
    const int rblocks = 1 + (hist ? (cells + 255) / 256 : 0);
    hipLaunchKernelGGL(vegas_reduce_kernel, dim3(rblocks), dim3(256), 0, stream, pl.blocks, cells, pl.groups,
                       v.part, gh, gc, d_sums, d_hist, d_count);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

int nlo_dis_gauss_legendre(const nlo_dis_desc* desc, int32_t contribution, int32_t points, int32_t precision,
                           double* d_out, hipStream_t stream) {
    clear_error();
    NLO_REQUIRE(contribution >= 0 && contribution < NLO_DIS_KINDS && kDims[contribution] == 2,
                "Gauss-Legendre needs a two-dimensional contribution (got %d)", contribution);
    NLO_REQUIRE(points >= 4 && points <= 128, "points must be 4..128");
    NLO_REQUIRE(precision == NLO_PREC_FP64 || precision == NLO_PREC_FP32, "precision must be FP64 or FP32");
    NLO_REQUIRE(d_out != nullptr, "NULL output");
    DisArgs a;
    int rc = make_args(desc, a);
    if (rc != NLO_OK) return rc;
#define DIS_GL(K)                                                                                          \
    case K:                                                                                                \
        if (precision == NLO_PREC_FP64)                                                                    \
            hipLaunchKernelGGL((gl_kernel<double, K>), dim3(2), dim3(256), 0, stream, a, points, d_out); \
        else                                                                                               \
            hipLaunchKernelGGL((gl_kernel<float, K>), dim3(2), dim3(256), 0, stream, a, points, d_out);  \
        break;
    switch (contribution) {
        DIS_GL(NLO_DIS_LO)
        DIS_GL(NLO_DIS_DIP_T0)
        DIS_GL(NLO_DIS_DIP_L_CONST)
        default:
            break;
    }
#undef DIS_GL
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

size_t nlo_dis_batch_workspace(const nlo_dis_batch_desc* desc) {
    BatchPlan P;
    if (batch_plan(desc, P, false) != NLO_OK) return 0;
    return P.total;
}

int nlo_dis_batch_run(const nlo_dis_batch_desc* desc, double* d_results, void* d_workspace, size_t workspace_bytes,
                      hipStream_t stream) {
    clear_error();
    BatchPlan P;
    int rc = batch_plan(desc, P, true);
    if (rc != NLO_OK) return rc;
    NLO_REQUIRE(d_results != nullptr, "nlo_dis_batch_run: NULL results");
    NLO_REQUIRE(P.vi.empty() || desc->d_initial_edges, "nlo_dis_batch_run: NULL initial edges");
    if (!d_workspace || workspace_bytes < P.total) {
        set_error("nlo_dis_batch_run: workspace %zu bytes < %zu", workspace_bytes, P.total);
        return NLO_ERR_WORKSPACE;
    }
    unsigned char* w = static_cast<unsigned char*>(d_workspace);
    // the records, staged in one host buffer and copied once
    std::vector<unsigned char> host(P.o_part, 0);
    auto put = [&](size_t at, const void* src, size_t bytes) {
        if (bytes) std::memcpy(host.data() + at, src, bytes);
    };
    put(P.o_args, P.args.data(), sizeof(DisArgs) * P.args.size());
    put(P.o_vi, P.vi.data(), sizeof(VItem) * P.vi.size());
    put(P.o_tables, desc->tables, sizeof(nlo_sobol_table) * (size_t)desc->n_tables);
    put(P.o_work, P.work.data(), sizeof(int2) * P.work.size());
    put(P.o_rows, P.rows.data(), sizeof(int2) * P.rows.size());
    put(P.o_gl_work, P.gl_work.data(), sizeof(int4) * P.gl_work.size());
    put(P.o_map, P.map.data(), sizeof(int4) * P.map.size());
    NLO_HIP_CHECK(hipMemcpyAsync(w, host.data(), host.size(), hipMemcpyHostToDevice, stream));
    const DisArgs* args = reinterpret_cast<const DisArgs*>(w + P.o_args);
    BatchDev bd{args,
                reinterpret_cast<const VItem*>(w + P.o_vi),
                reinterpret_cast<VState*>(w + P.o_st),
                reinterpret_cast<const nlo_sobol_table*>(w + P.o_tables),
                reinterpret_cast<double*>(w + P.o_part),
                reinterpret_cast<double*>(w + P.o_grid),
                desc->bins};
    double* gl_out = reinterpret_cast<double*>(w + P.o_gl_out);
    const int nv = (int)P.vi.size();
    if (nv > 0) {
        hipLaunchKernelGGL(vegas_batch_init_kernel, dim3(nv), dim3(256), 0, stream, bd, desc->d_initial_edges);
        NLO_LAUNCH_CHECK();
    }
    for (const auto& L : P.gl_launches) {
        const int4* gw = reinterpret_cast<const int4*>(w + P.o_gl_work);
        if (L.K == NLO_DIS_LO) {
            if (L.prec == NLO_PREC_FP64) launch_gl_batch<double, NLO_DIS_LO>(L, args, gw, gl_out, stream);
            else launch_gl_batch<float, NLO_DIS_LO>(L, args, gw, gl_out, stream);
        } else if (L.K == NLO_DIS_DIP_T0) {
            if (L.prec == NLO_PREC_FP64) launch_gl_batch<double, NLO_DIS_DIP_T0>(L, args, gw, gl_out, stream);
            else launch_gl_batch<float, NLO_DIS_DIP_T0>(L, args, gw, gl_out, stream);
        } else {
            if (L.prec == NLO_PREC_FP64) launch_gl_batch<double, NLO_DIS_DIP_L_CONST>(L, args, gw, gl_out, stream);
            else launch_gl_batch<float, NLO_DIS_DIP_L_CONST>(L, args, gw, gl_out, stream);
        }
        NLO_LAUNCH_CHECK();
    }
    const double eps = desc->precision == NLO_PREC_FP32 ? (double)FLT_EPSILON : DBL_EPSILON;
    const double tiny = desc->precision == NLO_PREC_FP32 ? (double)FLT_MIN : DBL_MIN;
    const int2* work = reinterpret_cast<const int2*>(w + P.o_work);
    const int2* rows = reinterpret_cast<const int2*>(w + P.o_rows);
    hipStream_t cs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int status = NLO_OK;
    if (nv > 0) {
        // one iteration of every integral, captured once and replayed
        NLO_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            status = enqueue_batch_iteration(P, bd, work, rows, eps, tiny, desc->epsrel, desc->epsabs, cs);
            const hipError_t e2 = hipStreamEndCapture(cs, &graph);
            if (status == NLO_OK && e2 != hipSuccess) e = e2;
        }
        if (status == NLO_OK && e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        for (int it = 0; status == NLO_OK && e == hipSuccess && it < P.max_iterations; ++it)
            e = hipGraphLaunch(exec, stream);
        if (status == NLO_OK && e != hipSuccess) {
            set_error("nlo_dis_batch_run: iteration graph: %s", hipGetErrorString(e));
            status = NLO_ERR_HIP;
        }
    }
    if (status == NLO_OK) {
        hipLaunchKernelGGL(batch_results_kernel, dim3((desc->n_items + 255) / 256), dim3(256), 0, stream, bd,
                           reinterpret_cast<const int4*>(w + P.o_map), desc->n_items, gl_out, desc->epsrel,
                           desc->epsabs, d_results);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_error("nlo_dis_batch_run: results: %s", hipGetErrorString(e));
            status = NLO_ERR_HIP;
        }
    }
    // the staging buffer and the graph are released once the stream is done
    const hipError_t es = hipStreamSynchronize(stream);
    if (status == NLO_OK && es != hipSuccess) {
        set_error("nlo_dis_batch_run: %s", hipGetErrorString(es));
        status = NLO_ERR_HIP;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cs) (void)hipStreamDestroy(cs);
    return status;
}

}  // extern "C"
