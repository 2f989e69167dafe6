// Scrambled Sobol generation, VEGAS map, deterministic histogram and grid
// refinement for gfx950.  HBM-streaming integer/byte work: no MFMA, coalesced
// wave64 stores, one point per lane.
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "vegas_grid.hpp"

// The VEGAS map and refinement reproduce the reference's FP64 operation
// sequence bit for bit, so no multiply-add in this file may be fused.
#pragma clang fp contract(off)

namespace nlo {

namespace {

// ------------------------------------------------------------------ Sobol

// One wave covers 64 consecutive sample indices n = hi * 64 + lane per group.
// Then gray(n) = (gray(hi) << 6) ^ (lane ^ (lane >> 1) ^ ((hi & 1) << 5)), so
// the state splits into a wave-uniform part (direction bits 6..29, scalar
// ALU) and a per-lane part (bits 0..5) that takes one of two values by the
// parity of hi.  Each wave walks kSobolGroups consecutive groups: the
// uniform part is built once from gray(hi) and then updated by one XOR per
// group (gray(hi) ^ gray(hi - 1) = the lowest set bit of hi), so the work per
// point is a select, the conversion and the store, and a wave writes one
// contiguous span.
constexpr int kSobolGroups = 16;

template <int DIM>
__global__ void __launch_bounds__(256) sobol_draw_kernel(nlo_sobol_table t, int64_t first,
                                                          int64_t n, int32_t kind, int vec, void* out) {
    const int lane = threadIdx.x & 63;
    const int64_t wave_global = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    const int64_t hi_begin = first >> 6;
    const int64_t hi_end = (first + n + 63) >> 6;
    // per-lane low part for even and odd hi
    const uint32_t glo = (uint32_t)(lane ^ (lane >> 1));
    uint32_t lo0[DIM], lo1[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        uint32_t v = 0u;
#pragma unroll
        for (int b = 0; b < 6; ++b) v ^= ((glo >> b) & 1u) ? t.direction[d][b] : 0u;
        lo0[d] = v;
        lo1[d] = v ^ t.direction[d][5];
    }
    for (int64_t base = hi_begin + wave_global * kSobolGroups; base < hi_end; base += waves * kSobolGroups) {
        // wave-uniform high part: shift ^ direction bits 6..29 of gray(base)
        uint32_t qh[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) qh[d] = t.shift[d];
        {
            uint64_t g = (uint64_t)base ^ ((uint64_t)base >> 1);
            for (int b = 6; g && b < NLO_SOBOL_BITS; ++b, g >>= 1)
                if (g & 1u) {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) qh[d] ^= t.direction[d][b];
                }
        }
        const int64_t stop = base + kSobolGroups < hi_end ? base + kSobolGroups : hi_end;
        for (int64_t hi = base; hi < stop; ++hi) {
            if (hi > base) {
                const int b = 6 + __builtin_ctzll((unsigned long long)hi);
                if (b < NLO_SOBOL_BITS) {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) qh[d] ^= t.direction[d][b];
                }
            }
            const uint64_t idx = ((uint64_t)hi << 6) | (uint64_t)lane;
            const int64_t row = (int64_t)idx - first;
            if (!(row >= 0 && row < n)) continue;
            const bool odd = (hi & 1) != 0;
            uint32_t q[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) q[d] = qh[d] ^ (odd ? lo1[d] : lo0[d]);
            const size_t o = (size_t)row * DIM;
            if constexpr (DIM == 4) {
                if (vec) {
                    // 16-byte aligned output: one 16- or 32-byte store per lane, so
                    // the wave writes one contiguous span
                    if (kind == NLO_SOBOL_U32) {
                        reinterpret_cast<uint4*>(static_cast<uint32_t*>(out) + o)[0] = make_uint4(q[0], q[1], q[2], q[3]);
                    } else if (kind == NLO_SOBOL_F32) {
                        reinterpret_cast<float4*>(static_cast<float*>(out) + o)[0] =
                            make_float4((float)q[0] * 0x1p-30f, (float)q[1] * 0x1p-30f, (float)q[2] * 0x1p-30f,
                                        (float)q[3] * 0x1p-30f);
                    } else {
                        double v[4];
                        for (int d = 0; d < 4; ++d) v[d] = (double)(idx == 0 ? t.first[d] : q[d]) * 0x1p-30;
                        double2* po = reinterpret_cast<double2*>(static_cast<double*>(out) + o);
                        po[0] = make_double2(v[0], v[1]);
                        po[1] = make_double2(v[2], v[3]);
                    }
                    continue;
                }
            }
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                if (kind == NLO_SOBOL_U32) {
                    static_cast<uint32_t*>(out)[o + d] = q[d];
                } else if (kind == NLO_SOBOL_F32) {
                    static_cast<float*>(out)[o + d] = (float)q[d] * 0x1p-30f;
                } else {
                    const uint32_t qq = idx == 0 ? t.first[d] : q[d];
                    static_cast<double*>(out)[o + d] = (double)qq * 0x1p-30;
                }
            }
        }
    }
}

// ------------------------------------------------------------------ VEGAS

// map_importance_grid (S/dis/fixed_sobol.py:100-113) in IEEE order (no
// contraction in this file), so bins and coordinates equal the FP64
// reference bit for bit.
__global__ void __launch_bounds__(256) vegas_map_kernel(int dim, int bins, int64_t n,
                                                         const double* __restrict__ u,
                                                         const double* __restrict__ edges,
                                                         const double* __restrict__ lo,
                                                         const double* __restrict__ width,
                                                         double* __restrict__ x,
                                                         double* __restrict__ jac,
                                                         uint8_t* __restrict__ bin, int vec) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (vec) {
            // dim = 4 with 16-byte aligned u and x and 4-byte aligned bins:
            // vector loads and stores, one contiguous span per wave; the same
            // arithmetic in the same order as below
            const double2* u2 = reinterpret_cast<const double2*>(u + i * 4);
            const double2 ua = u2[0], ub = u2[1];
            const double uv[4] = {ua.x, ua.y, ub.x, ub.y};
            double xv[4];
            uint32_t packed = 0u;
            double j = 1.0, vol = 1.0;
            for (int d = 0; d < 4; ++d) {
                const double scaled = uv[d] * (double)bins;
                int b = (int)floor(scaled);
                b = b < 0 ? 0 : (b > bins - 1 ? bins - 1 : b);
                const double frac = scaled - (double)b;
                const double left = edges[d * (bins + 1) + b];
                const double right = edges[d * (bins + 1) + b + 1];
                const double w = right - left;
                xv[d] = lo[d] + width[d] * (left + w * frac);
                j = j * ((double)bins * w);
                vol = vol * width[d];
                packed |= (uint32_t)b << (8 * d);
            }
            double2* x2 = reinterpret_cast<double2*>(x + i * 4);
            x2[0] = make_double2(xv[0], xv[1]);
            x2[1] = make_double2(xv[2], xv[3]);
            if (bin) reinterpret_cast<uint32_t*>(bin)[i] = packed;
            jac[i] = vol * j;
            continue;
        }
        double j = 1.0, vol = 1.0;
        for (int d = 0; d < dim; ++d) {
            const double scaled = u[i * dim + d] * (double)bins;
            int b = (int)floor(scaled);
            b = b < 0 ? 0 : (b > bins - 1 ? bins - 1 : b);
            const double frac = scaled - (double)b;
            const double left = edges[d * (bins + 1) + b];
            const double right = edges[d * (bins + 1) + b + 1];
            const double w = right - left;
            x[i * dim + d] = lo[d] + width[d] * (left + w * frac);
            j = j * ((double)bins * w);
            vol = vol * width[d];
            if (bin) bin[i * dim + d] = (uint8_t)b;
        }
        jac[i] = vol * j;
    }
}

constexpr int kHistThreads = 256;
constexpr int kHistChunk = 4096;  // samples per block: 2^12, sums stay below 2^63

// Per-block histogram.  Each block first finds the largest |w| of its chunk,
// then adds every |w| as a 64-bit fixed-point integer scaled so that the
// chunk total cannot exceed 2^63: integer addition is exact and
// order-independent, so LDS atomics give the same bits every run.
// Per block of kHistChunk samples: |w| scaled to 64-bit fixed point by the
// block maximum, then integer sums per (dim, bin) in LDS (exact and
// order-free).  With `waves` > 1 every wave fills its own copy (fewer atomic
// collisions on the same bin) and the copies are added in wave order.
__global__ void __launch_bounds__(kHistThreads) vegas_hist_block_kernel(
    int dim, int bins, int64_t n, int waves, const uint8_t* __restrict__ bin, const double* __restrict__ absw,
    double* __restrict__ part_hist, double* __restrict__ part_count) {
    extern __shared__ unsigned long long smem[];
    const int cells = dim * bins;
    unsigned long long* h = smem;                                               // waves x cells
    unsigned int* c = reinterpret_cast<unsigned int*>(smem + (size_t)waves * cells);  // waves x cells
    __shared__ double wave_max_v[kHistThreads / 64];
    const int64_t begin = (int64_t)blockIdx.x * kHistChunk;
    const int64_t end = begin + kHistChunk < n ? begin + kHistChunk : n;
    for (int k = threadIdx.x; k < waves * cells; k += blockDim.x) {
        h[k] = 0ull;
        c[k] = 0u;
    }
    double m = 0.0;
    for (int64_t i = begin + threadIdx.x; i < end; i += blockDim.x) m = fmax(m, fabs(absw[i]));
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) wave_max_v[threadIdx.x >> 6] = m;
    __syncthreads();
    m = 0.0;
    for (int k = 0; k < kHistThreads / 64; ++k) m = fmax(m, wave_max_v[k]);
    const int e = m > 0.0 ? ilogb(m) : 0;
    const int s = 50 - e;  // m * 2^s < 2^51; 4096 terms < 2^63
    const int copy = waves > 1 ? (int)(threadIdx.x >> 6) : 0;
    unsigned long long* hw = h + (size_t)copy * cells;
    unsigned int* cw = c + (size_t)copy * cells;
    for (int64_t i = begin + threadIdx.x; i < end; i += blockDim.x) {
        const unsigned long long v = (unsigned long long)rint(ldexp(fabs(absw[i]), s));
        for (int d = 0; d < dim; ++d) {
            const int b = min((int)bin[i * dim + d], bins - 1);
            atomicAdd(&hw[d * bins + b], v);
            atomicAdd(&cw[d * bins + b], 1u);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < cells; k += blockDim.x) {
        unsigned long long th = 0ull;
        unsigned int tc = 0u;
        for (int w = 0; w < waves; ++w) {
            th += h[(size_t)w * cells + k];
            tc += c[(size_t)w * cells + k];
        }
        part_hist[(size_t)blockIdx.x * cells + k] = ldexp((double)th, -s);
        part_count[(size_t)blockIdx.x * cells + k] = (double)tc;
    }
}

// Fixed-order sums of the per-block partials in two levels: group g of
// consecutive blocks per (dim, bin) (lanes on consecutive cells, coalesced),
// then the groups in order.
constexpr int kHistGroups = 64;

__global__ void __launch_bounds__(64) vegas_hist_group_kernel(int values, int blocks, int per_group,
                                                             const double* __restrict__ ph,
                                                             const double* __restrict__ pc,
                                                             double* __restrict__ gh, double* __restrict__ gc) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    const int g = blockIdx.y;
    if (k >= values) return;
    const int b0 = g * per_group, b1 = min(b0 + per_group, blocks);
    double sh = 0.0, sc = 0.0;
    for (int b = b0; b < b1; ++b) {
        sh += ph[(size_t)b * values + k];
        sc += pc[(size_t)b * values + k];
    }
    gh[(size_t)g * values + k] = sh;
    gc[(size_t)g * values + k] = sc;
}

__global__ void vegas_hist_reduce_kernel(int values, int groups, const double* __restrict__ gh,
                                         const double* __restrict__ gc, double* __restrict__ hist,
                                         double* __restrict__ count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= values) return;
    double sh = 0.0, sc = 0.0;
    for (int g = 0; g < groups; ++g) {
        sh += gh[(size_t)g * values + k];
        sc += gc[(size_t)g * values + k];
    }
    hist[k] = sh;
    count[k] = sc;
}

// _adapt_vegas_edges (S/numerics/integration.py:835-866) for one (row, dim)
// per thread (vegas_adapt_row, vegas_grid.hpp).
__global__ void vegas_adapt_kernel(int rows, int dim, int bins, const double* __restrict__ edges,
                                   const double* __restrict__ hist,
                                   const double* __restrict__ count, double eps,
                                   double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= rows * dim) return;
    vegas_adapt_row(edges + (size_t)k * (bins + 1), hist + (size_t)k * bins, count + (size_t)k * bins, bins,
                    eps, out + (size_t)k * (bins + 1));
}

}  // namespace

// ------------------------------------------------------------ error state

namespace {
thread_local std::string g_error;
}

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}

void clear_error() { g_error.clear(); }

const char* last_error() { return g_error.c_str(); }

int launch_vegas_adapt(int rows, int dim, int bins, const double* edges, const double* hist,
                       const double* count, double eps, double* out, hipStream_t stream) {
    const int items = rows * dim;
    hipLaunchKernelGGL(vegas_adapt_kernel, dim3((items + 63) / 64), dim3(64), 0, stream, rows, dim,
                       bins, edges, hist, count, eps, out);
    NLO_LAUNCH_CHECK();
    return NLO_OK;
}

}  // namespace nlo

// ------------------------------------------------------------------- C ABI

extern "C" {

int nlo_api_version(void) { return NLO_API_VERSION; }

const char* nlo_last_error(void) { return nlo::last_error(); }

int nlo_gauss_legendre(int32_t n, double* x, double* w) {
    NLO_REQUIRE(n >= 1 && n <= 512, "nlo_gauss_legendre: n=%d outside [1, 512]", n);
    NLO_REQUIRE(x && w, "nlo_gauss_legendre: null output");
    // Newton iteration on P_n from the Chebyshev-like initial guess
    // cos(pi (i + 0.75) / (n + 0.5)), three-term recurrence in long double.
    for (int i = 0; i < n; ++i) {
        long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
        long double dp = 1.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) p0 = 1.0L;
            dp = n * (z * p1 - p0) / (z * z - 1.0L);
            const long double dz = p1 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-30L) break;
        }
        long double p0 = 1.0L, p1 = z;
        for (int k = 2; k <= n; ++k) {
            const long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1.0L);
        // ascending order: node i (largest first from the guess) goes to n-1-i
        x[n - 1 - i] = (double)z;
        w[n - 1 - i] = (double)(2.0L / ((1.0L - z * z) * dp * dp));
    }
    nlo::clear_error();
    return NLO_OK;
}

int nlo_sobol_draw(const nlo_sobol_table* table, int64_t first_index, int64_t n, int32_t out_kind,
                   void* d_out, hipStream_t stream) {
    NLO_REQUIRE(table, "nlo_sobol_draw: null table");
    NLO_REQUIRE(table->dim >= 1 && table->dim <= NLO_SOBOL_MAXDIM,
                "nlo_sobol_draw: dim=%d outside [1, %d]", table->dim, NLO_SOBOL_MAXDIM);
    NLO_REQUIRE(first_index >= 0 && n >= 0 && first_index + n <= (int64_t(1) << 30),
                "nlo_sobol_draw: indices [%lld, %lld) outside [0, 2^30)", (long long)first_index,
                (long long)(first_index + n));
    NLO_REQUIRE(out_kind == NLO_SOBOL_U32 || out_kind == NLO_SOBOL_F32 || out_kind == NLO_SOBOL_F64,
                "nlo_sobol_draw: unknown out_kind %d", out_kind);
    NLO_REQUIRE(d_out || n == 0, "nlo_sobol_draw: null output");
    if (n == 0) return NLO_OK;
    const int64_t groups = ((first_index + n + 63) >> 6) - (first_index >> 6);
    const int64_t waves_per_block = 4;
    const int64_t chunks = (groups + nlo::kSobolGroups - 1) / nlo::kSobolGroups;
    int64_t blocks = (chunks + waves_per_block - 1) / waves_per_block;
    if (blocks > 65536) blocks = 65536;
    const int vec = table->dim == 4 && (reinterpret_cast<uintptr_t>(d_out) & 15) == 0;
#define NLO_SOBOL_DIM(D)                                                                                \
    case D:                                                                                             \
        hipLaunchKernelGGL(nlo::sobol_draw_kernel<D>, dim3((unsigned)blocks), dim3(256), 0, stream, *table, \
                           first_index, n, out_kind, vec, d_out);                                        \
        break;
    switch (table->dim) {
        NLO_SOBOL_DIM(1)
        NLO_SOBOL_DIM(2)
        NLO_SOBOL_DIM(3)
        NLO_SOBOL_DIM(4)
        NLO_SOBOL_DIM(5)
        NLO_SOBOL_DIM(6)
        NLO_SOBOL_DIM(7)
        NLO_SOBOL_DIM(8)
        default:
            break;
    }
#undef NLO_SOBOL_DIM
    NLO_LAUNCH_CHECK();
    nlo::clear_error();
    return NLO_OK;
}

int nlo_vegas_map(int32_t dim, int32_t bins, int64_t n, const double* d_u, const double* d_edges,
                  const double* d_lo, const double* d_width, double* d_x, double* d_jac,
                  uint8_t* d_bin, hipStream_t stream) {
    NLO_REQUIRE(dim >= 1 && dim <= 16, "nlo_vegas_map: dim=%d outside [1, 16]", dim);
    NLO_REQUIRE(bins >= 2 && bins <= 255, "nlo_vegas_map: bins=%d outside [2, 255]", bins);
    NLO_REQUIRE(n >= 0, "nlo_vegas_map: negative n");
    NLO_REQUIRE(n == 0 || (d_u && d_edges && d_lo && d_width && d_x && d_jac),
                "nlo_vegas_map: null pointer");
    if (n == 0) return NLO_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    const int vec = dim == 4 && (reinterpret_cast<uintptr_t>(d_u) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(d_x) & 15) == 0 && (reinterpret_cast<uintptr_t>(d_bin) & 3) == 0;
    hipLaunchKernelGGL(nlo::vegas_map_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, dim, bins,
                       n, d_u, d_edges, d_lo, d_width, d_x, d_jac, d_bin, vec);
    NLO_LAUNCH_CHECK();
    nlo::clear_error();
    return NLO_OK;
}

namespace {

struct HistPlan {
    int blocks, per_group, groups;
    size_t part, group, total;
};

HistPlan hist_plan(int dim, int bins, int64_t n) {
    HistPlan p{};
    p.blocks = (int)((n + nlo::kHistChunk - 1) / nlo::kHistChunk);
    if (p.blocks < 1) p.blocks = 1;
    p.per_group = (p.blocks + nlo::kHistGroups - 1) / nlo::kHistGroups;
    p.groups = (p.blocks + p.per_group - 1) / p.per_group;
    p.part = (size_t)p.blocks * dim * bins * 2 * sizeof(double);
    p.group = (size_t)p.groups * dim * bins * 2 * sizeof(double);
    p.total = p.part + p.group;
    return p;
}

}  // namespace

size_t nlo_vegas_histogram_workspace(int32_t dim, int32_t bins, int64_t n) {
    if (dim < 1 || bins < 1 || n < 0) return 0;
    return hist_plan(dim, bins, n).total;
}

int nlo_vegas_histogram(int32_t dim, int32_t bins, int64_t n, const uint8_t* d_bin,
                        const double* d_absw, double* d_hist, double* d_count, void* d_workspace,
                        size_t workspace_bytes, hipStream_t stream) {
    NLO_REQUIRE(dim >= 1 && dim <= 16, "nlo_vegas_histogram: dim=%d outside [1, 16]", dim);
    NLO_REQUIRE(bins >= 2 && bins <= 255, "nlo_vegas_histogram: bins=%d outside [2, 255]", bins);
    NLO_REQUIRE(n >= 1, "nlo_vegas_histogram: n must be positive");
    NLO_REQUIRE(d_bin && d_absw && d_hist && d_count, "nlo_vegas_histogram: null pointer");
    const size_t need = nlo_vegas_histogram_workspace(dim, bins, n);
    if (workspace_bytes < need || !d_workspace) {
        nlo::set_error("nlo_vegas_histogram: workspace %zu bytes < %zu", workspace_bytes, need);
        return NLO_ERR_WORKSPACE;
    }
    const HistPlan pl = hist_plan(dim, bins, n);
    const int blocks = pl.blocks;
    const int values = dim * bins;
    double* ph = static_cast<double*>(d_workspace);
    double* pc = ph + (size_t)blocks * values;
    double* gh = pc + (size_t)blocks * values;
    double* gc = gh + (size_t)pl.groups * values;
    // one sub-histogram per wave while they fit in 48 KiB of LDS
    const size_t cell_bytes = sizeof(unsigned long long) + sizeof(unsigned int);
    const int waves = (size_t)(nlo::kHistThreads / 64) * values * cell_bytes <= 48 * 1024 ? nlo::kHistThreads / 64 : 1;
    const size_t smem = (size_t)waves * values * cell_bytes + 8;
    hipLaunchKernelGGL(nlo::vegas_hist_block_kernel, dim3(blocks), dim3(nlo::kHistThreads), smem, stream,
                       dim, bins, n, waves, d_bin, d_absw, ph, pc);
    NLO_LAUNCH_CHECK();
    hipLaunchKernelGGL(nlo::vegas_hist_group_kernel, dim3((values + 63) / 64, pl.groups), dim3(64), 0, stream,
                       values, blocks, pl.per_group, ph, pc, gh, gc);
    NLO_LAUNCH_CHECK();
    hipLaunchKernelGGL(nlo::vegas_hist_reduce_kernel, dim3((values + 255) / 256), dim3(256), 0, stream,
                       values, pl.groups, gh, gc, d_hist, d_count);
    NLO_LAUNCH_CHECK();
    nlo::clear_error();
    return NLO_OK;
}

int nlo_vegas_adapt(int32_t rows, int32_t dim, int32_t bins, const double* d_edges,
                    const double* d_hist, const double* d_count, double eps, double* d_edges_out,
                    hipStream_t stream) {
    NLO_REQUIRE(rows >= 1 && dim >= 1 && bins >= 2 && bins <= 4096,
                "nlo_vegas_adapt: rows=%d dim=%d bins=%d invalid", rows, dim, bins);
    NLO_REQUIRE(d_edges && d_hist && d_count && d_edges_out, "nlo_vegas_adapt: null pointer");
    NLO_REQUIRE(d_edges != d_edges_out, "nlo_vegas_adapt: in-place refinement is not supported");
    const int rc = nlo::launch_vegas_adapt(rows, dim, bins, d_edges, d_hist, d_count, eps,
                                           d_edges_out, stream);
    if (rc == NLO_OK) nlo::clear_error();
    return rc;
}

}  // extern "C"
