// Shared host and device helpers of libnlo_gfx950 (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "nlo_gfx950.h"

namespace nlo {

// ------------------------------------------------------------ error state

void set_error(const char* fmt, ...);
void clear_error();

#define NLO_REQUIRE(cond, ...)                      \
    do {                                            \
        if (!(cond)) {                              \
            ::nlo::set_error(__VA_ARGS__);          \
            return NLO_ERR_INVALID;                 \
        }                                           \
    } while (0)

#define NLO_HIP_CHECK(expr)                                                           \
    do {                                                                              \
        hipError_t nlo_status_ = (expr);                                              \
        if (nlo_status_ != hipSuccess) {                                              \
            ::nlo::set_error("%s failed: %s", #expr, hipGetErrorString(nlo_status_)); \
            return NLO_ERR_HIP;                                                       \
        }                                                                             \
    } while (0)

// Checks the launch that was just enqueued.
#define NLO_LAUNCH_CHECK() NLO_HIP_CHECK(hipGetLastError())

constexpr int kWave = 64;

__host__ __device__ inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// ------------------------------------------------- compensated FP64 sums
//
// A running sum is the unevaluated pair s + c.  Each addition is Knuth's
// TwoSum, whose error term is the exact rounding error of s + v (for any
// order and magnitudes of the operands, barring overflow); c gathers those
// errors.  After n additions s + c is within about n u^2 sum |v| of the exact
// sum (u = 2^-53), where plain recursive summation is within n u sum |v|; a
// final rounding s + c adds u |sum|.

__device__ __forceinline__ void comp_add(double& s, double& c, double v) {
    const double t = s + v;
    const double z = t - s;
    c += (s - (t - z)) + (v - z);
    s = t;
}

// (s, c) += (s2, c2).  Symmetric bit for bit: TwoSum's error term is exact
// whatever the order of s and s2, and c + c2 is formed before it is added, so
// both partners of a butterfly hold the same pair.
__device__ __forceinline__ void comp_merge(double& s, double& c, double s2, double c2) {
    const double t = s + s2;
    const double z = t - s;
    const double e = (s - (t - z)) + (s2 - z);
    c = (c + c2) + e;
    s = t;
}

// Fixed-order butterfly of compensated pairs over the 64 lanes; every lane
// receives the total pair.
__device__ __forceinline__ void wave_comp_sum(double& s, double& c) {
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) {
        const double so = __shfl_xor(s, offset, kWave), co = __shfl_xor(c, offset, kWave);
        comp_merge(s, c, so, co);
    }
}

// Compensated sums of NV per-thread pairs (t[k], e[k]) over a block of at most
// 256 threads: a lane butterfly per wave, then the waves in order through
// red (shared, [waves][NV][2]); thread k < NV writes the rounded total of sum
// k to out[k].  Every thread of the block must call it.
template <int NV>
__device__ __forceinline__ void block_comp_sums(double (&t)[NV], double (&e)[NV], double (*red)[NV][2],
                                                double* __restrict__ out) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        wave_comp_sum(t[k], e[k]);
        if (lane == 0) {
            red[wave][k][0] = t[k];
            red[wave][k][1] = e[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = red[0][threadIdx.x][0], c = red[0][threadIdx.x][1];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) comp_merge(s, c, red[w][threadIdx.x][0], red[w][threadIdx.x][1]);
        out[threadIdx.x] = s + c;
    }
}

// --------------------------------------------------------- wave maxima

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) v = fmaxf(v, __shfl_xor(v, offset, kWave));
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) v = fmax(v, __shfl_xor(v, offset, kWave));
    return v;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ------------------------------------------------------------------ Sobol

// Integer state of Sobol point n: shift ^ XOR of direction[b] over the set
// bits b of gray(n) = n ^ (n >> 1) (gray-code order of SobolEngine).
__device__ __forceinline__ uint32_t sobol_state(const uint32_t* direction, uint32_t shift,
                                                uint64_t n) {
    uint64_t g = n ^ (n >> 1);
    uint32_t q = shift;
    int b = 0;
    while (g) {
        if (g & 1u) q ^= direction[b];
        g >>= 1;
        ++b;
    }
    return q;
}

}  // namespace nlo
