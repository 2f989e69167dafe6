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

// --------------------------------------------------------- wave reductions

// Fixed-order butterfly sum over the 64 lanes of a wavefront; every lane
// receives the total.  The order of additions depends only on lane ids, so
// the result is reproducible run to run.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, kWave);
    return v;
}

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
