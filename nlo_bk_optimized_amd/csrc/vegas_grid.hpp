// VEGAS grid refinement for one dimension of one grid, shared by the
// standalone refinement kernel and the batched DIS integrator: the
// reference's operation order with no contraction, so both give the same
// bits.
#pragma once

#include <hip/hip_runtime.h>

namespace nlo {

// Importance of bin i: sum |w| / max(count, 1).
__device__ __forceinline__ double vegas_importance(const double* h, const double* c, int i) {
    return h[i] / fmax(c[i], 1.0);
}

// Smoothed importance of bin i (S/numerics/integration.py:828-833): 3-point
// mean inside, 2-point mean at the ends, plus eps.
__device__ __forceinline__ double vegas_smoothed(const double* h, const double* c, int bins, int i,
                                                 double eps) {
#pragma clang fp contract(off)
    double v;
    if (i == 0) {
        v = (vegas_importance(h, c, 0) + vegas_importance(h, c, 1)) / 2.0;
    } else if (i == bins - 1) {
        v = (vegas_importance(h, c, bins - 2) + vegas_importance(h, c, bins - 1)) / 2.0;
    } else {
        v = (vegas_importance(h, c, i - 1) + vegas_importance(h, c, i) + vegas_importance(h, c, i + 1)) /
            3.0;
    }
    return v + eps;
}

// _adapt_vegas_edges (S/numerics/integration.py:835-866) of one dimension:
// edges e[bins + 1], histogram hh[bins] and counts cc[bins] in, refined
// edges o[bins + 1] out (o must not alias e).  The linspace targets
// increase, so searchsorted(cumulative, target, right) - 1 is found by a
// forward sweep over the cumulative sums: no scratch memory.
// sm(i) gives the smoothed importance of bin i.
template <typename Smoothed>
__device__ inline void vegas_adapt_row_with(const double* e, const Smoothed& sm, int bins, double* o) {
#pragma clang fp contract(off)
    double total = 0.0;
    for (int i = 0; i < bins; ++i) total = total + sm(i);
    const double step = total / (double)bins;
    // j: largest index with cumulative[j] <= target; cum_j = cumulative[j];
    // sm_j = smoothed[j] (cumulative[j + 1] = cum_j + sm_j).
    int j = 0;
    double cum_j = 0.0;
    double sm_j = sm(0);
    double cum_prev = 0.0, sm_prev = sm_j;  // values at j - 1 (valid when j == bins)
    for (int t = 0; t <= bins; ++t) {
        const double target =
            t < (bins + 1) / 2 ? (double)t * step : fma(-step, (double)(bins - t), total);
        while (j < bins) {
            const double cum_next = cum_j + sm_j;
            if (!(target >= cum_next)) break;
            cum_prev = cum_j;
            sm_prev = sm_j;
            cum_j = cum_next;
            ++j;
            if (j < bins) sm_j = sm(j);
        }
        const int b = j < bins ? j : bins - 1;
        const double cb = j < bins ? cum_j : cum_prev;
        const double sb = j < bins ? sm_j : sm_prev;
        const double frac = (target - cb) / sb;
        o[t] = e[b] + frac * (e[b + 1] - e[b]);
    }
    o[0] = 0.0;
    o[bins] = 1.0;
}

__device__ inline void vegas_adapt_row(const double* e, const double* hh, const double* cc, int bins, double eps,
                                       double* o) {
    vegas_adapt_row_with(e, [&](int i) { return vegas_smoothed(hh, cc, bins, i, eps); }, bins, o);
}

}  // namespace nlo
