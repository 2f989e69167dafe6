// J1 and I1 in FP64 for the double-log resummation of K1
// (S/bk_kernels/k1.py:108-132 uses torch.special.bessel_j1 / modified_bessel_i1).
// Both return f(x) / (x / 2) so that the caller's J1(2a)/a and I1(2a)/a are
// direct.
//   J1, x < 128: the piecewise Chebyshev table of bessel_j1_i1_tables.inc
//       (pieces of width 4, degree 16; the caller passes the table, staged in
//       LDS); beyond: Miller's backward recurrence J_{k-1} = (2k / x) J_k -
//       J_{k+1} from k = N ~ 1.25 x + 32, normalised by J_0 + 2 sum_k J_2k = 1
//       (Abramowitz & Stegun 9.1.27 and 9.1.46).
//   I1, x < 8: the ascending series sum (x^2/4)^k / (k! (k+1)!), all terms
//       positive, to k = 20 (the tail is below 4e-18 relative at x = 8) by
//       Horner's rule; x >= 8: e^x / sqrt(x) times the Chebyshev series of
//       e^-x sqrt(x) I1(x) in 16/x - 1.
#pragma once

#include <hip/hip_runtime.h>

namespace nlo {

#include "bessel_j1_i1_tables.inc"

// 1 / (k! (k+1)!), k = 0..20
__device__ constexpr double kI1Series[21] = {
    1.0, 0.5, 0.08333333333333333,
    0.006944444444444444, 0.00034722222222222224, 1.1574074074074073e-05,
    2.755731922398589e-07, 4.920949861426052e-09, 6.834652585313961e-11,
    7.594058428126623e-13, 6.903689480115112e-15, 5.230067787965994e-17,
    3.352607556388458e-19, 1.842092063949702e-21, 8.771866971189057e-24,
    3.654944571328774e-26, 1.3437296218120491e-28, 4.391273273895586e-31,
    1.2839980333028028e-33, 3.3789421929021126e-36, 8.045100459290744e-39,
};

// J1(2a)/a for s = -a^2 <= 0 and I1(2a)/a for s = a^2 >= 0: both are the one
// power series sum_k s^k / (k! (k+1)!) (A&S 9.1.10, 9.6.10), here to k = 20 by
// Horner's rule.  For |s| <= 16 (2a <= 8) the tail is below 4e-18 of the
// I1 sum; on the J1 side the alternating terms (at most ~23) leave an
// absolute rounding error below 1e-14.
__host__ __device__ inline __attribute__((noinline)) double j1_i1_series(double s) {
    double sum = kI1Series[20];
#pragma unroll
    for (int k = 19; k >= 0; --k) sum = fma(sum, s, kI1Series[k]);
    return sum;
}

__host__ __device__ inline __attribute__((noinline)) double i1_over_half_x(double x) {
    if (x < 8.0) {
        const double t = 0.25 * x * x;
        double sum = kI1Series[20];
#pragma unroll
        for (int k = 19; k >= 0; --k) sum = fma(sum, t, kI1Series[k]);
        return sum;
    }
    const double s = 16.0 / x - 1.0, s2 = s + s;
    double b0 = 0.0, b1 = 0.0;
    for (int m = kI1Deg; m >= 1; --m) {
        const double t = s2 * b0 - b1 + kI1Cheb[m];
        b1 = b0;
        b0 = t;
    }
    const double g = s * b0 - b1 + kI1Cheb[0];
    return g * exp(x) / sqrt(x) / (0.5 * x);
}

__host__ __device__ inline __attribute__((noinline)) double j1_miller_over_half_x(double x) {
    const int N = 2 * ((int)(0.625 * x) + 16);  // even, > 1.25 x + 30
    const double two_over_x = 2.0 / x;
    double jp = 0.0, j = 1e-30;  // J_{k+1}, J_k (unnormalised)
    double norm = 0.0, j1 = 0.0;
    for (int k = N; k >= 1; --k) {
        const double jm = (double)k * two_over_x * j - jp;  // J_{k-1}
        jp = j;
        j = jm;
        if (k == 2) j1 = j;                                    // J_1, produced by the step from k = 2
        if (((k - 1) & 1) == 0 && k - 1 > 0) norm += 2.0 * jm;  // even order >= 2
        if (fabs(j) > 1e250) {  // rescale to stay in range
            j *= 1e-250;
            jp *= 1e-250;
            norm *= 1e-250;
            j1 *= 1e-250;
        }
    }
    norm += j;  // J_0
    return (j1 / norm) / (0.5 * x);
}

// table: kJ1Cheb (kJ1Pieces x (kJ1Deg + 1) doubles), e.g. staged in LDS.
__host__ __device__ inline double j1_over_half_x(double x, const double* table) {
    if (!(x < 4.0 * kJ1Pieces)) return j1_miller_over_half_x(x);
    const int k = (int)(0.25 * x);
    const double s = 0.5 * (x - 4.0 * k) - 1.0, s2 = s + s;
    const double* c = table + k * (kJ1Deg + 1);
    double b0 = 0.0, b1 = 0.0;
    for (int m = kJ1Deg; m >= 1; --m) {
        const double t = s2 * b0 - b1 + c[m];
        b1 = b0;
        b0 = t;
    }
    return s * b0 - b1 + c[0];
}

}  // namespace nlo
