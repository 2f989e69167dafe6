// The NLO DIS integrands on the unit cube, shared by the per-point, VEGAS,
// Gauss-Legendre and batch kernels and by the fixed-Sobol estimator.
// Reference prefixes: S/ = torch_nlo_clean/src/nlo_torch/, K/ = S/custom_kernels/.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "dis_device.hpp"
#include "nlo_gfx950.h"

namespace nlo {
namespace dis {

constexpr double kZmin = 1e-6;       // ZMIN (S/dis/observables.py:46)
constexpr double kCF = 4.0 / 3.0;    // (Nc^2 - 1) / (2 Nc), Nc = 3
constexpr double kNc = 3.0;
constexpr double kQ0sqr = 1.0;       // Q0sqr (S/dis/config.py:13)
constexpr int kMaxNested = 128;

constexpr int kDims[NLO_DIS_KINDS] = {2, 2, 3, 4, 2, 3, 4, 5, 6, 7, 5, 5};

// Gauss-Legendre node i (ascending) and weight of the n-point rule on [-1, 1]:
// Newton on the three-term recurrence from the Chebyshev-like initial guess.
__device__ inline void gl_node(int n, int i, double& x, double& w) {
    double z = -cos(kPiD * (i + 0.75) / (n + 0.5));
    double p0 = 1.0, p1 = z, dp = 1.0;
    for (int it = 0; it < 100; ++it) {
        p0 = 1.0;
        p1 = z;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        const double dz = p1 / dp;
        z -= dz;
        if (fabs(dz) <= 1e-16) break;
    }
    p0 = 1.0;
    p1 = z;
    for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    x = z;
    w = 2.0 / ((1.0 - z * z) * dp * dp);
}

// The nested I2/I3 inner rule in LDS: y = t^3, t = (x + 1)/2, weight
// w/2 * 3 t^2 (K/dis/i2_i3/extension.py:21-36).
struct RuleView {
    int n;
    const double* y;
    const double* w;
};

__device__ inline RuleView stage_nested_rule(int n, double* y, double* w) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double x, wt;
        gl_node(n, i, x, wt);
        const double t = 0.5 * (x + 1.0);
        y[i] = t * t * t;
        w[i] = 0.5 * wt * 3.0 * t * t;
    }
    __syncthreads();
    return {n, y, w};
}

// Everything an integrand reads, by value in the kernel arguments.
struct DisArgs {
    double Q2, xbj;
    int transverse, nq;
    double mass[NLO_DIS_MAX_QUARKS], charge[NLO_DIS_MAX_QUARKS];
    double maxr;
    int finite_nc, parent_coupling;
    DisCoupling coup;
    double y_lo;    // LO rapidity max(ln(X0 / xbj), 0)
    double y_dip;   // dipole-sector rapidity of quark 0 (S/dis/observables.py:55-68)
    double z2min;   // Q0^2 / W^2 (S/dis/observables.py:71-73)
    double min_r, max_r;  // the dipole's table range (inside_table, :145-147)
    DisDipole dip;
    int nested_points;
};

// ------------------------------------------------------------- integrands

template <typename T>
__device__ __forceinline__ T amp(const DisArgs& a, T r, T Y) {
    return dipole_amplitude<T, false>(r, Y, a.dip);
}

// S/dis/observables.py:86-104
template <typename T>
__device__ __forceinline__ T tripole_amp(const DisArgs& a, T x01, T x02, T x21, T Y) {
    const T S01 = T(1) - amp(a, x01, Y), S02 = T(1) - amp(a, x02, Y), S12 = T(1) - amp(a, x21, Y);
    if (a.finite_nc) return T(1) - T(kNc / (2.0 * kCF)) * (S02 * S12 - S01 / T(kNc * kNc));
    return T(1) - S02 * S12;
}

// Phase-space point of the qqbar-g integrands (S/dis/observables.py:186-212,
// 279-297): momentum fractions, transverse sizes, Jacobian, rapidity, the
// tripole amplitude and the coupling factor.  Requires z2min <= 1.
template <typename T>
struct QgPoint {
    T z1, z2, x01, x02, x01sq, x02sq, x21sq, jac, Y, Ye, mf, m2, N012, af;
};

template <typename T>
__device__ __forceinline__ QgPoint<T> qg_point(const DisArgs& a, const double* u) {
    QgPoint<T> p;
    const T Q2 = (T)a.Q2;
    const T z2min = (T)a.z2min;
    p.z1 = (T(1) - z2min) * (T)u[0];
    p.z2 = (T(1) - p.z1 - z2min) * (T)u[1] + z2min;
    p.x01 = (T)(a.maxr * u[2]);
    p.x02 = (T)(a.maxr * u[3]);
    const T phi = (T)(2.0 * kPiD * u[4]);
    p.x01sq = p.x01 * p.x01;
    p.x02sq = p.x02 * p.x02;
    p.x21sq = fmax(p.x01sq + p.x02sq - T(2) * p.x01 * p.x02 * cos(phi), T(0));
    const T x21 = dis_sqrt(p.x21sq);
    const T maxr = (T)a.maxr;
    p.jac = (T(1) - z2min) * (T(1) - p.z1 - z2min) * p.x01 * p.x02 * maxr * maxr * T(2.0 * kPiD);
    p.Y = log((Q2 / (T)a.xbj) * p.z2 / T(kQ0sqr));
    p.Ye = fmax(p.Y, T(0));
    p.mf = (T)a.mass[0];
    p.m2 = p.mf * p.mf;
    p.N012 = tripole_amp<T>(a, p.x01, p.x02, x21, p.Ye);
    const T ra = a.parent_coupling ? p.x01 : fmin(p.x01, fmin(p.x02, x21));
    p.af = dis_alpha<T>(ra, a.coup) * T(kCF / kPiD);
    return p;
}

// res x Jacobian x coupling / z2, zero below Y = 0 or when not finite.
template <typename T>
__device__ __forceinline__ T qg_finish(const QgPoint<T>& p, T res) {
    res = res * p.jac * p.af / p.z2;
    return (p.Y >= T(0) && isfinite(res)) ? res : T(0);
}

// The G-value moments integrated over y_t on the nested inner rule.
template <typename T>
__device__ __forceinline__ TMoments<T> nested_moments(const Tripole<T>& v, T m2, const RuleView& rule) {
    TMoments<T> acc{T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
    for (int i = 0; i < rule.n; ++i) t_moments_axpy(acc, (T)rule.w[i], t_moments<T>(v, m2, (T)rule.y[i]));
    return acc;
}

template <typename T, int K>
__device__ __forceinline__ T dis_value(const DisArgs& a, const double* u, const RuleView& rule) {
    const T Q2 = (T)a.Q2;
    if constexpr (K == NLO_DIS_LO) {
        // S/dis/observables.py:345-352 and S/dis/lo.py:12-51
        const T z = (T)u[0];
        const T r = (T)(a.maxr * u[1]);
        T tot = T(0);
        for (int q = 0; q < a.nq; ++q)
            tot += lo_wavefunction<T>(r, z, Q2, (T)a.mass[q], (T)a.charge[q], a.transverse != 0);
        return T(2.0 * kPiD) * r * (T)a.maxr * tot * amp(a, r, (T)a.y_lo);
    } else if constexpr (K <= NLO_DIS_DIP_L_CD) {
        // integrand_dip_massive (S/dis/observables.py:107-172)
        const T z1 = (T)(kZmin + u[0] * (1.0 - 2.0 * kZmin));
        const double x01d = a.maxr * u[1];
        const bool inside = x01d >= a.min_r && x01d <= a.max_r;
        const T x01 = (T)x01d;
        const T x01e = (T)fmin(fmax(x01d, a.min_r), a.max_r);
        const T mf = (T)a.mass[0];
        const T N01 = amp(a, x01e, (T)a.y_dip);
        T res;
        if constexpr (K == NLO_DIS_DIP_T0) res = it_dip_0<T>(Q2, z1, x01e, mf);
        else if constexpr (K == NLO_DIS_DIP_T1) res = it_dip_1<T>(Q2, z1, x01e, mf, (T)u[2]);
        else if constexpr (K == NLO_DIS_DIP_T2) res = it_dip_2<T>(Q2, z1, x01e, mf, (T)u[2], (T)u[3]);
        else if constexpr (K == NLO_DIS_DIP_L_CONST) res = il_dip_const<T>(Q2, z1, x01e, mf);
        else if constexpr (K == NLO_DIS_DIP_L_AB) res = il_dip_ab<T>(Q2, z1, x01e, mf, (T)u[2]);
        else res = il_dip_cd<T>(Q2, z1, x01e, mf, (T)u[2], (T)u[3]);
        res = N01 * res * (x01 * (T)a.maxr * T(1.0 - 2.0 * kZmin)) * (dis_alpha<T>(x01e, a.coup) * T(kCF / kPiD));
        return (inside && isfinite(res)) ? res : T(0);
    } else {
        // integrand_qgunsub_massive (S/dis/observables.py:175-250) and the nested
        // form (:253-331)
        if (a.z2min > 1.0) return T(0);
        const QgPoint<T> p = qg_point<T>(a, u);
        const Tripole<T> v = tripole_vars<T>(Q2, p.mf, p.z1, p.z2, p.x01sq, p.x02sq, p.x21sq);
        const T m2 = p.m2, z1 = p.z1, z2 = p.z2;
        T res;
        if constexpr (K == NLO_DIS_QG_I1) {
            const TFixed<T> f = t_fixed<T>(v, m2, z1, z2);
            const T N01 = amp(a, p.x01, p.Ye);
            if (a.transverse)
                res = N01 * it_qg_uvsub<T>(v, m2, z1, z2, p.x01sq) + p.N012 * it_qg_i1<T>(v, f, m2, z1, z2);
            else
                res = N01 * il_qg_uvsub<T>(v, Q2, z1, z2, p.x01sq) + p.N012 * il_qg_i1<T>(v, f, Q2, z1, z2);
        } else if constexpr (K == NLO_DIS_QG_I2 || K == NLO_DIS_QG_I3) {
            const TFixed<T> f = t_fixed<T>(v, m2, z1, z2);
            const TMoments<T> m1 = t_moments<T>(v, m2, (T)u[5]);
            if constexpr (K == NLO_DIS_QG_I2) {
                res = p.N012 * (a.transverse ? it_qg_i2<T>(v, f, m2, z1, z2, m1) : il_qg_i2<T>(v, f, Q2, z1, z2, m1));
            } else {
                const TMoments<T> mm2 = t_moments<T>(v, m2, (T)u[6]);
                res = p.N012 * (a.transverse ? it_qg_i3<T>(v, f, m2, z1, z2, m1, mm2)
                                             : il_qg_i3<T>(v, Q2, m2, z1, z2, m1, mm2));
            }
        } else {
            // y_t integrals on the mapped inner rule: I2 is linear and I3
            // bilinear in the G values (K/dis/i2_i3/kernel.cu:768-912)
            const TFixed<T> f = t_fixed<T>(v, m2, z1, z2);
            const TMoments<T> acc = nested_moments<T>(v, m2, rule);
            if constexpr (K == NLO_DIS_QG_I2_NESTED)
                res = p.N012 * (a.transverse ? it_qg_i2<T>(v, f, m2, z1, z2, acc) : il_qg_i2<T>(v, f, Q2, z1, z2, acc));
            else
                res = p.N012 * (a.transverse ? it_qg_i3<T>(v, f, m2, z1, z2, acc, acc)
                                             : il_qg_i3<T>(v, Q2, m2, z1, z2, acc, acc));
        }
        return qg_finish<T>(p, res);
    }
}

// The nested I2 and I3 at one point (integrand_qg_nested_massive "I2" and
// "I3", S/dis/observables.py:253-331) from one pass over the inner rule: both
// follow from the same integrated moments, as in the reference's combined
// kernel (K/dis/i2_i3/kernel.cu:768-912).
template <typename T>
__device__ __forceinline__ void qg_nested_pair(const DisArgs& a, const double* u, const RuleView& rule, T& i2,
                                               T& i3) {
    i2 = T(0);
    i3 = T(0);
    if (a.z2min > 1.0) return;
    const T Q2 = (T)a.Q2;
    const QgPoint<T> p = qg_point<T>(a, u);
    const Tripole<T> v = tripole_vars<T>(Q2, p.mf, p.z1, p.z2, p.x01sq, p.x02sq, p.x21sq);
    const T m2 = p.m2, z1 = p.z1, z2 = p.z2;
    const TFixed<T> f = t_fixed<T>(v, m2, z1, z2);
    const TMoments<T> acc = nested_moments<T>(v, m2, rule);
    if (a.transverse) {
        i2 = p.N012 * it_qg_i2<T>(v, f, m2, z1, z2, acc);
        i3 = p.N012 * it_qg_i3<T>(v, f, m2, z1, z2, acc, acc);
    } else {
        i2 = p.N012 * il_qg_i2<T>(v, f, Q2, z1, z2, acc);
        i3 = p.N012 * il_qg_i3<T>(v, Q2, m2, z1, z2, acc, acc);
    }
    i2 = qg_finish<T>(p, i2);
    i3 = qg_finish<T>(p, i3);
}

// ------------------------------------------------------------ host helpers

DisDipole make_dipole(const nlo_dis_dipole& d);
int check_dipole(const nlo_dis_dipole* d);
// ends: the BK table's first and last r when the caller has them (a batch
// reads them once), else NULL and they are copied from the device.
int make_args(const nlo_dis_desc* desc, DisArgs& a, const double* ends = nullptr);

}  // namespace dis
}  // namespace nlo
