// kf_k0.h -- K0 of the modified Helmholtz kernel in FP64.
//
// Host-compilable on its own (the CPU tests build it with g++; KF_HD is then plain inline).
#pragma once
#include <cmath>

#include "k0_cheb.h"
#include "kf_logt.h"

#ifndef KF_HD
#define KF_HD inline
#endif

static constexpr double KF_EULER_GAMMA = 0.57721566490153286061;

// K0(x) for x > KF_K0_SPLIT: Chebyshev expansion of sqrt(x) e^x K0(x) in u = 2*split/x - 1.
// On the device it is a function of its own (not inlined): inlined into a kernel's pair loop,
// its 38 coefficients were hoisted out of the loop into 76 vector registers, which halved the
// occupancy of every modified Helmholtz kernel, while the loops take this branch only for
// pairs farther apart than 1/k.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __attribute__((noinline))
#else
inline
#endif
double k0_far(double x) {
    double u = 2.0 * KF_K0_SPLIT / x - 1.0, u2 = 2.0 * u, b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int k = KF_K0_CHEB_N - 1; k >= 1; --k) {
        double b0 = fma(u2, b1, KF_K0_CHEB[k] - b2);
        b2 = b1;
        b1 = b0;
    }
    double f = fma(u, b1, KF_K0_CHEB[0] - b2);
    return exp(-x) * f / sqrt(x);
}

// K0(x) in FP64: power series for x <= KF_K0_SPLIT, k0_far beyond (coefficients from
// gen_k0_cheb.py).  K0(0) = inf, K0(x<0) = nan, matching misc/numba_special_functions.py:167-181.
KF_HD double k0_f64(double x) {
    if (x <= KF_K0_SPLIT) {
        if (x < 0.0) return NAN;
        if (x == 0.0) return INFINITY;
        double q = 0.25 * x * x, i0 = 0.0, s = 0.0;
#pragma unroll
        for (int k = KF_I0_SERIES_N - 1; k >= 0; --k) {
            i0 = fma(i0, q, KF_I0_SERIES[k]);
            s = fma(s, q, KF_K0_SERIES_H[k]);
        }
        return fma(-(log(0.5 * x) + KF_EULER_GAMMA), i0, s);
    }
    return k0_far(x);
}

// K0(k sqrt(r2)) from the squared distance, for the plan's kernels (tab: the table-driven log's
// table, as for the Laplace kernel).  On the power-series branch, x^2 = k^2 r2 <= split^2, it
// needs neither the square root nor the library log: the series runs in x^2 / 4, and
// log(x / 2) = log(r2) / 2 + log(k / 2), with lk = log(k / 2) + gamma precomputed and log(r2)
// from kf_logt.  That branch is the one the near field, P2M and L2P take wherever k times the
// leaf width is below about 1/3.  Beyond it, k0_f64.  r2 = 0 gives +inf (kf_logt(0) = -inf).
// The two power series' coefficients as a value: carried in the kernel arguments, they sit in
// scalar registers and each Horner step is one FMA with a scalar operand (as compile-time
// constants, the compiler kept them in vector registers and copied one into the accumulator
// before every step: a move per term and 56 vector registers).
struct K0Series {
    double i0[KF_I0_SERIES_N], h[KF_I0_SERIES_N];
};
static_assert(KF_I0_SERIES_N == KF_K0_SERIES_H_N, "the two series have equal length");
constexpr K0Series k0_series() {
    K0Series c{};
    for (int j = 0; j < KF_I0_SERIES_N; ++j) {
        c.i0[j] = KF_I0_SERIES[j];
        c.h[j] = KF_K0_SERIES_H[j];
    }
    return c;
}
static constexpr K0Series KF_K0_SERIES = k0_series();

KF_HD double k0_r2(double r2, double k, double lk, const double* tab, const LogtCoef& lc,
                   const K0Series& sc = KF_K0_SERIES) {
    const double x2 = (k * k) * r2;
    if (x2 <= KF_K0_SPLIT * KF_K0_SPLIT) {
        double q = 0.25 * x2, i0 = 0.0, s = 0.0;
#pragma unroll
        for (int j = KF_I0_SERIES_N - 1; j >= 0; --j) {
            i0 = fma(i0, q, sc.i0[j]);
            s = fma(s, q, sc.h[j]);
        }
        return fma(-fma(0.5, kf_logt(r2, tab, lc), lk), i0, s);
    }
    return k0_f64(k * sqrt(r2));
}
