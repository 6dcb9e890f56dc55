// kf_common.h -- Green's functions and small helpers shared by host and device code.
//
// Kernels (kf_kernel_desc.kind):
//   KF_LAPLACE  G(s,t) = -(1/4pi) log |t-s|^2          (reference kernels/laplace.py:5-8)
//   KF_MODHELM  G(s,t) = scale * K0(k |t-s|)            (reference kernels/modified_helmholtz.py:7-9;
//                                                        scale = 1 there, 1/(2pi) in the examples)
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "../../include/kifmm2d.h"

#define KF_HD __host__ __device__ __forceinline__
#include "kf_logt.h"
#include "kf_k0.h"

static constexpr double KF_PI = 3.14159265358979323846;
static constexpr double KF_LAP_SCALE = -0.25 / KF_PI;

// FP32 K0: the same expansions truncated to single precision.
static constexpr int KF_I0_SERIES_N32 = 8;
static constexpr int KF_K0_CHEB_N32 = 16;
KF_HD float k0_f32(float x) {
    if (x <= (float)KF_K0_SPLIT) {
        if (x < 0.0f) return NAN;
        if (x == 0.0f) return INFINITY;
        float q = 0.25f * x * x, i0 = 0.0f, s = 0.0f;
#pragma unroll
        for (int k = KF_I0_SERIES_N32 - 1; k >= 0; --k) {
            i0 = fmaf(i0, q, (float)KF_I0_SERIES[k]);
            s = fmaf(s, q, (float)KF_K0_SERIES_H[k]);
        }
        return fmaf(-(logf(0.5f * x) + (float)KF_EULER_GAMMA), i0, s);
    }
    float u = (float)(2.0 * KF_K0_SPLIT) / x - 1.0f, u2 = 2.0f * u, b1 = 0.0f, b2 = 0.0f;
#pragma unroll
    for (int k = KF_K0_CHEB_N32 - 1; k >= 1; --k) {
        float b0 = fmaf(u2, b1, (float)KF_K0_CHEB[k] - b2);
        b2 = b1;
        b1 = b0;
    }
    float f = fmaf(u, b1, (float)KF_K0_CHEB[0] - b2);
    return expf(-x) * f / sqrtf(x);
}

struct KfKernel {
    int kind;      // KF_LAPLACE or KF_MODHELM
    double k;      // modified Helmholtz wavenumber
    double scale;  // modified Helmholtz prefactor

    // G as a function of the squared distance.
    KF_HD double of_r2(double r2) const {
        if (kind == KF_LAPLACE) return KF_LAP_SCALE * log(r2);
        return scale * k0_f64(k * sqrt(r2));
    }
    KF_HD float of_r2f(float r2) const {
        if (kind == KF_LAPLACE) return (float)KF_LAP_SCALE * logf(r2);
        return (float)scale * k0_f32((float)k * sqrtf(r2));
    }
    KF_HD double eval(double sx, double sy, double tx, double ty) const {
        double dx = tx - sx, dy = ty - sy;
        return of_r2(fma(dx, dx, dy * dy));
    }
};

// Compile-time kernel kinds for the evaluation kernels: g(r2) = G, raw(r2) = G without the
// constant prefactor pref() (the near-field sums apply the prefactor once per target).
template <int KIND>
struct KerT;
// g(r2) uses the library log (the direct sums, which also state the FMM errors); the plan's
// kernels pass `tab`, their LDS copy of the table-driven log's table (kf_logt.h).
template <>
struct KerT<KF_LAPLACE> {
    double k, scale;
    LogtCoef lc = KF_LOGT_COEF;
    KF_HD double g(double r2) const { return KF_LAP_SCALE * log(r2); }
    KF_HD double g(double r2, const double* tab) const { return KF_LAP_SCALE * kf_logt(r2, tab, lc); }
    KF_HD double raw(double r2, const double* tab) const { return kf_logt(r2, tab, lc); }
    KF_HD double pref() const { return KF_LAP_SCALE; }
};
template <>
struct KerT<KF_MODHELM> {
    double k, scale;
    LogtCoef lc = KF_LOGT_COEF;
    double lk = log(0.5 * k) + KF_EULER_GAMMA;  // log(k/2) + gamma, for k0_r2
    KF_HD double g(double r2) const { return scale * k0_f64(k * sqrt(r2)); }
    KF_HD double g(double r2, const double* tab) const { return scale * k0_r2(r2, k, lk, tab, lc); }
    KF_HD double raw(double r2, const double* tab) const { return k0_r2(r2, k, lk, tab, lc); }
    KF_HD double pref() const { return scale; }
};
