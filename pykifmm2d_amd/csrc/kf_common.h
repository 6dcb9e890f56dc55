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

// A tabulated radial kernel (KF_TABLE, kf_ktable_create): piecewise polynomials in
// u = log2(r2) on the segments [j du, (j + 1) du), j = s0 .. s0 + nseg - 1, du = 1 / rdu a power
// of two, so y = u rdu, its floor and x = 2 (y - floor) - 1 are exact; u outside the segments
// is clamped to their ends (x = -1 or 1).  G(0) = g0.  c points to the coefficients in the
// memory of the code that reads them: a row of ld (even, >= deg + 1) per segment, from x^0 up,
// zero past x^deg, 16-byte aligned, so each pair of coefficients is one 16-byte load.
struct KfTable {
    const double* c = nullptr;
    int nseg = 0, deg = 0, ld = 0, s0 = 0;
    double rdu = 1.0, g0 = 0.0;
};
static constexpr double KF_INV_LN2 = 1.44269504088896340736;

KF_HD double ktab_poly(const KfTable& t, double u) {
    const double y = u * t.rdu, f = floor(y);
    const double j = f - (double)t.s0;
    const double jc = fmin(fmax(j, 0.0), (double)(t.nseg - 1));  // a nan u lands in segment 0
    const double x = j == jc ? 2.0 * (y - f) - 1.0 : (j < jc ? -1.0 : 1.0);
    const double* c = (const double*)__builtin_assume_aligned(t.c + (size_t)jc * t.ld, 16);
    // Horner from the top pair; the zero padding leaves g = c[deg] after the first pair
    double g = 0.0;
    for (int k = t.ld - 2; k >= 0; k -= 2) {
        g = fma(g, x, c[k + 1]);
        g = fma(g, x, c[k]);
    }
    return g;
}

struct KfKernel {
    int kind;      // KF_LAPLACE, KF_MODHELM or KF_TABLE
    double k;      // modified Helmholtz wavenumber
    double scale;  // modified Helmholtz prefactor
    KfTable th{}, td{};  // KF_TABLE: the table in host and in device memory
    // a complex kernel (KF_TABLE only): G = table + i table_im, the imaginary part's table in
    // host and device memory; of_r2 / eval give the real part
    int cplx = 0;
    KfTable th_im{}, td_im{};

    // The imaginary part of a complex kernel on the host, as of_r2 / eval give the real part.
    double of_r2_im(double r2) const { return r2 == 0.0 ? th_im.g0 : ktab_poly(th_im, std::log2(r2)); }
    double eval_im(double sx, double sy, double tx, double ty) const {
        double dx = tx - sx, dy = ty - sy;
        return of_r2_im(std::fma(dx, dx, dy * dy));
    }

    // G as a function of the squared distance.
    KF_HD double of_r2(double r2) const {
        if (kind == KF_LAPLACE) return KF_LAP_SCALE * log(r2);
        if (kind == KF_TABLE) {
#if defined(__HIP_DEVICE_COMPILE__)
            const KfTable& t = td;
#else
            const KfTable& t = th;
#endif
            return r2 == 0.0 ? t.g0 : ktab_poly(t, log2(r2));
        }
        return scale * k0_f64(k * sqrt(r2));
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
    // raw() where fast_ok(r2); loops evaluating several values check them together and take
    // raw() for the group otherwise
    KF_HD double raw_fast(double r2, const double* tab) const { return kf_logt_fast(r2, tab, lc); }
    KF_HD bool fast_ok(double r2) const { return kf_logt_normal(r2); }
    KF_HD double pref() const { return KF_LAP_SCALE; }
};
// A tabulated kernel (KF_TABLE): u = log2(r2) from the table-driven log in the plan's kernels
// (the library log2 in the direct sums); the polynomial coefficients are read from global
// memory (the table is small enough to stay in the L1/L2 caches).
template <>
struct KerT<KF_TABLE> {
    double k, scale;
    LogtCoef lc = KF_LOGT_COEF;
    KfTable t{};
    KF_HD double g(double r2) const { return r2 == 0.0 ? t.g0 : ktab_poly(t, log2(r2)); }
    KF_HD double g(double r2, const double* tab) const { return raw(r2, tab); }
    KF_HD double raw(double r2, const double* tab) const {
        return r2 == 0.0 ? t.g0 : ktab_poly(t, kf_logt(r2, tab, lc) * KF_INV_LN2);
    }
    KF_HD double raw_fast(double r2, const double* tab) const {
        return r2 == 0.0 ? t.g0 : ktab_poly(t, kf_logt_fast(r2, tab, lc) * KF_INV_LN2);
    }
    KF_HD bool fast_ok(double r2) const { return r2 == 0.0 || kf_logt_normal(r2); }
    KF_HD double pref() const { return 1.0; }
};
template <>
struct KerT<KF_MODHELM> {
    double k, scale;
    LogtCoef lc = KF_LOGT_COEF;
    double lk = log(0.5 * k) + KF_EULER_GAMMA;  // log(k/2) + gamma, for k0_r2
    KF_HD double g(double r2) const { return scale * k0_f64(k * sqrt(r2)); }
    K0Series sc = KF_K0_SERIES;  // in the kernel arguments: scalar-register coefficients
    KF_HD double g(double r2, const double* tab) const { return scale * k0_r2(r2, k, lk, tab, lc, sc); }
    KF_HD double raw(double r2, const double* tab) const { return k0_r2(r2, k, lk, tab, lc, sc); }
    KF_HD double raw_fast(double r2, const double* tab) const { return raw(r2, tab); }
    KF_HD bool fast_ok(double) const { return true; }
    KF_HD double pref() const { return scale; }
};

// N kernel values raw(r2[i]) at once: raw_fast for every input, without a branch (so that the N
// evaluations can be interleaved by the compiler), then raw() again for the lanes where an
// input is one raw_fast does not take (a wave none of whose lanes has one skips that).  The
// values are raw()'s.
template <int KIND, int N>
KF_HD void raw_group(const KerT<KIND>& ker, const double (&r2)[N], const double* tab, double (&v)[N]) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        v[i] = ker.raw_fast(r2[i], tab);
        ok = ok && ker.fast_ok(r2[i]);
    }
    if (__builtin_expect(!ok, 0)) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = ker.raw(r2[i], tab);
    }
}
