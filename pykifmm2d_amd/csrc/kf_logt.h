// kf_logt.h -- the table-driven FP64 logarithm of the near-field, P2M and L2P kernels.
//
// Host-compilable on its own (the CPU tests build it with g++; KF_HD is then plain inline).
#pragma once
#include <cmath>
#include <cstdint>

#include "kf_log_tab.h"

#ifndef KF_HD
#define KF_HD inline
#endif

// Table-driven log (Tang's method; the layout, table and polynomial are those of
// kf_log_tab.h, made by gen_log_tab.py): with u = hi(x) - C, e = u >> 20, j = bits 13..19 of
// u and m = x 2^-e in [C', 2C'), log x = e ln2 + L_j + log1p(r), r = m c_j - 1 (|r| <= 2^-8),
// log1p(r) = r + r^2 Q(r).  The evaluation starts from w = K - hi(x) = 0xfffff - u: its
// arithmetic shift by 20 is -e, its bits 13..19 are 127 - j (the table is stored from the last
// interval down), and m = ldexp(x, -e), so the exponent, the table index and the reduced
// argument take one subtraction, three shifts/masks and one ldexp.  The table entry (c_j, L_j)
// is one 16-byte read from `tab` (KF_LOGT_TAB or, in device code, a copy in LDS).  x outside
// the positive normal doubles (0, subnormals, inf, nan, negatives) goes to the library log
// (log(0) = -inf); no lane of the near-field loops normally takes that branch.
// The polynomial coefficients and ln2 as a value: kernels receive a copy in their arguments, so
// the device holds them in scalar registers and every Horner step is one FMA.
struct LogtCoef {
    double q[KF_LOGT_QN];
    double ln2;
};
static constexpr LogtCoef KF_LOGT_COEF = {{KF_LOGT_Q[0], KF_LOGT_Q[1], KF_LOGT_Q[2], KF_LOGT_Q[3], KF_LOGT_Q[4]},
                                          KF_LOGT_LN2};
static_assert(KF_LOGT_QN == 5, "KF_LOGT_COEF lists five coefficients");

// The library log for the inputs kf_logt does not take (no near-field pair normally does): a
// function of its own on the device, so that the registers of its extended-precision evaluation
// are not reserved across the loops that call kf_logt.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __attribute__((noinline))
#else
inline
#endif
double kf_log_rare(double x) {
    return log(x);
}

// kf_logt for positive normal x only (no check; other inputs give garbage): loops that evaluate
// several logs at once check them together with kf_logt_normal and redo the group with kf_logt
// in the rare case, keeping the common path free of branches.
KF_HD double kf_logt_fast(double x, const double* tab, const LogtCoef& c = KF_LOGT_COEF) {
    const uint32_t w = KF_LOGT_K - (uint32_t)(__builtin_bit_cast(uint64_t, x) >> 32);
    const int ne = (int32_t)w >> 20;  // -e
    const int k = (int)((w >> (20 - KF_LOGT_BITS)) & ((1u << KF_LOGT_BITS) - 1));  // 127 - j
    const double m = __builtin_ldexp(x, ne);  // exact for normal x
    const double r = fma(m, tab[2 * k], -1.0);
    double q = c.q[KF_LOGT_QN - 1];
#pragma unroll
    for (int i = KF_LOGT_QN - 2; i >= 0; --i) q = fma(q, r, c.q[i]);
    return fma(-(double)ne, c.ln2, tab[2 * k + 1] + fma(r * r, q, r));
}

KF_HD bool kf_logt_normal(double x) {
#if defined(__clang__)
    return __builtin_isfpclass(x, 0x100);  // positive normal
#else
    return x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308;
#endif
}

KF_HD double kf_logt(double x, const double* tab, const LogtCoef& c = KF_LOGT_COEF) {
    const uint32_t w = KF_LOGT_K - (uint32_t)(__builtin_bit_cast(uint64_t, x) >> 32);
    const int ne = (int32_t)w >> 20;  // -e
    const int k = (int)((w >> (20 - KF_LOGT_BITS)) & ((1u << KF_LOGT_BITS) - 1));  // 127 - j
    const double m = __builtin_ldexp(x, ne);  // exact for normal x
    const double r = fma(m, tab[2 * k], -1.0);
    double q = c.q[KF_LOGT_QN - 1];
#pragma unroll
    for (int i = KF_LOGT_QN - 2; i >= 0; --i) q = fma(q, r, c.q[i]);
    double y = fma(-(double)ne, c.ln2, tab[2 * k + 1] + fma(r * r, q, r));
#if defined(__clang__)
    const bool normal = __builtin_isfpclass(x, 0x100);  // positive normal
#else
    const bool normal = x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308;
#endif
    if (__builtin_expect(!normal, 0)) y = kf_log_rare(x);
    return y;
}

