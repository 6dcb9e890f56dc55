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
// log1p(r) = r + r^2 Q(r).  Integer field extraction and six FMA-class operations replace the
// reciprocal, Newton steps and degree-7 polynomial of kf_log; the table entry (c_j, L_j) is
// one 16-byte read from `tab` (KF_LOGT_TAB or, in device code, a copy in LDS).  x outside the
// positive normal doubles (0, subnormals, inf, nan, negatives) goes to the library log
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

KF_HD double kf_logt(double x, const double* tab, const LogtCoef& c = KF_LOGT_COEF) {
    const uint64_t bits = __builtin_bit_cast(uint64_t, x);
    const uint32_t u = (uint32_t)(bits >> 32) - KF_LOGT_C;
    const int e = (int32_t)u >> 20;
    const int j = (int)((u >> (20 - KF_LOGT_BITS)) & ((1u << KF_LOGT_BITS) - 1));
    const double m = __builtin_bit_cast(double, (bits & 0xffffffffull) | ((uint64_t)(KF_LOGT_C + (u & 0xfffffu)) << 32));
    const double r = fma(m, tab[2 * j], -1.0);
    double q = c.q[KF_LOGT_QN - 1];
#pragma unroll
    for (int k = KF_LOGT_QN - 2; k >= 0; --k) q = fma(q, r, c.q[k]);
    double y = fma((double)e, c.ln2, tab[2 * j + 1] + fma(r * r, q, r));
#if defined(__clang__)
    const bool normal = __builtin_isfpclass(x, 0x100);  // positive normal
#else
    const bool normal = x >= 2.2250738585072014e-308 && x <= 1.7976931348623157e308;
#endif
    if (__builtin_expect(!normal, 0)) y = log(x);
    return y;
}

