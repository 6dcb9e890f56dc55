// CPU check of the table-driven logarithm (pykifmm2d_amd/csrc/kf_logt.h) against the 80-bit
// long double logl: maximum error in units of the last place over log-uniform samples of the
// whole positive normal range, of the squared distances the near field sees, of the
// neighbourhood of 1 (where the c = 1 table entry keeps the relative accuracy) and of every
// table interval's end points; exact special values; the table entries themselves.
// Prints one line "max_ulp <v> near1_ulp <v> bound_ulp <v> table_ok <0|1> special_ok <0|1>".
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "../pykifmm2d_amd/csrc/kf_logt.h"

static double ulp_err(double x) {
    const long double ref = logl((long double)x);
    const double r = (double)ref;
    if (r == 0.0) return kf_logt(x, KF_LOGT_TAB) == 0.0 ? 0.0 : 1e9;
    const double ulp = std::fabs(std::nextafter(r, INFINITY) - r);
    return (double)(std::fabs((long double)kf_logt(x, KF_LOGT_TAB) - ref) / ulp);
}

static double from_hi(uint32_t hi, uint32_t lo) {
    const uint64_t b = ((uint64_t)hi << 32) | lo;
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}

int main() {
    std::mt19937_64 rng(11);
    double max_ulp = 0, near1 = 0, bound = 0;
    std::uniform_real_distribution<double> wide(-1021.9, 1023.9), r2(-60.0, 10.0);
    for (int i = 0; i < 2000000; ++i) max_ulp = std::fmax(max_ulp, ulp_err(std::exp2(wide(rng))));
    for (int i = 0; i < 2000000; ++i) max_ulp = std::fmax(max_ulp, ulp_err(std::exp2(r2(rng))));
    // within 2^-12 of 1, and the first 10^5 doubles on either side of it
    std::uniform_real_distribution<double> d1(-1.0 / 4096, 1.0 / 4096);
    for (int i = 0; i < 1000000; ++i) near1 = std::fmax(near1, ulp_err(1.0 + d1(rng)));
    double up = 1.0, dn = 1.0;
    for (int i = 0; i < 100000; ++i) {
        up = std::nextafter(up, 2.0);
        dn = std::nextafter(dn, 0.0);
        near1 = std::fmax(near1, std::fmax(ulp_err(up), ulp_err(dn)));
    }
    // interval end points (upper words C + j 2^13, low word 0 and all ones) at several exponents
    for (int ex = -40; ex <= 40; ex += 8)
        for (uint32_t j = 0; j <= (1u << KF_LOGT_BITS); ++j)
            for (uint32_t lo : {0u, 0xffffffffu}) {
                const double m = from_hi(KF_LOGT_C + (j << (20 - KF_LOGT_BITS)), lo);
                for (double s : {1.0, 1.0 + DBL_EPSILON, 1.0 - DBL_EPSILON / 2})
                    bound = std::fmax(bound, ulp_err(std::ldexp(m * s, ex)));
            }
    max_ulp = std::fmax(max_ulp, std::fmax(near1, bound));
    // table: c = 1, L = 0 at the interval holding 1; L_j = -log(c_j) to half an ulp
    int table_ok = 1;
    for (int j = 0; j < (1 << KF_LOGT_BITS); ++j) {
        const int k = (1 << KF_LOGT_BITS) - 1 - j;  // stored from the last interval down
        const double c = KF_LOGT_TAB[2 * k], L = KF_LOGT_TAB[2 * k + 1];
        const long double ref = -logl((long double)c);
        if (L != 0.0 && std::fabs((long double)L - ref) > 0.5L * std::fabs(std::nextafter(L, INFINITY) - L)) table_ok = 0;
        const double A = from_hi(KF_LOGT_C + ((uint32_t)j << (20 - KF_LOGT_BITS)), 0);
        const double B = from_hi(KF_LOGT_C + ((uint32_t)(j + 1) << (20 - KF_LOGT_BITS)), 0);
        if (A <= 1.0 && 1.0 < B && !(c == 1.0 && L == 0.0)) table_ok = 0;
        if (std::fabs((long double)A * c - 1) > 1.0L / 256 || std::fabs((long double)B * c - 1) > 1.0L / 256) table_ok = 0;
    }
    const double sub = 1e-310;
    const int special_ok = std::isinf(kf_logt(0.0, KF_LOGT_TAB)) && kf_logt(0.0, KF_LOGT_TAB) < 0 &&
                           std::isnan(kf_logt(NAN, KF_LOGT_TAB)) && std::isnan(kf_logt(-1.0, KF_LOGT_TAB)) &&
                           std::isinf(kf_logt(INFINITY, KF_LOGT_TAB)) && kf_logt(1.0, KF_LOGT_TAB) == 0.0 &&
                           std::fabs(kf_logt(sub, KF_LOGT_TAB) - std::log(sub)) <= 1e-12 &&
                           std::fabs(kf_logt(DBL_MIN, KF_LOGT_TAB) - std::log(DBL_MIN)) <= 1e-12;
    std::printf("max_ulp %.4f near1_ulp %.4f bound_ulp %.4f table_ok %d special_ok %d\n", max_ulp, near1, bound,
                table_ok, special_ok);
    return 0;
}
