// ktable.hip -- tabulated radial kernels (KF_TABLE): the C-ABI that creates, evaluates and
// destroys them.  The table format is documented with kf_ktable_create in include/kifmm2d.h.
// Creating and evaluating a table on the host make no HIP call.
#include <cmath>
#include <vector>

#include "kf_internal.h"

using namespace kf;

extern "C" {

int kf_ktable_create(const double* coef, int32_t nseg, int32_t deg, double u0, double du, double g0,
                     kf_ktable** out) {
    return guarded([&] {
        KF_REQUIRE(coef && out, KF_ERR_ARG, "null argument");
        KF_REQUIRE(nseg >= 1 && deg >= 0 && deg <= 24, KF_ERR_ARG, "table needs nseg >= 1 and 0 <= deg <= 24");
        int e = 0;
        KF_REQUIRE(std::isfinite(du) && std::frexp(du, &e) == 0.5 && e >= -9 && e <= 5, KF_ERR_ARG,
                   "table needs du a power of two in [2^-10, 16]");
        const double s0 = u0 / du;
        KF_REQUIRE(std::isfinite(u0) && s0 == std::floor(s0) && std::fabs(s0) <= (double)(1 << 20), KF_ERR_ARG,
                   "table needs u0 a multiple of du, |u0 / du| <= 2^20");
        KF_REQUIRE((double)nseg + s0 <= (double)(1 << 21), KF_ERR_ARG, "table too long");
        const size_t n = (size_t)nseg * (deg + 1);
        for (size_t i = 0; i < n; ++i) KF_REQUIRE(std::isfinite(coef[i]), KF_ERR_ARG, "table coefficients must be finite");
        std::unique_ptr<kf_ktable> t(new kf_ktable());
        const int ld = (deg + 2) & ~1;
        t->c.assign((size_t)nseg * ld, 0.0);
        for (int j = 0; j < nseg; ++j)
            for (int k = 0; k <= deg; ++k) t->c[(size_t)j * ld + k] = coef[(size_t)j * (deg + 1) + k];
        t->ld = ld;
        t->nseg = nseg;
        t->deg = deg;
        t->s0 = (int)s0;
        t->du = du;
        t->g0 = g0;
        *out = t.release();
    });
}

void kf_ktable_destroy(kf_ktable* table) { delete table; }

int kf_ktable_eval(const kf_ktable* table, const double* r2, int64_t n, double* out) {
    return guarded([&] {
        KF_REQUIRE(table && (n == 0 || (r2 && out)), KF_ERR_ARG, "null argument");
        KfKernel k{KF_TABLE, 0.0, 1.0};
        k.th = table->view(false);
        for (int64_t i = 0; i < n; ++i) out[i] = k.of_r2(r2[i]);
    });
}

}  // extern "C"
