// CPU check of the plan kernels' K0 (pykifmm2d_amd/csrc/kf_k0.h: k0_r2, the squared-distance
// form with the table-driven log on its power-series branch).  Reads lines "k r2" from stdin and
// prints K0(k sqrt(r2)) as k0_r2 computes it, one per line, with 17 significant digits.
#include <cmath>
#include <cstdio>

#include "../pykifmm2d_amd/csrc/kf_k0.h"

int main() {
    double k, r2;
    while (std::scanf("%lf %lf", &k, &r2) == 2) {
        const double lk = std::log(0.5 * k) + KF_EULER_GAMMA;
        std::printf("%.17g\n", k0_r2(r2, k, lk, KF_LOGT_TAB, KF_LOGT_COEF));
    }
    return 0;
}
