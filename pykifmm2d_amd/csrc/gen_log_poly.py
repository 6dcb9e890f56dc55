"""Generate kf_log_poly.h: polynomial P(z) ~ atanh(sqrt z)/sqrt z on [0, zmax] for the FP64 log.

log(m) = 2 atanh(s) = 2 s P(s^2), s = (m-1)/(m+1), m in [sqrt(1/2), sqrt(2)) so
zmax = ((sqrt2-1)/(sqrt2+1))^2.  P is the Chebyshev interpolant of the exact function
(mpmath, 50 digits) at degree DEG, re-expanded in monomials; the script prints its maximum
relative error on a fine grid.  Run: python gen_log_poly.py
"""
import os

import mpmath as mp

mp.mp.dps = 50
DEG = 7


def main():
    zmax = ((mp.sqrt(2) - 1) / (mp.sqrt(2) + 1)) ** 2
    f = lambda z: mp.atanh(mp.sqrt(z)) / mp.sqrt(z) if z > 0 else mp.mpf(1)
    n = DEG + 1
    nodes = [zmax / 2 * (1 + mp.cos(mp.pi * (j + mp.mpf(1) / 2) / n)) for j in range(n)]
    # interpolating polynomial in monomial form (Vandermonde solve in 50-digit arithmetic)
    A = mp.matrix([[z ** k for k in range(n)] for z in nodes])
    b = mp.matrix([f(z) for z in nodes])
    c = mp.lu_solve(A, b)
    worst = mp.mpf(0)
    for i in range(2001):
        z = zmax * i / 2000
        p = sum(c[k] * z ** k for k in range(n))
        worst = max(worst, abs(p / f(z) - 1))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kf_log_poly.h")
    with open(path, "w") as fh:
        fh.write("// Generated by gen_log_poly.py -- do not edit.\n")
        fh.write("// P(z) ~ atanh(sqrt z)/sqrt z on [0, %s], max rel. error %s\n" % (mp.nstr(zmax, 8), mp.nstr(worst, 3)))
        fh.write("#pragma once\n")
        fh.write("static constexpr int KF_LOG_POLY_N = %d;\n" % n)
        fh.write("static constexpr double KF_LOG_POLY[%d] = {\n" % n)
        fh.write(",\n".join("    %s" % mp.nstr(c[k], 20, min_fixed=-1, max_fixed=-1) for k in range(n)))
        fh.write("\n};\n")
    print("wrote", path, "max rel err", mp.nstr(worst, 3))


if __name__ == "__main__":
    main()
