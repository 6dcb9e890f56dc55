"""Generate kf_log_tab.h: the table and polynomial of the table-driven FP64 logarithm.

Method (P. T. P. Tang, "Table-driven implementation of the logarithm function in IEEE
floating-point arithmetic", ACM TOMS 16(4), 1990), in this layout:

  u = hi(x) - C (hi = upper 32 bits), e = u >> 20 (arithmetic), j = bits 13..19 of u,
  m = x 2^-e (upper word C + (u & 0xfffff)), so m lies in [C', 2C'), C' ~ sqrt(1/2);
  interval j of m is [A_j, B_j) with upper words C + j 2^13 and C + (j+1) 2^13;
  r = m c_j - 1 (one fma), |r| <= 2^-8;  log x = e ln2 + L_j + r + r^2 Q(r),  L_j = -log c_j.

The evaluation works from w = K - hi(x), K = C + 0xfffff, which is 0xfffff - u: w >> 20 is -e
and bits 13..19 of w are 127 - j, so the table is stored from the last interval down (entry
k = 127 - j) and m = ldexp(x, -e) takes one instruction.

C puts 1.0 at the hi-word midpoint of its interval, whose entry is c = 1, L = 0: there
r = x - 1 exactly and log x = log1p(r) keeps its relative accuracy.  The other c_j are the
doubles nearest 2/(A_j + B_j); L_j is -log of that double, rounded.  Q is the Chebyshev
interpolant of (log1p(r) - r)/r^2 over the union of the r ranges (mpmath, 50 digits),
re-expanded in monomials.  Run: python gen_log_tab.py
"""
import os
import struct

import mpmath as mp

mp.mp.dps = 50
BITS = 7
DEG = 4


def hi_to_double(hi):
    return struct.unpack("<d", struct.pack("<Q", (hi & 0xFFFFFFFF) << 32))[0]


def main():
    one_hi = 0x3FF00000
    step = 1 << (20 - BITS)
    # C near hi(sqrt(1/2)) with hi(1.0) at the middle of an interval
    target = 0x3FE6A09E
    k = round((one_hi - step // 2 - target) / step)
    C = one_hi - step // 2 - k * step
    rows = []
    rmin, rmax = mp.mpf(0), mp.mpf(0)
    for j in range(1 << BITS):
        A = hi_to_double(C + j * step)
        B = hi_to_double(C + (j + 1) * step)
        if A <= 1.0 < B:
            c, L = 1.0, 0.0
        else:
            c = float(2 / (mp.mpf(A) + mp.mpf(B)))
            L = float(-mp.log(mp.mpf(c)))
        lo = mp.mpf(A) * mp.mpf(c) - 1
        hi = mp.mpf(B) * mp.mpf(c) - 1
        rmin, rmax = min(rmin, lo), max(rmax, hi)
        rows.append((c, L))
    f = lambda r: (mp.log1p(r) - r) / r ** 2 if r != 0 else mp.mpf(-0.5)
    n = DEG + 1
    a, b = rmin, rmax
    nodes = [(a + b) / 2 + (b - a) / 2 * mp.cos(mp.pi * (i + mp.mpf(1) / 2) / n) for i in range(n)]
    V = mp.matrix([[z ** p for p in range(n)] for z in nodes])
    coef = mp.lu_solve(V, mp.matrix([f(z) for z in nodes]))
    worst = mp.mpf(0)
    for i in range(4001):
        r = a + (b - a) * i / 4000
        if r == 0:
            continue
        q = sum(coef[p] * r ** p for p in range(n))
        exact = mp.log1p(r)
        worst = max(worst, abs((r + r * r * q) - exact) / abs(exact))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kf_log_tab.h")
    with open(path, "w") as fh:
        fh.write("// Generated by gen_log_tab.py -- do not edit.\n")
        fh.write("// Table-driven log: C = 0x%08X, r in [%s, %s], Q of degree %d, max rel. error of\n"
                 "// r + r^2 Q(r) against log1p(r): %s\n" % (C, mp.nstr(rmin, 6), mp.nstr(rmax, 6), DEG,
                                                             mp.nstr(worst, 3)))
        fh.write("#pragma once\n")
        fh.write("static constexpr unsigned KF_LOGT_C = 0x%08Xu;\n" % C)
        fh.write("static constexpr unsigned KF_LOGT_K = 0x%08Xu;  // C + 0xfffff\n" % (C + 0xFFFFF))
        fh.write("static constexpr int KF_LOGT_BITS = %d;\n" % BITS)
        fh.write("static constexpr int KF_LOGT_QN = %d;\n" % n)
        fh.write("static constexpr double KF_LOGT_LN2 = %s;\n" % mp.nstr(mp.log(2), 20))
        fh.write("static constexpr double KF_LOGT_Q[%d] = {\n" % n)
        fh.write(",\n".join("    %s" % mp.nstr(coef[p], 20, min_fixed=-1, max_fixed=-1) for p in range(n)))
        fh.write("\n};\n")
        fh.write("// entry k: c_j, -log(c_j) of interval j = %d - k\n" % ((1 << BITS) - 1))
        fh.write("static constexpr double KF_LOGT_TAB[%d] = {\n" % (2 << BITS))
        fh.write(",\n".join("    %r, %r" % (c, L) for c, L in reversed(rows)))
        fh.write("\n};\n")
    print("wrote", path, "C = 0x%08X" % C, "r range", mp.nstr(rmin, 6), mp.nstr(rmax, 6), "max rel err", mp.nstr(worst, 3))


if __name__ == "__main__":
    main()
