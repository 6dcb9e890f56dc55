"""Generate k0_cheb.h: coefficients for the modified Bessel function K0 used on device.

Method (textbook, Abramowitz & Stegun 9.6.13 and a Chebyshev economisation):
  * 0 < x <= X0: K0(x) = -(ln(x/2) + gamma) I0(x) + sum_{k>=1} q^k H_k / (k!)^2,
                 I0(x) = sum_{k>=0} q^k / (k!)^2,  q = x^2/4, H_k = 1 + 1/2 + ... + 1/k.
                 Both power series are truncated where the next term is below 1e-18.
  * x > X0:      K0(x) = exp(-x)/sqrt(x) * f(u),  u = 2*X0/x - 1 in (-1, 1],
                 f = sqrt(x) e^x K0(x) expanded in Chebyshev polynomials T_n(u); the
                 coefficients come from mpmath at 40 digits and are truncated below 1e-18.
Run:  python gen_k0_cheb.py  (writes k0_cheb.h next to this file).
"""
import os

import mpmath as mp

mp.mp.dps = 40
X0 = 1  # split point between the power series and the Chebyshev expansion


def series_coeffs(tol=mp.mpf("1e-19")):
    a, s = [], []
    k = 0
    fact2 = mp.mpf(1)
    h = mp.mpf(0)
    while True:
        if k > 0:
            fact2 *= k * k
            h += mp.mpf(1) / k
        term = 1 / fact2
        a.append(term)
        s.append(term * h)
        if term < tol and k > 2:
            break
        k += 1
    return a, s


def cheb_coeffs(nmax=80, tol=mp.mpf("1e-19")):
    # f(u) = sqrt(x) e^x K0(x) with x = 2*X0/(u+1); Chebyshev-Gauss nodes on [-1, 1]
    n = nmax
    nodes = [mp.cos(mp.pi * (j + mp.mpf(1) / 2) / n) for j in range(n)]
    vals = []
    for u in nodes:
        x = 2 * mp.mpf(X0) / (u + 1)
        vals.append(mp.sqrt(x) * mp.exp(x) * mp.besselk(0, x))
    c = []
    for m in range(n):
        acc = mp.mpf(0)
        for j in range(n):
            acc += vals[j] * mp.cos(mp.pi * m * (j + mp.mpf(1) / 2) / n)
        c.append(2 * acc / n)
    c[0] /= 2
    last = max(i for i, v in enumerate(c) if abs(v) > tol)
    return c[: last + 1]


def emit(name, vals):
    body = ",\n".join("    %s" % mp.nstr(v, 20, min_fixed=-1, max_fixed=-1) for v in vals)
    return "static constexpr int %s_N = %d;\nstatic constexpr double %s[%d] = {\n%s\n};\n" % (
        name, len(vals), name, len(vals), body)


def main():
    a, s = series_coeffs()
    c = cheb_coeffs()
    out = ["// Generated by gen_k0_cheb.py -- do not edit.",
           "// K0 coefficients: power series on (0, X0], Chebyshev series in u = 2*X0/x - 1 on (X0, inf).",
           "#pragma once", "",
           "static constexpr double KF_K0_SPLIT = %r;" % float(X0), "",
           emit("KF_I0_SERIES", a), emit("KF_K0_SERIES_H", s), emit("KF_K0_CHEB", c)]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "k0_cheb.h")
    with open(path, "w") as fh:
        fh.write("\n".join(out))
    print("wrote", path, len(a), len(s), len(c))


if __name__ == "__main__":
    main()
