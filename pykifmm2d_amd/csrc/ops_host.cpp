// ops_host.cpp -- FP64 translation operators per level (host).
//
// The operators are the reference's (pykifmm2d/fmm.py:249-347, class_fmm.py:188-250):
// check-to-equivalent E, the children-to-parent M2MC and the 40 M2L blocks.  E is a
// symmetric circulant matrix (both circles carry p equispaced points at the same angles and
// G is radial), so instead of an LU factorisation the solve uses its eigen-decomposition in
// the real Fourier basis, E = F diag(lambda) F^T, applied as two GEMMs and a diagonal scale
// on the device.  lambda is the DFT of the matrix's mean circulant diagonal, accumulated in
// long double.
#include "ops_host.h"

#include <cmath>
#include <cstring>

namespace kf {

void level_circles(double width, int p, std::vector<double>& sx, std::vector<double>& sy,
                   std::vector<double>& lx, std::vector<double>& ly) {
    const double dd = 0.01;
    double r1 = 0.5 * width * (std::sqrt(2.0) + dd);
    double r2 = 0.5 * width * (4.0 - std::sqrt(2.0) - 2.0 * dd);
    double step = (2.0 * KF_PI) / p;
    sx.resize(p);
    sy.resize(p);
    lx.resize(p);
    ly.resize(p);
    for (int k = 0; k < p; ++k) {
        double th = k * step, c = std::cos(th), s = std::sin(th);
        sx[k] = r1 * c;
        sy[k] = r1 * s;
        lx[k] = r2 * c;
        ly[k] = r2 * s;
    }
}

void c2e_matrix(const KfKernel& ker, double width, int p, std::vector<double>& E) {
    std::vector<double> sx, sy, lx, ly;
    level_circles(width, p, sx, sy, lx, ly);
    E.assign((size_t)p * p, 0.0);
    for (int j = 0; j < p; ++j)
        for (int i = 0; i < p; ++i) E[(size_t)j * p + i] = ker.eval(sx[i], sy[i], lx[j], ly[j]);
}

void build_level_ops(const KfKernel& ker, double W, int l, int p, int P, bool has_parent,
                     bool has_children, double* out) {
    std::memset(out, 0, LevelOps::size(P) * sizeof(double));
    double w = std::ldexp(W, -l);
    std::vector<double> sx, sy, lx, ly;
    level_circles(w, p, sx, sy, lx, ly);
    // 40 M2L blocks, stored transposed: M2LT[i][j] = G(source small_i + o w, target small_j)
    for (int o = 0; o < 40; ++o) {
        int ox, oy;
        m2l_offset_of(o, ox, oy);
        double* B = out + LevelOps::m2lt(P, o);
        for (int i = 0; i < p; ++i) {
            double ax = sx[i] + ox * w, ay = sy[i] + oy * w;
            for (int j = 0; j < p; ++j) B[(size_t)i * P + j] = ker.eval(ax, ay, sx[j], sy[j]);
        }
    }
    // children small circles (at +-w_child/2) against this level's large circle
    // M2MC of parent level lp (p x 4p): child q's small circle, shifted by (-,-), (-,+),
    // (+,-), (+,+) half child widths, against the parent's large circle (fmm.py:280-297)
    auto m2mc = [&](int lp, std::vector<double>& M) {
        double wp = std::ldexp(W, -lp), wc = std::ldexp(W, -(lp + 1));
        std::vector<double> csx, csy, clx, cly, psx, psy, plx, ply;
        level_circles(wc, p, csx, csy, clx, cly);
        level_circles(wp, p, psx, psy, plx, ply);
        M.assign((size_t)p * 4 * p, 0.0);
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < p; ++i) {
                double ex = (q < 2) ? csx[i] - 0.5 * wc : csx[i] + 0.5 * wc;
                double ey = (q & 1) ? csy[i] + 0.5 * wc : csy[i] - 0.5 * wc;
                for (int j = 0; j < p; ++j) M[(size_t)j * 4 * p + q * p + i] = ker.eval(ex, ey, plx[j], ply[j]);
            }
    };
    std::vector<double> M;
    if (has_parent) {  // L2L from level l-1: child check (small_j of child q) from parent large_i
        m2mc(l - 1, M);
        for (int q = 0; q < 4; ++q) {
            double* B = out + LevelOps::l2lb(P, q);
            for (int i = 0; i < p; ++i)
                for (int j = 0; j < p; ++j) B[(size_t)i * P + j] = M[(size_t)i * 4 * p + q * p + j];
        }
    }
    if (has_children) {  // M2M into level l: parent check j from child q equivalent i
        m2mc(l, M);
        for (int q = 0; q < 4; ++q) {
            double* B = out + LevelOps::m2mb(P, q);
            for (int i = 0; i < p; ++i)
                for (int j = 0; j < p; ++j) B[(size_t)i * P + j] = M[(size_t)j * 4 * p + q * p + i];
        }
    }
    // C2E: E[j][i] = G(small_i, large_j) is symmetric circulant; g(d) = mean diagonal d
    std::vector<long double> g(p, 0.0L);
    for (int i = 0; i < p; ++i)
        for (int d = 0; d < p; ++d) g[d] += (long double)ker.eval(sx[i], sy[i], lx[(i + d) % p], ly[(i + d) % p]);
    for (int d = 0; d < p; ++d) g[d] /= p;
    std::vector<long double> gs(p);
    for (int d = 0; d < p; ++d) gs[d] = 0.5L * (g[d] + g[(p - d) % p]);
    const long double two_pi = 6.283185307179586476925286766559L;
    auto cosk = [&](long long m) { return cosl(two_pi * (long double)(m % p) / p); };
    auto sink = [&](long long m) { return sinl(two_pi * (long double)(m % p) / p); };
    int nfreq = p / 2 + 1;
    std::vector<long double> lam(nfreq, 0.0L);
    for (int m = 0; m < nfreq; ++m)
        for (int d = 0; d < p; ++d) lam[m] += gs[d] * cosk((long long)m * d);
    double* F = out + LevelOps::F(P);
    double* FT = out + LevelOps::FT(P);
    double* il = out + LevelOps::invlam(P);
    long double c0 = 1.0L / sqrtl((long double)p), c1 = sqrtl(2.0L / p);
    int col = 0;
    auto put = [&](int c, int m, int kind) {  // kind 0: constant / cos, 1: sin, 2: Nyquist
        for (int j = 0; j < p; ++j) {
            long double v;
            if (m == 0) v = c0;
            else if (kind == 2) v = (j & 1) ? -c0 : c0;
            else if (kind == 0) v = c1 * cosk((long long)m * j);
            else v = c1 * sink((long long)m * j);
            F[(size_t)j * P + c] = (double)v;
            FT[(size_t)c * P + j] = (double)v;
        }
        il[c] = (double)(1.0L / lam[m]);
    };
    put(col++, 0, 0);
    for (int m = 1; 2 * m < p; ++m) {
        put(col++, m, 0);
        put(col++, m, 1);
    }
    if (p % 2 == 0) put(col++, p / 2, 2);
}

// Eigenvalues of the symmetric circulant C2E matrix of one real kernel part, by mode
// m = 0 .. p/2 (DFT of the mean circulant diagonal, long double), as build_level_ops.
template <class G>
static std::vector<long double> circulant_eigs(int p, const std::vector<double>& sx, const std::vector<double>& sy,
                                               const std::vector<double>& lx, const std::vector<double>& ly,
                                               G&& gfun) {
    std::vector<long double> g(p, 0.0L);
    for (int i = 0; i < p; ++i)
        for (int d = 0; d < p; ++d) g[d] += (long double)gfun(sx[i], sy[i], lx[(i + d) % p], ly[(i + d) % p]);
    for (int d = 0; d < p; ++d) g[d] /= p;
    std::vector<long double> gs(p);
    for (int d = 0; d < p; ++d) gs[d] = 0.5L * (g[d] + g[(p - d) % p]);
    const long double two_pi = 6.283185307179586476925286766559L;
    const int nfreq = p / 2 + 1;
    std::vector<long double> lam(nfreq, 0.0L);
    for (int m = 0; m < nfreq; ++m)
        for (int d = 0; d < p; ++d) lam[m] += gs[d] * cosl(two_pi * (long double)(((long long)m * d) % p) / p);
    return lam;
}

void build_level_ops_complex(const KfKernel& ker, double W, int l, int p, int Ph, bool has_parent,
                             bool has_children, double* out) {
    const int P = 2 * Ph;
    std::memset(out, 0, LevelOps::size(P) * sizeof(double));
    double w = std::ldexp(W, -l);
    std::vector<double> sx, sy, lx, ly;
    level_circles(w, p, sx, sy, lx, ly);
    // block B (P x P) of the complex p x p operator M = Mr + i Mi in the row convention y = x M
    // on rows [re | im]: [[Mr, Mi], [-Mi, Mr]]
    auto put = [&](double* B, int i, int j, double re, double im) {
        B[(size_t)i * P + j] = re;
        B[(size_t)i * P + Ph + j] = im;
        B[(size_t)(Ph + i) * P + j] = -im;
        B[(size_t)(Ph + i) * P + Ph + j] = re;
    };
    for (int o = 0; o < 40; ++o) {
        int ox, oy;
        m2l_offset_of(o, ox, oy);
        double* B = out + LevelOps::m2lt(P, o);
        for (int i = 0; i < p; ++i) {
            double ax = sx[i] + ox * w, ay = sy[i] + oy * w;
            for (int j = 0; j < p; ++j) put(B, i, j, ker.eval(ax, ay, sx[j], sy[j]), ker.eval_im(ax, ay, sx[j], sy[j]));
        }
    }
    // M2MC of parent level lp as in build_level_ops, real and imaginary parts
    auto m2mc = [&](int lp, std::vector<double>& Mr, std::vector<double>& Mi) {
        double wp = std::ldexp(W, -lp), wc = std::ldexp(W, -(lp + 1));
        std::vector<double> csx, csy, clx, cly, psx, psy, plx, ply;
        level_circles(wc, p, csx, csy, clx, cly);
        level_circles(wp, p, psx, psy, plx, ply);
        Mr.assign((size_t)p * 4 * p, 0.0);
        Mi.assign((size_t)p * 4 * p, 0.0);
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < p; ++i) {
                double ex = (q < 2) ? csx[i] - 0.5 * wc : csx[i] + 0.5 * wc;
                double ey = (q & 1) ? csy[i] + 0.5 * wc : csy[i] - 0.5 * wc;
                for (int j = 0; j < p; ++j) {
                    Mr[(size_t)j * 4 * p + q * p + i] = ker.eval(ex, ey, plx[j], ply[j]);
                    Mi[(size_t)j * 4 * p + q * p + i] = ker.eval_im(ex, ey, plx[j], ply[j]);
                }
            }
    };
    std::vector<double> Mr, Mi;
    if (has_parent) {
        m2mc(l - 1, Mr, Mi);
        for (int q = 0; q < 4; ++q) {
            double* B = out + LevelOps::l2lb(P, q);
            for (int i = 0; i < p; ++i)
                for (int j = 0; j < p; ++j)
                    put(B, i, j, Mr[(size_t)i * 4 * p + q * p + j], Mi[(size_t)i * 4 * p + q * p + j]);
        }
    }
    if (has_children) {
        m2mc(l, Mr, Mi);
        for (int q = 0; q < 4; ++q) {
            double* B = out + LevelOps::m2mb(P, q);
            for (int i = 0; i < p; ++i)
                for (int j = 0; j < p; ++j)
                    put(B, i, j, Mr[(size_t)j * 4 * p + q * p + i], Mi[(size_t)j * 4 * p + q * p + i]);
        }
    }
    // C2E: E = Er + i Ei with Er, Ei real symmetric circulants, both diagonal in the real
    // Fourier basis F: E^{-1} = F diag(d) F^T, d = 1 / (mu_r + i mu_i).  The device applies
    // (x F) * s then * FT; here F is blockdiag(F, F), s = 1 on the p columns of each half, and
    // FT carries diag(d) F^T, realified.
    const std::vector<long double> mr = circulant_eigs(p, sx, sy, lx, ly, [&](double a, double b, double c, double d) {
        return ker.eval(a, b, c, d);
    });
    const std::vector<long double> mi = circulant_eigs(p, sx, sy, lx, ly, [&](double a, double b, double c, double d) {
        return ker.eval_im(a, b, c, d);
    });
    const long double two_pi = 6.283185307179586476925286766559L;
    auto cosk = [&](long long m) { return cosl(two_pi * (long double)(m % p) / p); };
    auto sink = [&](long long m) { return sinl(two_pi * (long double)(m % p) / p); };
    double* F = out + LevelOps::F(P);
    double* FT = out + LevelOps::FT(P);
    double* il = out + LevelOps::invlam(P);
    long double c0 = 1.0L / sqrtl((long double)p), c1 = sqrtl(2.0L / p);
    int col = 0;
    auto basis = [&](int c, int m, int kind) {  // kind 0: constant / cos, 1: sin, 2: Nyquist
        const long double den = mr[m] * mr[m] + mi[m] * mi[m];
        const long double dr = mr[m] / den, di = -mi[m] / den;  // d = 1 / (mu_r + i mu_i)
        for (int j = 0; j < p; ++j) {
            long double v;
            if (m == 0) v = c0;
            else if (kind == 2) v = (j & 1) ? -c0 : c0;
            else if (kind == 0) v = c1 * cosk((long long)m * j);
            else v = c1 * sink((long long)m * j);
            F[(size_t)j * P + c] = (double)v;
            F[(size_t)(Ph + j) * P + Ph + c] = (double)v;
            // (diag(d) F^T)[c][j] = d_c F[j][c], realified
            FT[(size_t)c * P + j] = (double)(dr * v);
            FT[(size_t)c * P + Ph + j] = (double)(di * v);
            FT[(size_t)(Ph + c) * P + j] = (double)(-di * v);
            FT[(size_t)(Ph + c) * P + Ph + j] = (double)(dr * v);
        }
        il[c] = 1.0;
        il[Ph + c] = 1.0;
    };
    basis(col++, 0, 0);
    for (int m = 1; 2 * m < p; ++m) {
        basis(col++, m, 0);
        basis(col++, m, 1);
    }
    if (p % 2 == 0) basis(col++, p / 2, 2);
}

}  // namespace kf
