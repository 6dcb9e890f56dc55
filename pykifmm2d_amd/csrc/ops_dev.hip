// ops_dev.hip -- the translation operators of every (class, level) built on the GPU.
//
// The same operators as ops_host.cpp (reference pykifmm2d/fmm.py:249-347, class_fmm.py:188-250),
// in the LevelOps layout: the 40 M2L blocks, the L2L and M2M blocks and the factored
// check-to-equivalent inverse F diag(1/lambda) F^T.  Every block entry is one kernel
// evaluation at circle points the host computes (ops_host.cpp level_circles), so the device and
// host blocks agree to the kernel's rounding.  lambda is the DFT of the mean circulant diagonal of
// E; the host sums it in long double, here in double-double arithmetic (Dekker / Knuth two-sum
// and an fma two-product), with the cosines split into double-double on the host, so 1/lambda
// keeps its accuracy in the high Fourier modes, where lambda is smallest (1e-9 of the largest at
// p = 64) and the solve amplifies its error.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>

#include "kf_internal.h"
#include "ops_dev.h"
#include "ops_host.h"
#include "plan_lists.h"

namespace kf {

namespace {

struct DD {
    double hi, lo;
};
__device__ __forceinline__ DD two_sum(double a, double b) {
    const double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DD dd_add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    const double lo = s.lo + a.lo + b.lo;
    return two_sum(s.hi, lo);
}
__device__ __forceinline__ DD dd_add_d(DD a, double b) {
    DD s = two_sum(a.hi, b);
    return two_sum(s.hi, s.lo + a.lo);
}
__device__ __forceinline__ DD dd_mul(DD a, DD b) {
    const double p = a.hi * b.hi;
    const double e = fma(a.hi, b.hi, -p);
    return two_sum(p, e + (a.hi * b.lo + a.lo * b.hi));
}
__device__ __forceinline__ DD dd_scale(DD a, double s) {  // s a power of two or an exact small factor
    const double p = a.hi * s;
    const double e = fma(a.hi, s, -p);
    return two_sum(p, e + a.lo * s);
}
// 1 / a in double-double: one Newton step from the double reciprocal.
__device__ __forceinline__ DD dd_recip(DD a) {
    const double r = 1.0 / a.hi;
    // e = 1 - a r (double-double), r' = r + r e
    const DD ar = dd_mul(a, DD{r, 0.0});
    const DD e = dd_add(DD{1.0, 0.0}, DD{-ar.hi, -ar.lo});
    return dd_add(DD{r, 0.0}, dd_mul(DD{r, 0.0}, e));
}

// Circle points of op o: small (x, y) then large (x, y), P doubles each (first p used).
__device__ __forceinline__ const double* circ(const double* c, int o, int k, int P) { return c + ((size_t)o * 4 + k) * P; }

// Blocks 0-39 M2L, 40-43 L2L (from the parent level), 44-47 M2M (into this level); entries past
// p, and blocks a level does not use, are zero.  Grid: one thread per (op, block, i, j).
__global__ void k_ops_blocks(KfKernel ker, int n_ops, int L, int p, int P, const double* __restrict__ cw,
                             const double* __restrict__ circles, const int8_t* __restrict__ offs, double* __restrict__ ops) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t PP = (int64_t)P * P;
    if (e >= (int64_t)n_ops * 48 * PP) return;
    const int o = (int)(e / (48 * PP));
    const int blk = (int)((e / PP) % 48);
    const int i = (int)((e % PP) / P), j = (int)(e % P);
    const int l = o % L;
    double v = 0.0;
    if (l >= 1 && i < p && j < p) {
        const double W = cw[o / L];
        if (blk < 40) {  // M2LT[i][j] = G(source small_i + off * w, target small_j)
            const double w = ldexp(W, -l);
            const double* s = circ(circles, o, 0, P);
            const double* t = circ(circles, o, 1, P);
            v = ker.eval(s[i] + offs[2 * blk] * w, t[i] + offs[2 * blk + 1] * w, s[j], t[j]);
        } else {
            // M2MC of parent level lp: child q's small circle (level lp + 1) shifted by half a
            // child width against the parent's large circle (fmm.py:280-297)
            const bool l2l = blk < 44;
            const int q = (blk - 40) & 3;
            const int lp = l2l ? l - 1 : l;
            if ((l2l && l >= 2) || (!l2l && l <= L - 2)) {
                const int oc = o - l + lp + 1, opar = o - l + lp;
                const double wc = ldexp(W, -(lp + 1));
                const double* csx = circ(circles, oc, 0, P);
                const double* csy = circ(circles, oc, 1, P);
                const double* plx = circ(circles, opar, 2, P);
                const double* ply = circ(circles, opar, 3, P);
                // L2LB[q][i][j] = M2MC_{l-1}[i][q p + j]; M2MB[q][i][j] = M2MC_l[j][q p + i]
                const int ci = l2l ? j : i, pj = l2l ? i : j;
                const double ex = q < 2 ? csx[ci] - 0.5 * wc : csx[ci] + 0.5 * wc;
                const double ey = (q & 1) ? csy[ci] + 0.5 * wc : csy[ci] - 0.5 * wc;
                v = ker.eval(ex, ey, plx[pj], ply[pj]);
            }
        }
    }
    ops[(size_t)o * LevelOps::size(P) + (size_t)blk * PP + (size_t)i * P + j] = v;
}

// Check-to-equivalent factors of op o (one workgroup of P threads): the mean circulant diagonal
// g(d) of E[j][i] = G(small_i, large_j), its symmetric part, lambda_m = sum_d g(d) cos(2 pi m d / p)
// (double-double), the Fourier basis copied from `basis` (F then F^T) and 1/lambda in the
// basis's column order (constant, cos m and sin m for m < p/2, Nyquist).
__global__ void k_ops_c2e(KfKernel ker, int L, int p, int P, const double* __restrict__ circles,
                          const double* __restrict__ cosdd, const double* __restrict__ basis, double* __restrict__ ops) {
    const int o = blockIdx.x;
    const int l = o % L;
    double* out = ops + (size_t)o * LevelOps::size(P);
    const size_t PP = (size_t)P * P;
    for (size_t k = threadIdx.x; k < 2 * PP; k += blockDim.x) out[LevelOps::F(P) + k] = l >= 1 ? basis[k] : 0.0;
    __shared__ double ghi[64], glo[64];
    __shared__ double lhi[33];
    const int d = threadIdx.x;
    if (l >= 1 && d < p) {
        const double* sx = circ(circles, o, 0, P);
        const double* sy = circ(circles, o, 1, P);
        const double* lx = circ(circles, o, 2, P);
        const double* ly = circ(circles, o, 3, P);
        DD g{0.0, 0.0};
        for (int i = 0; i < p; ++i) {
            const int jj = (i + d) % p;
            g = dd_add_d(g, ker.eval(sx[i], sy[i], lx[jj], ly[jj]));
        }
        // g / p
        const double q = g.hi / p;
        const DD qp = dd_mul(DD{q, 0.0}, DD{(double)p, 0.0});
        const DD rem = dd_add(g, DD{-qp.hi, -qp.lo});
        const DD gp = two_sum(q, rem.hi / p);
        ghi[d] = gp.hi;
        glo[d] = gp.lo;
    }
    __syncthreads();
    const int nf = p / 2 + 1;
    if (l >= 1 && d < nf) {
        DD lam{0.0, 0.0};
        for (int k = 0; k < p; ++k) {
            const int k2 = (p - k) % p;
            const DD gs = dd_scale(dd_add(DD{ghi[k], glo[k]}, DD{ghi[k2], glo[k2]}), 0.5);
            const size_t c = ((size_t)d * p + k) * 2;
            lam = dd_add(lam, dd_mul(gs, DD{cosdd[c], cosdd[c + 1]}));
        }
        const DD r = dd_recip(lam);
        lhi[d] = r.hi + r.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* il = out + LevelOps::invlam(P);
        for (int c = 0; c < P; ++c) il[c] = 0.0;
        if (l >= 1) {
            int col = 0;
            il[col++] = lhi[0];
            for (int m = 1; 2 * m < p; ++m) {
                il[col++] = lhi[m];
                il[col++] = lhi[m];
            }
            if (p % 2 == 0) il[col++] = lhi[p / 2];
        }
    }
}

// Largest |entry| of the 40 M2L blocks of each op (one workgroup per op).
__global__ void k_m2l_maxabs(const double* __restrict__ ops, int P, double* __restrict__ mx) {
    const int o = blockIdx.x;
    const double* b = ops + (size_t)o * LevelOps::size(P);
    double m = 0.0;
    for (size_t k = threadIdx.x; k < (size_t)40 * P * P; k += blockDim.x) m = fmax(m, fabs(b[k]));
    __shared__ double sm[256];
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) mx[o] = sm[0] > 0.0 ? sm[0] : 1.0;
}

__global__ void k_m2l_f32(const double* __restrict__ ops, int n_ops, int L, int P, const double* __restrict__ sc,
                          float* __restrict__ o32) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)40 * P * P;
    if (e >= (int64_t)n_ops * per) return;
    const int o = (int)(e / per);
    const int64_t k = e % per;
    o32[e] = o % L == 0 ? 0.0f : (float)(ops[(size_t)o * LevelOps::size(P) + k] / sc[o]);
}

// FP16 split of the scaled M2L blocks: per op and block, two planes [m][k] (the transposed
// block, k padded with zeros to PK), hi = fp16(v) and lo = fp16(v - hi) of v = block / scale;
// hi + lo carries v to 2^-24 of the op's largest entry (subnormal halves included).
__global__ void k_m2l_f16(const double* __restrict__ ops, int n_ops, int L, int P, int PK,
                          const double* __restrict__ sc, _Float16* __restrict__ o16) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)40 * P * PK;  // one plane of one op
    if (e >= (int64_t)n_ops * per) return;
    const int o = (int)(e / per);
    const int64_t r = e % per;
    const int blk = (int)(r / ((int64_t)P * PK)), mk = (int)(r % ((int64_t)P * PK)), m = mk / PK, k = mk % PK;
    double v = 0.0;
    if (o % L != 0 && k < P) v = ops[(size_t)o * LevelOps::size(P) + (size_t)blk * P * P + (size_t)k * P + m] / sc[o];
    const _Float16 hi = (_Float16)(float)v;
    const _Float16 lo = (_Float16)(float)(v - (double)(float)hi);
    _Float16* base = o16 + ((size_t)o * 40 + blk) * 2 * (size_t)P * PK;
    base[mk] = hi;
    base[(size_t)P * PK + mk] = lo;
}

// Host tables that depend on p only: cos(2 pi m d / p) as double-double pairs for m <= p/2,
// and the real Fourier basis F (P x P, then F^T), both from long double.
struct PTables {
    std::vector<double> cosdd, basis;
};
const PTables& p_tables(int p, int P) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, PTables> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({p, P});
    if (it != cache.end()) return it->second;
    PTables t;
    const long double two_pi = 6.283185307179586476925286766559L;
    auto cosk = [&](long long m) { return cosl(two_pi * (long double)(m % p) / p); };
    auto sink = [&](long long m) { return sinl(two_pi * (long double)(m % p) / p); };
    const int nf = p / 2 + 1;
    t.cosdd.resize((size_t)nf * p * 2);
    for (int m = 0; m < nf; ++m)
        for (int d = 0; d < p; ++d) {
            const long double v = cosk((long long)m * d);
            const double hi = (double)v;
            t.cosdd[((size_t)m * p + d) * 2] = hi;
            t.cosdd[((size_t)m * p + d) * 2 + 1] = (double)(v - (long double)hi);
        }
    t.basis.assign((size_t)2 * P * P, 0.0);
    double* F = t.basis.data();
    double* FT = F + (size_t)P * P;
    const long double c0 = 1.0L / sqrtl((long double)p), c1 = sqrtl(2.0L / p);
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
    };
    put(col++, 0, 0);
    for (int m = 1; 2 * m < p; ++m) {
        put(col++, m, 0);
        put(col++, m, 1);
    }
    if (p % 2 == 0) put(col++, p / 2, 2);
    return cache.emplace(std::make_pair(p, P), std::move(t)).first->second;
}

inline unsigned nblk(int64_t n, int t) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

}  // namespace

void build_ops16_dev(const DevBuf<double>& ops, int n_ops, int L, int P, const DevBuf<double>& scale,
                     DevBuf<_Float16>& ops16, hipStream_t s) {
    const int PK = (P + 31) / 32 * 32;
    ops16.alloc((size_t)n_ops * 40 * 2 * P * PK);
    const int64_t n = (int64_t)n_ops * 40 * P * PK;
    k_m2l_f16<<<nblk(n, 256), 256, 0, s>>>(ops.p, n_ops, L, P, PK, scale.p, ops16.p);
    KF_HIP_CHECK(hipGetLastError());
}

void build_ops_dev(const KfKernel& ker, const std::vector<double>& class_width, int L, int p, int P,
                   DevBuf<double>& ops, DevBuf<double>& surf, hipStream_t s) {
    const int nc = (int)class_width.size(), n_ops = nc * L;
    // circles of every op (host, as the host operators use them)
    std::vector<double> circles((size_t)n_ops * 4 * P, 0.0), hs((size_t)n_ops * 2 * P, 0.0);
    for (int c = 0; c < nc; ++c)
        for (int l = 0; l < L; ++l) {
            const int o = c * L + l;
            std::vector<double> a, b, lx, ly;
            level_circles(std::ldexp(class_width[c], -l), p, a, b, lx, ly);
            for (int j = 0; j < p; ++j) {
                circles[((size_t)o * 4 + 0) * P + j] = a[j];
                circles[((size_t)o * 4 + 1) * P + j] = b[j];
                circles[((size_t)o * 4 + 2) * P + j] = lx[j];
                circles[((size_t)o * 4 + 3) * P + j] = ly[j];
                hs[(size_t)(2 * o) * P + j] = lx[j];
                hs[(size_t)(2 * o + 1) * P + j] = ly[j];
            }
        }
    std::vector<int8_t> offs(80);
    for (int k = 0; k < 40; ++k) {
        int ox, oy;
        m2l_offset_of(k, ox, oy);
        offs[2 * k] = (int8_t)ox;
        offs[2 * k + 1] = (int8_t)oy;
    }
    const PTables& pt = p_tables(p, P);
    DevBuf<double> d_circ, d_cw, d_cos, d_basis;
    DevBuf<int8_t> d_offs;
    d_circ.upload(circles, s);
    d_cw.upload(class_width, s);
    d_cos.upload(pt.cosdd, s);
    d_basis.upload(pt.basis, s);
    d_offs.upload(offs, s);
    surf.upload(hs, s);
    ops.alloc((size_t)n_ops * LevelOps::size(P));
    const int64_t nb = (int64_t)n_ops * 48 * P * P;
    k_ops_blocks<<<nblk(nb, 256), 256, 0, s>>>(ker, n_ops, L, p, P, d_cw.p, d_circ.p, d_offs.p, ops.p);
    k_ops_c2e<<<n_ops, 64, 0, s>>>(ker, L, p, P, d_circ.p, d_cos.p, d_basis.p, ops.p);
    KF_HIP_CHECK(hipGetLastError());
    KF_HIP_CHECK(hipStreamSynchronize(s));  // the host arrays above are the uploads' sources
}

void build_ops32_dev(const DevBuf<double>& ops, int n_ops, int L, int P, DevBuf<float>& ops32, DevBuf<double>& scale,
                     hipStream_t s) {
    scale.alloc(n_ops);
    ops32.alloc((size_t)n_ops * 40 * P * P);
    k_m2l_maxabs<<<n_ops, 256, 0, s>>>(ops.p, P, scale.p);
    const int64_t n = (int64_t)n_ops * 40 * P * P;
    k_m2l_f32<<<nblk(n, 256), 256, 0, s>>>(ops.p, n_ops, L, P, scale.p, ops32.p);
    KF_HIP_CHECK(hipGetLastError());
}

}  // namespace kf
