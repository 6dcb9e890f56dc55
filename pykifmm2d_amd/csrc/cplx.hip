// cplx.hip -- the point passes of a plan of a complex kernel (cplx.h).  Straightforward VALU
// forms: one thread per check point or target, each kernel value's two parts from one log.
// The reference's complex FMM is class_fmm.py with dtype complex (class_fmm.py:167, 174): its
// numba loops (class_fmm.py:58-113, fmm.py:30-64 and 91-115) with a complex Kernel_Eval.
#include <algorithm>

#include "cplx.h"
#include "kf_internal.h"
#include "kf_kerlds.h"

namespace kf {

struct CKer {
    KfTable re, im;
    LogtCoef lc = KF_LOGT_COEF;
    // both parts at squared distance r2 (the tables' g0 at r2 = 0)
    __device__ __forceinline__ void g(double r2, const double* logt, double& gr, double& gi) const {
        if (r2 == 0.0) {
            gr = re.g0;
            gi = im.g0;
            return;
        }
        const double u = kf_logt(r2, logt, lc) * KF_INV_LN2;
        gr = ktab_poly(re, u);
        gi = ktab_poly(im, u);
    }
};

__device__ __forceinline__ void stage_logt_c(double* lds) {
    for (int i = threadIdx.x; i < (2 << KF_LOGT_BITS); i += blockDim.x) lds[i] = KF_LOGT_TAB[i];
}

// P2M (class_fmm.py:58-71): check point k of box b, right-hand side r:
// U[row][r][k] = sum_s tau[s][r] Gr, U[row][r][Ph + k] = sum_s tau[s][r] Gi (zero padding).
__global__ void __launch_bounds__(64) k_p2m_c(CKer ker, CplxArgs a) {
    __shared__ __attribute__((aligned(16))) double logt[2 << KF_LOGT_BITS];
    stage_logt_c(logt);
    __syncthreads();
    const int32_t b = a.p2m[blockIdx.x];
    const int Ph = a.P / 2, l = a.bop[b];
    const int32_t b0 = a.bot[b], b1 = a.top[b];
    for (int k = threadIdx.x; k < Ph; k += blockDim.x) {
        const bool on = k < a.p;
        const double tx = on ? a.cx[b] + a.surf[(size_t)(2 * l) * a.P + k] : 0.0;
        const double ty = on ? a.cy[b] + a.surf[(size_t)(2 * l + 1) * a.P + k] : 0.0;
        for (int r = 0; r < a.R; ++r) {
            double ar = 0.0, ai = 0.0;
            if (on)
                for (int32_t s = b0; s < b1; ++s) {
                    const double dx = tx - a.xs[s], dy = ty - a.ys[s];
                    double gr, gi;
                    ker.g(fma(dx, dx, dy * dy), logt, gr, gi);
                    const double t = a.tau[(size_t)s * a.R + r];
                    ar = fma(gr, t, ar);
                    ai = fma(gi, t, ai);
                }
            double* row = a.U + ((size_t)a.rid[b] * a.R + r) * a.P;
            row[k] = ar;
            row[Ph + k] = ai;
        }
    }
}

// Near field (fmm.py:30-64): target t of leaf li, every source of its colleagues (the target
// itself skipped), complex sums into near[2 (t R + r) + c].
__global__ void __launch_bounds__(64) k_p2p_c(CKer ker, CplxArgs a, int lo, int n) {
    __shared__ __attribute__((aligned(16))) double logt[2 << KF_LOGT_BITS];
    stage_logt_c(logt);
    __syncthreads();
    for (int li = lo + blockIdx.x; li < lo + n; li += gridDim.x) {
        const int32_t b = a.leaves[li];
        const int32_t tb = a.bot[b], te = a.top[b];
        for (int32_t t = tb + threadIdx.x; t < te; t += blockDim.x) {
            const double x = a.xs[t], y = a.ys[t];
            for (int r = 0; r < a.R; ++r) {
                double ar = 0.0, ai = 0.0;
                for (int j = 0; j < 9; ++j) {
                    const int32_t c = a.leaf_coll[(size_t)li * 9 + j];
                    if (c < 0) continue;
                    for (int32_t s = a.bot[c]; s < a.top[c]; ++s) {
                        if (s == t) continue;
                        const double dx = x - a.xs[s], dy = y - a.ys[s];
                        double gr, gi;
                        ker.g(fma(dx, dx, dy * dy), logt, gr, gi);
                        const double v = a.tau[(size_t)s * a.R + r];
                        ar = fma(gr, v, ar);
                        ai = fma(gi, v, ai);
                    }
                }
                a.near[2 * ((size_t)t * a.R + r)] = ar;
                a.near[2 * ((size_t)t * a.R + r) + 1] = ai;
            }
        }
    }
}

struct OutPtrs {
    double* p[64];
};

// L2P (fmm.py:91-115) of leaf li plus its near field, written in input order: out[r][2 perm[t] + c].
__global__ void __launch_bounds__(64) k_l2p_c(CKer ker, CplxArgs a, OutPtrs out) {
    __shared__ __attribute__((aligned(16))) double logt[2 << KF_LOGT_BITS];
    stage_logt_c(logt);
    __syncthreads();
    const int32_t b = a.leaves[blockIdx.x];
    const bool far = a.lvl[b] >= 2;
    const int Ph = a.P / 2, l = a.bop[b];
    const double bx = a.cx[b], by = a.cy[b];
    for (int32_t t = a.bot[b] + threadIdx.x; t < a.top[b]; t += blockDim.x) {
        const double rx = a.xs[t] - bx, ry = a.ys[t] - by;
        const int32_t o = a.perm[t];
        for (int r = 0; r < a.nvalid; ++r) {
            double ur = 0.0, ui = 0.0;
            if (far) {
                const double* ps = a.Psi + ((size_t)a.rid[b] * a.R + r) * a.P;
                for (int k = 0; k < a.p; ++k) {
                    const double dx = rx - a.surf[(size_t)(2 * l) * a.P + k];
                    const double dy = ry - a.surf[(size_t)(2 * l + 1) * a.P + k];
                    double gr, gi;
                    ker.g(fma(dx, dx, dy * dy), logt, gr, gi);
                    const double pr = ps[k], pi = ps[Ph + k];
                    ur = fma(gr, pr, fma(-gi, pi, ur));
                    ui = fma(gr, pi, fma(gi, pr, ui));
                }
            }
            out.p[r][2 * (size_t)o] = ur + a.near[2 * ((size_t)t * a.R + r)];
            out.p[r][2 * (size_t)o + 1] = ui + a.near[2 * ((size_t)t * a.R + r) + 1];
        }
    }
}

// class_fmm.py:379-399 evaluate_to_points at arbitrary targets (class_fmm.py:73-113): the leaf's
// local expansion, then its colleagues' sources (check: skip coincident points).
__global__ void __launch_bounds__(256) k_targets_c(CKer ker, CplxArgs a, const double* __restrict__ tx,
                                                   const double* __restrict__ ty, int64_t m,
                                                   const int32_t* __restrict__ loc,
                                                   const int32_t* __restrict__ coll, int check,
                                                   double* __restrict__ pot) {
    __shared__ __attribute__((aligned(16))) double logt[2 << KF_LOGT_BITS];
    stage_logt_c(logt);
    __syncthreads();
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int r = blockIdx.y, Ph = a.P / 2;
    const double x = tx[i], y = ty[i];
    const int32_t b = loc[i];
    const int l = a.bop[b];
    double ur = 0.0, ui = 0.0;
    if (a.lvl[b] >= 2) {
        const double rx = x - a.cx[b], ry = y - a.cy[b];
        const double* ps = a.Psi + ((size_t)a.rid[b] * a.R + r) * a.P;
        for (int k = 0; k < a.p; ++k) {
            const double dx = rx - a.surf[(size_t)(2 * l) * a.P + k];
            const double dy = ry - a.surf[(size_t)(2 * l + 1) * a.P + k];
            double gr, gi;
            ker.g(fma(dx, dx, dy * dy), logt, gr, gi);
            ur = fma(gr, ps[k], fma(-gi, ps[Ph + k], ur));
            ui = fma(gr, ps[Ph + k], fma(gi, ps[k], ui));
        }
    }
    for (int slot = 0; slot < 9; ++slot) {
        const int32_t c = coll[(size_t)b * 9 + slot];
        if (c < 0) continue;
        for (int32_t s = a.bot[c]; s < a.top[c]; ++s) {
            const double dx = x - a.xs[s], dy = y - a.ys[s];
            if (check && dx == 0.0 && dy == 0.0) continue;
            double gr, gi;
            ker.g(fma(dx, dx, dy * dy), logt, gr, gi);
            const double v = a.tau[(size_t)s * a.R + r];
            ur = fma(gr, v, ur);
            ui = fma(gi, v, ui);
        }
    }
    pot[2 * ((size_t)r * m + i)] = ur;
    pot[2 * ((size_t)r * m + i) + 1] = ui;
}

static CKer cker_of(const KfKernel& k) {
    CKer c;
    c.re = k.td;
    c.im = k.td_im;
    return c;
}

void launch_cplx(const KfKernel& ker, const CplxArgs& a, int which, int lo, int hi, double* const* out,
                 hipStream_t s) {
    const CKer c = cker_of(ker);
    if (which == 0 && a.n_p2m) k_p2m_c<<<a.n_p2m, 64, 0, s>>>(c, a);
    if (which == 1 && hi > lo) k_p2p_c<<<hi - lo, 64, 0, s>>>(c, a, lo, hi - lo);
    if (which == 2 && a.n_leaves) {
        OutPtrs o{};
        for (int r = 0; r < a.nvalid; ++r) o.p[r] = out[r];
        k_l2p_c<<<a.n_leaves, 64, 0, s>>>(c, a, o);
    }
    KF_HIP_CHECK(hipGetLastError());
}

void launch_cplx_targets(const KfKernel& ker, const CplxArgs& a, const double* tx, const double* ty, int64_t m,
                         const int32_t* loc, const int32_t* coll, int check, double* pot, hipStream_t s) {
    const unsigned nb = (unsigned)((m + 255) / 256);
    k_targets_c<<<dim3(nb, (unsigned)a.nvalid), 256, 0, s>>>(cker_of(ker), a, tx, ty, m, loc, coll, check, pot);
    KF_HIP_CHECK(hipGetLastError());
}

}  // namespace kf
