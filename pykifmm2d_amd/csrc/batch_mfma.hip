// batch_mfma.hip -- near field, P2M and L2P of batches of 16, 32 or 64 right-hand sides on the
// FP64 MFMA (the kernels behind launch_batch_mfma, batch_mfma.h).
#include <algorithm>

#include "batch_mfma.h"
#include "kf_internal.h"
#include "kf_kerlds.h"
#include "kf_mfma.h"
#include "kf_xcd.h"

namespace kf {

// ---------------------------------------------------------- batched forms on the MFMA ----
// For batches of R = RB x 16 right-hand sides (16, 32 or 64) the near field, P2M and L2P form
// each kernel value once on the vector ALU and contract it with the right-hand sides on FP64
// MFMA: v_mfma_f64_16x16x4 computes D[i][j] += sum_k A[i][k] B[k][j] with lane l supplying
// A[l%16][l/16] and B[l/16][l%16] and holding D[4v + l/16][l%16] in register v (kf_mfma.h).
// Lane l of a step forms the kernel value of its own (l%16, l/16) pair, so no value is formed
// twice.  The kernel values are bitwise those of the vector forms; the sums run in MFMA order
// over steps of 4 sources (or equivalent points) taken in a fixed order, so a right-hand side
// gets bitwise the same potential at every width >= 16 and, to rounding, the vector forms'.
//
// One wave per work item (a leaf, or a box and a pair of check-point blocks), 64-thread
// workgroups, no workgroup-wide staging: each wave walks its sources on its own, reading the
// density rows (or Psi) straight from the L2 one step ahead of their use, so that the loads of
// step s + 1 are in flight while step s forms its kernel values and runs its MFMAs.  Column j
// of MFMA block rb is right-hand side RB j + rb: a lane's RB values of one row are contiguous
// (one 16- or 32-byte access) and 16 lanes cover a whole R-wide row.  A wave holds at most two
// 16-row blocks of targets (or check points) per pass, sharing each step's density loads.
// Padding lanes (past the last source, target or check point) evaluate the kernel at r^2 = 1
// and contribute exact zeros; every load reads a clamped, in-range address.
constexpr int KF_MW_CHUNK = 128;  // near-field sources listed in LDS per pass
constexpr int KF_MFMA_CHUNK = 64;  // P2M sources staged per pass
constexpr int KF_MW_RUN = 16;     // neighbouring work items per XCD run (xcd_chunked)
// 16-row blocks per wave pass: two share each step's loads; K0's evaluation takes too many
// registers to hold two blocks' accumulators beside it.
template <int KIND>
constexpr int mw_blocks() {
    return KIND == KF_MODHELM ? 1 : 2;
}

// RB contiguous doubles at p (RB * 8-byte aligned).
template <int RB>
__device__ __forceinline__ void ld_rb(double (&d)[RB], const double* __restrict__ p) {
    if constexpr (RB == 1) {
        d[0] = p[0];
    } else {
#pragma unroll
        for (int h = 0; h < RB / 2; ++h) {
            const double2 v = reinterpret_cast<const double2*>(p)[h];
            d[2 * h] = v.x;
            d[2 * h + 1] = v.y;
        }
    }
}
template <int RB>
__device__ __forceinline__ void st_rb(double* __restrict__ p, const double (&d)[RB]) {
    if constexpr (RB == 1) {
        p[0] = d[0];
    } else {
#pragma unroll
        for (int h = 0; h < RB / 2; ++h) reinterpret_cast<double2*>(p)[h] = make_double2(d[2 * h], d[2 * h + 1]);
    }
}

// Stages `m` sources (ids in sid[0..m)) of a chunk: coordinates and their R-wide density rows.
template <int R>
static __device__ __forceinline__ void mfma_stage_rows(int m, const int32_t* sid, const double* __restrict__ tau,
                                                double* st) {
    for (int e = threadIdx.x; e < m * R; e += blockDim.x) st[e] = tau[(size_t)sid[e / R] * R + e % R];
}

// P2M of one box per workgroup, wave w on check points 16w + (0..15) (P <= 64; a wave past P
// stages with the others and computes nothing).  The box's sources (about 26 on the 1M
// clustered tree) and their density rows are staged in LDS once for all check blocks: measured
// against one wave per check-block pair reading the rows from the L2 (0.65 against 1.52 ms at
// 64 right-hand sides on that tree, 9.1 against 10.0 ms on the 4M annulus with K0), the few
// steps per box leave that form's load latency exposed.
// U_b[r][k] = pref sum_s tau[s][r] G(source s, check point k)  (D = tau^T K on the MFMA).
template <int KIND, int RB>
static __global__ void __launch_bounds__(256) k_p2m_mfma(const int32_t* __restrict__ list, const int32_t* __restrict__ rid,
                                                  const int32_t* __restrict__ bop, const double* __restrict__ cx,
                                                  const double* __restrict__ cy, const int32_t* __restrict__ bot,
                                                  const int32_t* __restrict__ top, const double* __restrict__ xs,
                                                  const double* __restrict__ ys, const double* __restrict__ tau,
                                                  const double* __restrict__ surf, int p, int P, KerT<KIND> ker,
                                                  double* __restrict__ U) {
    constexpr int R = 16 * RB, CH = KF_MFMA_CHUNK;
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    __shared__ __attribute__((aligned(16))) double st[CH * R];
    __shared__ double sx[CH], sy[CH];
    __shared__ int32_t sid[CH];
    stage_logt<KIND>(logt);  // published by the first staging barrier
    pin_coef(ker);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
    const int32_t b = list[blockIdx.x];
    const int l = bop[b];
    const int32_t b0 = bot[b], b1 = top[b];
    const size_t row0 = (size_t)rid[b] * R;
    const bool wave_on = 16 * w < P;
    const int k = min(16 * w + i16, P - 1);
    const bool kon = 16 * w + i16 < p;
    const double tx = cx[b] + surf[(size_t)(2 * l) * P + k], ty = cy[b] + surf[(size_t)(2 * l + 1) * P + k];
    d4 acc[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) acc[rb] = d4{0, 0, 0, 0};
    for (int32_t c = b0; c < b1; c += CH) {
        const int m = min(CH, b1 - c);
        __syncthreads();
        if (threadIdx.x < m) {
            sid[threadIdx.x] = c + threadIdx.x;
            sx[threadIdx.x] = xs[c + threadIdx.x];
            sy[threadIdx.x] = ys[c + threadIdx.x];
        }
        __syncthreads();
        mfma_stage_rows<R>(m, sid, tau, st);
        __syncthreads();
        if (wave_on)
            for (int s0 = 0; s0 < m; s0 += 4) {
                const int s = s0 + k4, sc = min(s, m - 1);
                const bool on = s < m && kon;
                const double dx = tx - sx[sc], dy = ty - sy[sc];
                const double v = ker.raw(on ? fma(dx, dx, dy * dy) : 1.0, logt);
                const double kv = on ? v : 0.0;
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    const double a = st[sc * R + 16 * rb + i16];
                    acc[rb] = mfma_f64(s < m ? a : 0.0, kv, acc[rb]);
                }
            }
    }
    if (wave_on)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int v = 0; v < 4; ++v) U[(row0 + 16 * rb + 4 * v + k4) * P + k] = ker.pref() * acc[rb][v];
}

// L2P, target blocks q .. q + NT - 1 of a leaf (rows form: the totals stay in `near`):
// near[t][r] += pref sum_k G(target t, equivalent point k) Psi[r][k]  (D = K Psi^T: A[i][k] =
// G(target i, point k), B[k][j] = Psi of slot j at point k).
template <int KIND, int RB, int NT>
__device__ __forceinline__ void l2p_mw(const KerT<KIND>& ker, const double* logt, int q, int32_t tb, int nt,
                                       double bx, double by, const double* __restrict__ sfx,
                                       const double* __restrict__ sfy, int p, const double* __restrict__ xs,
                                       const double* __restrict__ ys, const double* __restrict__ psi, int P,
                                       double* __restrict__ near) {
    constexpr int R = 16 * RB;
    const int lane = threadIdx.x, i16 = lane & 15, k4 = lane >> 4;
    double rx[NT], ry[NT];
    bool ton[NT];
#pragma unroll
    for (int c = 0; c < NT; ++c) {
        const int t = 16 * (q + c) + i16, tc = min(t, nt - 1);
        ton[c] = t < nt;
        rx[c] = xs[tb + tc] - bx;
        ry[c] = ys[tb + tc] - by;
    }
    d4 acc[NT][RB];
#pragma unroll
    for (int c = 0; c < NT; ++c)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) acc[c][rb] = d4{0, 0, 0, 0};
    const double* prow = psi + (size_t)(RB * i16) * P;  // slot i16's first row
    int kn = min(k4, p - 1);
    double exn = sfx[kn], eyn = sfy[kn], pn[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) pn[rb] = prow[(size_t)rb * P + kn];
    for (int k0 = 0; k0 < p; k0 += 4) {
        const bool kon = k0 + k4 < p;
        const double ex = exn, ey = eyn;
        double pv[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) pv[rb] = pn[rb];
        if (k0 + 4 < p) {
            kn = min(k0 + 4 + k4, p - 1);
            exn = sfx[kn];
            eyn = sfy[kn];
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) pn[rb] = prow[(size_t)rb * P + kn];
        }
#pragma unroll
        for (int c = 0; c < NT; ++c) {
            const double dx = rx[c] - ex, dy = ry[c] - ey;
            const bool on = kon && ton[c];
            const double v = ker.raw(on ? fma(dx, dx, dy * dy) : 1.0, logt);
            const double a = on ? v : 0.0;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) acc[c][rb] = mfma_f64(a, pv[rb], acc[c][rb]);
        }
    }
    // D register v of block rb: target 16 (q + c) + 4v + k4, right-hand side RB i16 + rb
#pragma unroll
    for (int c = 0; c < NT; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int t = 16 * (q + c) + 4 * v + k4;
            if (t < nt) {
                double* o = near + (size_t)(tb + t) * R + RB * i16;
                double d[RB];
                ld_rb<RB>(d, o);
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) d[rb] = fma(ker.pref(), acc[c][rb][v], d[rb]);
                st_rb<RB>(o, d);
            }
        }
}

// L2P of the leaves at level >= 2, one wave per leaf (passes of two target blocks).
template <int KIND, int RB>
static __global__ void __launch_bounds__(64) k_l2p_mfma(const int32_t* __restrict__ leaves, int n_leaves,
                                                        const int32_t* __restrict__ rid, const int32_t* __restrict__ lvl,
                                                        const int32_t* __restrict__ bop, const double* __restrict__ cx,
                                                        const double* __restrict__ cy, const int32_t* __restrict__ bot,
                                                        const int32_t* __restrict__ top, const double* __restrict__ xs,
                                                        const double* __restrict__ ys, const double* __restrict__ surf,
                                                        const double* __restrict__ Psi, double* __restrict__ near, int p,
                                                        int P, KerT<KIND> ker) {
    constexpr int R = 16 * RB;
    const int32_t b = leaves[xcd_chunked(blockIdx.x, n_leaves, KF_MW_RUN)];
    if (lvl[b] < 2) return;  // no far field above level 2: the totals are the near field
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    stage_logt<KIND>(logt);
    pin_coef(ker);
    const int ob = bop[b];
    const double* psi = Psi + (size_t)rid[b] * R * P;
    const double* sfx = surf + (size_t)(2 * ob) * P;
    const double* sfy = surf + (size_t)(2 * ob + 1) * P;
    const double bx = cx[b], by = cy[b];
    const int32_t tb = bot[b], nt = top[b] - tb;
    const int nb = (nt + 15) >> 4;
    __syncthreads();  // the log table
    constexpr int NB = mw_blocks<KIND>();
    for (int q = 0; q < nb; q += NB) {
        if (NB == 2 && nb - q >= 2)
            l2p_mw<KIND, RB, NB>(ker, logt, q, tb, nt, bx, by, sfx, sfy, p, xs, ys, psi, P, near);
        else
            l2p_mw<KIND, RB, 1>(ker, logt, q, tb, nt, bx, by, sfx, sfy, p, xs, ys, psi, P, near);
    }
}

// Near field, target blocks q .. q + NT - 1 of a leaf against its n_src sources (the leaf's
// own points, then the other colleagues' in list order, as one virtual sequence over the
// ranges [cb, ce)), the target itself skipped: near[t][r] = pref sum_s G(t, s) tau[s][r]
// (D = K tau: A[i][k] = G(target i, source k), B[k][j] = tau of source k for slot j).  The
// sources are listed in LDS (ids and coordinates) KF_MW_CHUNK at a time.
template <int KIND, int RB, int NT>
__device__ __forceinline__ void p2p_mw(const KerT<KIND>& ker, const double* logt, int q, int32_t tb, int nt,
                                       int n_src, const int32_t* cb, const int32_t* ce,
                                       const double* __restrict__ xs, const double* __restrict__ ys,
                                       const double* __restrict__ tau, int32_t* sid, double2* sxy,
                                       double* __restrict__ near) {
    constexpr int R = 16 * RB;
    const int lane = threadIdx.x, i16 = lane & 15, k4 = lane >> 4;
    double tx[NT], ty[NT];
    int32_t tid[NT];
    bool ton[NT];
#pragma unroll
    for (int c = 0; c < NT; ++c) {
        const int t = 16 * (q + c) + i16, tc = min(t, nt - 1);
        ton[c] = t < nt;
        tid[c] = tb + t;
        tx[c] = xs[tb + tc];
        ty[c] = ys[tb + tc];
    }
    d4 acc[NT][RB];
#pragma unroll
    for (int c = 0; c < NT; ++c)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) acc[c][rb] = d4{0, 0, 0, 0};
    for (int c0 = 0; c0 < n_src; c0 += KF_MW_CHUNK) {
        const int m = min(KF_MW_CHUNK, n_src - c0);
        __syncthreads();  // the previous chunk's list has been read
        for (int e = lane; e < m; e += 64) {
            int v = c0 + e, j = 0;
            while (v >= ce[j] - cb[j]) {
                v -= ce[j] - cb[j];
                ++j;
            }
            const int32_t s = cb[j] + v;
            sid[e] = s;
            sxy[e] = make_double2(xs[s], ys[s]);
        }
        __syncthreads();
        int32_t idn = sid[min(k4, m - 1)];
        double tn[RB];
        ld_rb<RB>(tn, tau + (size_t)idn * R + RB * i16);
        for (int s0 = 0; s0 < m; s0 += 4) {
            const int k = s0 + k4;
            const bool kon = k < m;
            const int32_t id = idn;
            const double2 sp = sxy[min(k, m - 1)];
            double tv[RB];
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) tv[rb] = tn[rb];
            if (s0 + 4 < m) {
                idn = sid[min(k + 4, m - 1)];
                ld_rb<RB>(tn, tau + (size_t)idn * R + RB * i16);
            }
#pragma unroll
            for (int c = 0; c < NT; ++c) {
                const double dx = tx[c] - sp.x, dy = ty[c] - sp.y;
                const bool on = kon && ton[c] && id != tid[c];
                const double v = ker.raw(on ? fma(dx, dx, dy * dy) : 1.0, logt);
                const double a = on ? v : 0.0;
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) acc[c][rb] = mfma_f64(a, tv[rb], acc[c][rb]);
            }
        }
    }
    // D register v of block rb: target 16 (q + c) + 4v + k4, right-hand side RB i16 + rb
#pragma unroll
    for (int c = 0; c < NT; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int t = 16 * (q + c) + 4 * v + k4;
            if (t < nt) {
                double d[RB];
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) d[rb] = ker.pref() * acc[c][rb][v];
                st_rb<RB>(near + (size_t)(tb + t) * R + RB * i16, d);
            }
        }
}

// Near field of the leaves, one wave per leaf (passes of two target blocks); with a grid
// smaller than the leaf count each workgroup loops over the leaves.
template <int KIND, int RB>
static __global__ void __launch_bounds__(64) k_p2p_mfma(const int32_t* __restrict__ leaves, int n_leaves,
                                                        const int32_t* __restrict__ leaf_coll,
                                                        const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                                        const double* __restrict__ xs, const double* __restrict__ ys,
                                                        const double* __restrict__ tau, KerT<KIND> ker,
                                                        double* __restrict__ near) {
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    __shared__ double2 sxy[KF_MW_CHUNK];
    __shared__ int32_t sid[KF_MW_CHUNK], cb[10], ce[10];
    stage_logt<KIND>(logt);  // published by the barrier that opens each leaf
    pin_coef(ker);
    const bool mapped = (int)gridDim.x == n_leaves;
    for (int it = blockIdx.x; it < n_leaves; it += gridDim.x) {
        const int li = mapped ? xcd_chunked(it, n_leaves, KF_MW_RUN) : it;
        const int32_t b = leaves[li];
        const int32_t tb = bot[b], nt = top[b] - tb;
        __syncthreads();
        if (threadIdx.x < 10) {  // range 0: the leaf itself; 1..9: the other colleagues
            int32_t c0 = tb, c1 = tb + nt;
            if (threadIdx.x > 0) {
                const int32_t c = leaf_coll[(size_t)li * 9 + threadIdx.x - 1];
                const bool use = c >= 0 && c != b;
                c0 = use ? bot[c] : 0;
                c1 = use ? top[c] : 0;
            }
            cb[threadIdx.x] = c0;
            ce[threadIdx.x] = c1;
        }
        __syncthreads();
        int n_src = 0;
#pragma unroll
        for (int j = 0; j < 10; ++j) n_src += ce[j] - cb[j];
        const int nb = (nt + 15) >> 4;
        constexpr int NB = mw_blocks<KIND>();
        for (int q = 0; q < nb; q += NB) {
            if (NB == 2 && nb - q >= 2)
                p2p_mw<KIND, RB, NB>(ker, logt, q, tb, nt, n_src, cb, ce, xs, ys, tau, sid, sxy, near);
            else
                p2p_mw<KIND, RB, 1>(ker, logt, q, tb, nt, n_src, cb, ce, xs, ys, tau, sid, sxy, near);
        }
    }
}

template <int KIND, int RB>
static void launch_rb(const KfKernel& kd, const BatchMfmaArgs& a, int which, int lo, int hi, int near_grid,
                      hipStream_t s) {
    const KerT<KIND> ker = make_ker<KIND>(kd);
    if (which == 0 && a.n_p2m) {
        k_p2m_mfma<KIND, RB><<<a.n_p2m, 256, 0, s>>>(a.p2m, a.rid, a.bop, a.cx, a.cy, a.bot, a.top, a.xs, a.ys, a.tau,
                                                    a.surf, a.p, a.P, ker, a.U);
    }
    if (which == 1 && hi > lo) {
        const int n = hi - lo;
        const int grid = near_grid > 0 ? std::min(near_grid, n) : n;
        k_p2p_mfma<KIND, RB><<<grid, 64, 0, s>>>(a.leaves + lo, n, a.leaf_coll + (size_t)lo * 9, a.bot, a.top, a.xs,
                                                 a.ys, a.tau, ker, a.near);
    }
    if (which == 2 && a.n_leaves)
        k_l2p_mfma<KIND, RB><<<a.n_leaves, 64, 0, s>>>(a.leaves, a.n_leaves, a.rid, a.lvl, a.bop, a.cx, a.cy, a.bot,
                                                       a.top, a.xs, a.ys, a.surf, a.Psi, a.near, a.p, a.P, ker);
}

template <int KIND>
static void launch_kind(const KfKernel& kd, const BatchMfmaArgs& a, int which, int lo, int hi, int near_grid,
                        hipStream_t s) {
    if (a.R == 16) launch_rb<KIND, 1>(kd, a, which, lo, hi, near_grid, s);
    else if (a.R == 32) launch_rb<KIND, 2>(kd, a, which, lo, hi, near_grid, s);
    else launch_rb<KIND, 4>(kd, a, which, lo, hi, near_grid, s);
}

void launch_batch_mfma(const KfKernel& kd, const BatchMfmaArgs& a, int which, int lo, int hi, int near_grid,
                       hipStream_t s) {
    KF_REQUIRE(a.R == 16 || a.R == 32 || a.R == 64, KF_ERR_STATE, "batched MFMA forms take 16, 32 or 64 columns");
    KF_REQUIRE(a.P <= 64, KF_ERR_STATE, "batched MFMA forms take p <= 64");
    if (kd.kind == KF_LAPLACE) launch_kind<KF_LAPLACE>(kd, a, which, lo, hi, near_grid, s);
    else if (kd.kind == KF_TABLE) launch_kind<KF_TABLE>(kd, a, which, lo, hi, near_grid, s);
    else launch_kind<KF_MODHELM>(kd, a, which, lo, hi, near_grid, s);
}

}  // namespace kf
