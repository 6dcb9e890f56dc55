// batch_mfma.hip -- near field, P2M and L2P of batches of 16, 32 or 64 right-hand sides on the
// FP64 MFMA (the kernels behind launch_batch_mfma, batch_mfma.h).
#include <algorithm>

#include "batch_mfma.h"
#include "kf_internal.h"
#include "kf_kerlds.h"
#include "kf_mfma.h"

namespace kf {

// ---------------------------------------------------------- batched forms on the MFMA ----
// For batches of RB x 16 right-hand sides (16, 32 or 64) the near field, P2M and L2P form each
// kernel value once on the vector ALU and contract it with the right-hand sides on FP64 MFMA:
// v_mfma_f64_16x16x4 computes D[i][j] += sum_k A[i][k] B[k][j] with lane l supplying A[l%16][l/16]
// and B[l/16][l%16] and holding D[4v + l/16][l%16] in register v (kf_mfma.h).  Lane l of a step
// forms the kernel value of its own (l%16, l/16) pair, so no value is formed twice.  The kernel
// values are bitwise those of the vector forms; the sums run in MFMA order, so a right-hand side
// gets the same potential at every width >= 16 and, to rounding, the vector forms' below.
// Workgroups of four waves: one box (or leaf) per workgroup, the sources (or the local
// densities) staged through LDS once and shared by the waves, each wave on its own 16 check
// points (or targets).  Every load inside the step loops reads a clamped, in-range LDS address
// and the padding is zeroed by selects, so the loops have no branches around loads.
constexpr int KF_MFMA_CHUNK = 64;  // sources staged per pass

// Stages `m` sources (ids in sid[0..m)) of a chunk: coordinates and their R-wide density rows.
template <int R>
static __device__ __forceinline__ void mfma_stage_rows(int m, const int32_t* sid, const double* __restrict__ tau,
                                                double* st) {
    for (int e = threadIdx.x; e < m * R; e += blockDim.x) st[e] = tau[(size_t)sid[e / R] * R + e % R];
}

// P2M of one box per workgroup, wave w on check points 16w + (0..15) (P <= 64; a wave past P
// stages with the others and computes nothing):
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

// Waves per target block for a leaf of nb blocks of 16 targets: all four waves work on a leaf
// of one or two blocks by splitting the sum (over equivalent points or source steps) among
// them; the partial sums meet in LDS in a fixed order.
__device__ __forceinline__ int mfma_split(int nb) { return nb == 1 ? 4 : (nb == 2 ? 2 : 1); }

// Sums the ks partial 16 x R tiles of each of the nbp target blocks of a pass (red: wave w's
// tile at w * 16 R, row = target within the block, R doubles per row) and stores them
// (accumulate: adds to `near`) for the targets below nt; every thread takes part.
template <int R, bool ACC>
__device__ __forceinline__ void mfma_reduce_store(const double* red, int ks, int nbp, int B0, int nt, int32_t tb,
                                                  double pref, double* __restrict__ near) {
    for (int e = threadIdx.x; e < nbp * 16 * R; e += blockDim.x) {
        const int bi = e / (16 * R), rest = e - bi * 16 * R;
        double sum = 0.0;
        for (int kp = 0; kp < ks; ++kp) sum += red[(bi * ks + kp) * 16 * R + rest];
        const int tt = 16 * (B0 + bi) + rest / R;
        if (tt < nt) {
            const size_t nr = (size_t)(tb + tt) * R + rest % R;
            near[nr] = ACC ? pref * sum + near[nr] : pref * sum;
        }
    }
}

// Wave w's accumulators (D[4v + k4][i16] of block rb) into its tile of red.
template <int RB>
__device__ __forceinline__ void mfma_put_tile(double* red, const d4* acc) {
    constexpr int R = 16 * RB;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[w * 16 * R + (4 * v + k4) * R + 16 * rb + i16] = acc[rb][v];
}

// L2P of one leaf per workgroup (rows form: the totals stay in `near`, tree order):
// near[t][r] += pref sum_k G(target t, equivalent point k) Psi[r][k]  (D = K Psi^T).  Psi is
// staged in LDS transposed (k-major, rows of R + 1 so that the transposing writes spread over
// the banks).  Wave w takes target block B0 + w / ks and every ks-th step of 4 equivalent points.
template <int KIND, int RB>
static __global__ void __launch_bounds__(256) k_l2p_mfma(const int32_t* __restrict__ leaves, const int32_t* __restrict__ rid,
                                                         const int32_t* __restrict__ lvl, const int32_t* __restrict__ bop,
                                                         const double* __restrict__ cx, const double* __restrict__ cy,
                                                         const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                                         const double* __restrict__ xs, const double* __restrict__ ys,
                                                         const double* __restrict__ surf, const double* __restrict__ Psi,
                                                         double* __restrict__ near, int p, int P, KerT<KIND> ker) {
    constexpr int R = 16 * RB, RS = R + 1, PM = 64;  // P <= 64
    const int32_t b = leaves[blockIdx.x];
    if (lvl[b] < 2) return;  // no far field above level 2: the totals are the near field
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    __shared__ __attribute__((aligned(16))) double ps[PM * RS];  // later the partial tiles (64 R)
    __shared__ double ex[PM], ey[PM];
    stage_logt<KIND>(logt);
    pin_coef(ker);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
    const int ob = bop[b];
    const double* psi = Psi + (size_t)rid[b] * R * P;
    const double bx = cx[b], by = cy[b];
    const int32_t tb = bot[b], nt = top[b] - tb;
    const int nb = (nt + 15) >> 4, ks = mfma_split(nb), bpp = 4 / ks;
    for (int B0 = 0; B0 < nb; B0 += bpp) {
        __syncthreads();  // the previous pass has read its partial tiles
        for (int e = threadIdx.x; e < R * p; e += blockDim.x) {
            const int r = e / p, k = e - r * p;
            ps[k * RS + r] = psi[(size_t)r * P + k];
        }
        if (threadIdx.x < p) {
            ex[threadIdx.x] = surf[(size_t)(2 * ob) * P + threadIdx.x];
            ey[threadIdx.x] = surf[(size_t)(2 * ob + 1) * P + threadIdx.x];
        }
        __syncthreads();
        const int blk = B0 + w / ks, kp = w % ks;
        const int t = 16 * blk + i16, tc = min(t, nt - 1);
        const bool ton = t < nt;
        const double rx = xs[tb + tc] - bx, ry = ys[tb + tc] - by;
        d4 acc[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) acc[rb] = d4{0, 0, 0, 0};
        if (blk < nb)
            for (int k0 = 4 * kp; k0 < p; k0 += 4 * ks) {
                const int k = k0 + k4, kc = min(k, p - 1);
                const bool on = ton && k < p;
                const double dx = rx - ex[kc], dy = ry - ey[kc];
                const double v = ker.raw(on ? fma(dx, dx, dy * dy) : 1.0, logt);
                const double a = on ? v : 0.0;
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    const double x = ps[kc * RS + 16 * rb + i16];
                    acc[rb] = mfma_f64(a, k < p ? x : 0.0, acc[rb]);
                }
            }
        __syncthreads();  // every wave is done with the staged Psi
        mfma_put_tile<RB>(ps, acc);
        __syncthreads();
        mfma_reduce_store<R, true>(ps, ks, min(bpp, nb - B0), B0, nt, tb, ker.pref(), near);
    }
}

// Near field of one leaf per workgroup (grid-stride over the leaves): near[t][r] = pref sum
// over the colleagues' sources s (the leaf's own first, then the others in list order, the
// target itself skipped) of G(t, s) tau[s][r]  (D = K tau).  The sources are staged in chunks
// of 64 shared by the waves; wave w takes target block B0 + w / ks and every ks-th step of 4
// sources of each chunk.
template <int KIND, int RB>
static __global__ void __launch_bounds__(256) k_p2p_mfma(const int32_t* __restrict__ leaves, int n_leaves,
                                                         const int32_t* __restrict__ leaf_coll,
                                                         const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                                         const double* __restrict__ xs, const double* __restrict__ ys,
                                                         const double* __restrict__ tau, KerT<KIND> ker,
                                                         double* __restrict__ near) {
    constexpr int R = 16 * RB, CH = KF_MFMA_CHUNK;
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    __shared__ __attribute__((aligned(16))) double st[CH * R];  // later the partial tiles (64 R)
    __shared__ double sx[CH], sy[CH];
    __shared__ int32_t sid[CH], cbeg[10], cend[10];
    stage_logt<KIND>(logt);  // published by the barrier that opens each leaf
    pin_coef(ker);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, i16 = lane & 15, k4 = lane >> 4;
    for (int li = blockIdx.x; li < n_leaves; li += gridDim.x) {
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
            cbeg[threadIdx.x] = c0;
            cend[threadIdx.x] = c1;
        }
        __syncthreads();
        int n_src = 0;
#pragma unroll
        for (int j = 0; j < 10; ++j) n_src += cend[j] - cbeg[j];
        const int nb = (nt + 15) >> 4, ks = mfma_split(nb), bpp = 4 / ks;
        for (int B0 = 0; B0 < nb; B0 += bpp) {
            const int blk = B0 + w / ks, kp = w % ks;
            const int t = 16 * blk + i16, tc = min(t, nt - 1);
            const bool ton = t < nt, wave_on = blk < nb;
            const double tx = xs[tb + tc], ty = ys[tb + tc];
            d4 acc[RB];
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) acc[rb] = d4{0, 0, 0, 0};
            for (int v0 = 0; v0 < n_src; v0 += CH) {
                const int m = min(CH, n_src - v0);
                __syncthreads();  // the previous chunk (or pass) is done with the staging area
                if (threadIdx.x < m) {
                    int v = v0 + threadIdx.x, j = 0;
                    while (v >= cend[j] - cbeg[j]) {
                        v -= cend[j] - cbeg[j];
                        ++j;
                    }
                    const int32_t s = cbeg[j] + v;
                    sid[threadIdx.x] = s;
                    sx[threadIdx.x] = xs[s];
                    sy[threadIdx.x] = ys[s];
                }
                __syncthreads();
                mfma_stage_rows<R>(m, sid, tau, st);
                __syncthreads();
                if (wave_on)
                    for (int s0 = 4 * kp; s0 < m; s0 += 4 * ks) {
                        const int k = s0 + k4, kc = min(k, m - 1);
                        const bool on = k < m && ton && sid[kc] != tb + t;
                        const double dx = tx - sx[kc], dy = ty - sy[kc];
                        const double v = ker.raw(on ? fma(dx, dx, dy * dy) : 1.0, logt);
                        const double a = on ? v : 0.0;
#pragma unroll
                        for (int rb = 0; rb < RB; ++rb) {
                            const double x = st[kc * R + 16 * rb + i16];
                            acc[rb] = mfma_f64(a, k < m ? x : 0.0, acc[rb]);
                        }
                    }
            }
            __syncthreads();  // every wave is done with the last chunk
            mfma_put_tile<RB>(st, acc);
            __syncthreads();
            mfma_reduce_store<R, false>(st, ks, min(bpp, nb - B0), B0, nt, tb, ker.pref(), near);
        }
    }
}

template <int KIND, int RB>
static void launch_rb(const KfKernel& kd, const BatchMfmaArgs& a, int which, int lo, int hi, int near_grid,
                      hipStream_t s) {
    const KerT<KIND> ker = make_ker<KIND>(kd);
    if (which == 0 && a.n_p2m)
        k_p2m_mfma<KIND, RB><<<a.n_p2m, 256, 0, s>>>(a.p2m, a.rid, a.bop, a.cx, a.cy, a.bot, a.top, a.xs, a.ys, a.tau,
                                                    a.surf, a.p, a.P, ker, a.U);
    if (which == 1 && hi > lo) {
        const int n = hi - lo;
        const int grid = near_grid > 0 ? std::min(near_grid, n) : n;
        k_p2p_mfma<KIND, RB><<<grid, 256, 0, s>>>(a.leaves + lo, n, a.leaf_coll + (size_t)lo * 9, a.bot, a.top, a.xs,
                                                  a.ys, a.tau, ker, a.near);
    }
    if (which == 2 && a.n_leaves)
        k_l2p_mfma<KIND, RB><<<a.n_leaves, 256, 0, s>>>(a.leaves, a.rid, a.lvl, a.bop, a.cx, a.cy, a.bot, a.top, a.xs,
                                                        a.ys, a.surf, a.Psi, a.near, a.p, a.P, ker);
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
