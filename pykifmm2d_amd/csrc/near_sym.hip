// near_sym.hip -- near field of one right-hand side with each pair of mutual colleague leaves
// evaluated once (reference fmm.py:30-64 evaluates every directed pair).
//
// G is symmetric in its two points, so for two leaves A and B that are each other's colleagues
// one kernel value G(a, b) serves both sums: the row sum of target a over the sources of B and
// the column sum of target b over the sources of A.  84% of the near-field pairs of the 1M-point
// clustered tree lie in such blocks (a leaf's own block included).  The pair is owned by the
// leaf that comes first in the leaf list; its wave forms the block once and produces both sums.
//
// One wave per leaf A (its targets staged in LDS): its own block (each unordered pair once), the
// blocks of the mutual colleagues it owns, and its other colleagues (a colleague that is not a
// leaf with points: row sums only; the reverse direction belongs to the near field of that
// colleague's descendants).  A step forms G for 4 targets x 16 sources, lane l taking target
// 4 ia + l/16 and source 16 jb + l%16:
//   * row sums of A's targets: each lane accumulates its products in FP64, one accumulator per
//     target quartet ia (a partial over the sources with its l%16); a butterfly over the 16
//     lanes of each row group totals them once per leaf;
//   * column sums of B's sources: v_mfma_f64_16x16x4 with the step's kernel values as the A
//     operand (A[i][k] = G(target 4 ia + k, source 16 jb + i)) and the targets' densities as the
//     B operand contracts the 4 targets, accumulating over the target quartets of the block.
// Each wave writes its leaf's row sums to slot 0, its own block's column sums to slot 1 and the
// column sums of an owned colleague B to B's slot 2 + j (j: A's place in B's colleague list);
// k_near_sum adds a point's slots in a fixed order.  No atomics: the result does not depend on
// scheduling.
#include <algorithm>

#include "kf_internal.h"
#include "kf_kerlds.h"
#include "kf_mfma.h"
#include "near_sym.h"

namespace kf {

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_leaf_index(const int32_t* __restrict__ leaves, int n, int32_t* __restrict__ lidx) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) lidx[leaves[k]] = k;
}

__global__ void k_leaf_pairs(const int32_t* __restrict__ leaves, int n, const int32_t* __restrict__ lcoll,
                             const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                             const int32_t* __restrict__ lidx, int32_t* __restrict__ pinfo, int* __restrict__ max_leaf) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= 9 * (int64_t)n) return;
    const int li = (int)(e / 9);
    const int32_t b = leaves[li], c = lcoll[e];
    if (e % 9 == 0) atomicMax(max_leaf, top[b] - bot[b]);
    int32_t v = -1;
    if (c >= 0 && c != b) {
        const int32_t lc = lidx[c];
        if (lc >= 0)
            for (int jj = 0; jj < 9; ++jj)
                if (lcoll[9 * (int64_t)lc + jj] == b) v = lc * 16 + jj;
    }
    pinfo[e] = v;
}

void build_leaf_pairs(const int32_t* leaves, int n_leaves, const int32_t* lcoll, const int32_t* bot,
                      const int32_t* top, int64_t n_boxes, int32_t* pinfo, int* max_leaf, hipStream_t s) {
    DevBuf<int32_t> lidx;
    DevBuf<int> mx;
    lidx.alloc((size_t)std::max<int64_t>(1, n_boxes));
    mx.alloc(1);
    mx.zero(s);
    k_fill_i32<<<1024, 256, 0, s>>>(lidx.p, n_boxes, -1);
    if (n_leaves) {
        k_leaf_index<<<(n_leaves + 255) / 256, 256, 0, s>>>(leaves, n_leaves, lidx.p);
        const int64_t ne = 9 * (int64_t)n_leaves;
        k_leaf_pairs<<<(unsigned)((ne + 255) / 256), 256, 0, s>>>(leaves, n_leaves, lcoll, bot, top, lidx.p, pinfo,
                                                                  mx.p);
    }
    KF_HIP_CHECK(hipGetLastError());
    KF_HIP_CHECK(hipMemcpyAsync(max_leaf, mx.p, sizeof(int), hipMemcpyDeviceToHost, s));
    KF_HIP_CHECK(hipStreamSynchronize(s));  // lidx and mx are freed on return
}

enum { SYM_SELF = 0, SYM_PAIR = 1, SYM_ROWS = 2 };

// One step of leaf A's near field: A's na targets (LDS: axy coordinates, at densities, both
// padded to 64 entries with copies of the last target carrying density 0) against 16 sources,
// the lane's source at (bx, by) with density bt (0 past the block's end, where the coordinates
// copy the block's last source), index bs within a block of nb.  SYM_SELF: A's own block, the
// pairs a < b; SYM_PAIR: an owned mutual colleague, both sums; SYM_ROWS: row sums only.  Row
// shares go to racc (accumulator ia: target quartet ia); the column sums of the 16 sources are
// returned in the MFMA layout.  Padding needs no mask: padded targets have density 0 (no
// column share) and their row sums are never stored; padded sources have density 0 (no row
// share) and their column sums are never stored.  In A's own block the pairs a >= b (the
// diagonal among them) and the padded sources (there copies of A's last target, at r^2 = 0 from
// it) are masked, with r^2 = 1 in their place.
#ifndef KF_SYM_QG
#define KF_SYM_QG 2
#endif
template <int KIND, int MODE>
__device__ __forceinline__ d4 sym_step(const KerT<KIND>& ker, const double* logt, const double2* axy,
                                       const double* at, int na, double bx, double by, double bt, int jb16, int nb,
                                       double (&racc)[16]) {
    constexpr int QG = KF_SYM_QG;  // target quartets per group: their evaluations interleave
    const int g4 = (threadIdx.x & 63) >> 4, bs = jb16 + (threadIdx.x & 15);
    d4 D = d4{0, 0, 0, 0};
    const int lim = MODE == SYM_SELF ? min(na, jb16 + 16) : na;  // own block: targets below its last b
#pragma unroll
    for (int i0 = 0; i0 < KF_SYM_MAX_LEAF / 4; i0 += QG) {
        if (4 * i0 >= lim) break;  // a group past na only adds exact zeros (density 0, rows unread)
        double r2[QG], g[QG];
        bool on[QG], ok = true;
#pragma unroll
        for (int q = 0; q < QG; ++q) {
            const int a = 4 * (i0 + q) + g4;
            const double2 ap = axy[a];
            const double dx = ap.x - bx, dy = ap.y - by;
            r2[q] = fma(dx, dx, dy * dy);
            on[q] = MODE != SYM_SELF || (a < bs && bs < nb);
            if (MODE == SYM_SELF) r2[q] = on[q] ? r2[q] : 1.0;
            g[q] = ker.raw_fast(r2[q], logt);
            ok = ok && ker.fast_ok(r2[q]);
        }
        if (__builtin_expect(!ok, 0))  // a zero, subnormal, inf or nan distance in the group
#pragma unroll
            for (int q = 0; q < QG; ++q) g[q] = ker.raw(r2[q], logt);
#pragma unroll
        for (int q = 0; q < QG; ++q) {
            const int a = 4 * (i0 + q) + g4;
            const double gq = MODE == SYM_SELF ? (on[q] ? g[q] : 0.0) : g[q];
            racc[i0 + q] = fma(gq, bt, racc[i0 + q]);
            if (MODE != SYM_ROWS) D = mfma_f64(gq, at[a], D);
        }
    }
    return D;
}

// x from the lane whose index within its 16-lane row differs from this lane's in bit h (a
// swizzle: no address registers).
template <int H>
__device__ __forceinline__ double xor_lane(double x) {
    const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(x), 0x1F | (H << 10));
    const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(x), 0x1F | (H << 10));
    return __hiloint2double(hi, lo);
}

template <int H>
__device__ __forceinline__ void row_stage(double (&r)[16], int i16) {
    const bool up = (i16 & H) != 0;
#pragma unroll
    for (int k = 0; k < H; ++k) {
        const double send = up ? r[k] : r[k + H];
        const double keep = up ? r[k + H] : r[k];
        r[k] = keep + xor_lane<H>(send);
    }
}

// Totals of the 16 accumulators over the 16 lanes of each row group (lanes with equal l/16):
// four exchange stages, each halving the accumulators a lane keeps; the lane with l%16 = i
// ends with the total of accumulator i.  A fixed order of additions.
__device__ __forceinline__ double row_totals(double (&r)[16]) {
    const int i16 = threadIdx.x & 15;
    row_stage<8>(r, i16);
    row_stage<4>(r, i16);
    row_stage<2>(r, i16);
    row_stage<1>(r, i16);
    return r[0];
}

// A leaf's own block runs in steps of 16 of its own points (read from the staged targets);
// the sources of the colleagues it evaluates (owned pairs and rows-only colleagues, in slot
// order) form one list, staged in LDS KF_SYM_CHUNK at a time with each source's column slot
// (-1: rows only), so that a step of 16 sources is full whatever the colleagues' sizes.
constexpr int KF_SYM_CHUNK = 128;

template <int KIND>
__global__ void __launch_bounds__(64, 4) k_p2p_sym(const int32_t* __restrict__ leaves, int lo, int hi,
                                                   const int32_t* __restrict__ lcoll, const int32_t* __restrict__ pinfo,
                                                   const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                                   const double* __restrict__ xs, const double* __restrict__ ys,
                                                   const double* __restrict__ tau, KerT<KIND> ker,
                                                   double* __restrict__ near, int64_t N) {
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    __shared__ double2 axy[KF_SYM_MAX_LEAF], sxy[KF_SYM_CHUNK];
    __shared__ double at[KF_SYM_MAX_LEAF], st[KF_SYM_CHUNK];
    __shared__ int32_t spt[KF_SYM_CHUNK], sco[KF_SYM_CHUNK];  // point, column slot (-1: none)
    __shared__ int32_t cb0[9], cnb[9], cco[9];                // the colleague blocks in list order
    stage_logt<KIND>(logt);  // published by the barrier that opens each leaf
    pin_coef(ker);
    const int lane = threadIdx.x, i16 = lane & 15, g4 = lane >> 4;
    double* slot = near + N;  // slot s of point t at slot[s N + t]
    for (int li = lo + blockIdx.x; li < hi; li += gridDim.x) {
        const int32_t b = leaves[li];
        const int32_t ab = bot[b];
        const int na = top[b] - ab;
        // lane j < 9: colleague slot j, evaluated here unless a partner of lower index owns it
        bool use = false;
        int32_t c0 = 0, nb = 0, co = -1;
        if (lane < 9) {
            const int32_t c = lcoll[9 * (int64_t)li + lane];
            if (c >= 0 && c != b) {
                c0 = bot[c];
                nb = top[c] - c0;
                const int32_t pi = pinfo[9 * (int64_t)li + lane];
                use = nb > 0 && (pi < 0 || (pi >> 4) > li);
                co = pi < 0 ? -1 : 2 + (pi & 15);
            }
        }
        const uint64_t used = __ballot(use);
        const int ncb = __popcll(used);
        __syncthreads();  // the previous leaf's lists and targets have been read
        if (use) {
            const int k = __popcll(used & ((1ull << lane) - 1));
            cb0[k] = c0;
            cnb[k] = nb;
            cco[k] = co;
        }
        {  // every entry written: past na, copies of the last target with density 0
            const int32_t t = ab + min(lane, na - 1);
            axy[lane] = make_double2(xs[t], ys[t]);
            at[lane] = lane < na ? tau[t] : 0.0;
        }
        __syncthreads();
        double racc[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) racc[k] = 0.0;
        // the own block: the pairs a < b, column sums to slot 1
        for (int jb = 0; 16 * jb < na; ++jb) {
            const int bs = 16 * jb + i16;
            const double2 sp = axy[bs];
            const d4 D = sym_step<KIND, SYM_SELF>(ker, logt, axy, at, na, sp.x, sp.y, at[bs], 16 * jb, na, racc);
            if (i16 == 0)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int s = 16 * jb + 4 * v + g4;
                    if (s < na) slot[N + ab + s] = ker.pref() * D[v];
                }
        }
        // the colleagues' sources as one list
        int n_src = 0;
        for (int k = 0; k < ncb; ++k) n_src += cnb[k];
        for (int v0 = 0; v0 < n_src; v0 += KF_SYM_CHUNK) {
            const int m = min(KF_SYM_CHUNK, n_src - v0);
            __syncthreads();  // the previous chunk has been read
            for (int e = lane; e < m; e += 64) {
                int v = v0 + e, k = 0;
                while (v >= cnb[k]) {
                    v -= cnb[k];
                    ++k;
                }
                const int32_t pt = cb0[k] + v;
                sxy[e] = make_double2(xs[pt], ys[pt]);
                st[e] = tau[pt];
                spt[e] = pt;
                sco[e] = cco[k];
            }
            __syncthreads();
            for (int s16 = 0; s16 < m; s16 += 16) {
                const int e = s16 + i16, ec = min(e, m - 1);  // past m: a copy of the last source
                const double2 sp = sxy[ec];
                const double bt = e < m ? st[ec] : 0.0;
                const d4 D = sym_step<KIND, SYM_PAIR>(ker, logt, axy, at, na, sp.x, sp.y, bt, s16, m, racc);
                if (i16 == 0)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int es = s16 + 4 * v + g4;
                        if (es < m && sco[es] >= 0) slot[(int64_t)sco[es] * N + spt[es]] = ker.pref() * D[v];
                    }
            }
        }
        const double tot = row_totals(racc);
        const int a = 4 * i16 + g4;
        if (a < na) slot[ab + a] = ker.pref() * tot;
    }
}

// near[t] = slot 0 + slot 1 + the slots 2 + j whose colleague owns the pair, in increasing j.
__global__ void __launch_bounds__(64) k_near_sum(const int32_t* __restrict__ leaves, int n_leaves,
                                                 const int32_t* __restrict__ pinfo, const int32_t* __restrict__ bot,
                                                 const int32_t* __restrict__ top, double* __restrict__ near, int64_t N) {
    const double* slot = near + N;
    for (int li = blockIdx.x; li < n_leaves; li += gridDim.x) {
        const int32_t b = leaves[li];
        const int32_t tb = bot[b];
        const int nt = top[b] - tb;
        uint32_t in = 0;
        for (int j = 0; j < 9; ++j) {
            const int32_t pi = pinfo[9 * (int64_t)li + j];
            if (pi >= 0 && (pi >> 4) < li) in |= 1u << j;
        }
        for (int t = threadIdx.x; t < nt; t += 64) {
            double v = slot[tb + t] + slot[N + tb + t];
            for (int j = 0; j < 9; ++j)
                if (in >> j & 1u) v += slot[(int64_t)(2 + j) * N + tb + t];
            near[tb + t] = v;
        }
    }
}

template <int KIND>
static void launch_kind(const KfKernel& kd, const NearSymArgs& a, int lo, int hi, bool sum, int grid, hipStream_t s) {
    const KerT<KIND> ker = make_ker<KIND>(kd);
    if (hi > lo) {
        const int g = grid > 0 ? std::min(grid, hi - lo) : hi - lo;
        k_p2p_sym<KIND><<<g, 64, 0, s>>>(a.leaves, lo, hi, a.lcoll, a.pinfo, a.bot, a.top, a.xs, a.ys, a.tau, ker,
                                         a.near, a.N);
    }
    if (sum) {
        const int g = grid > 0 ? std::min(grid, a.n_leaves) : a.n_leaves;
        k_near_sum<<<g, 64, 0, s>>>(a.leaves, a.n_leaves, a.pinfo, a.bot, a.top, a.near, a.N);
    }
}

void launch_near_sym(const KfKernel& kd, const NearSymArgs& a, int lo, int hi, bool sum, int grid, hipStream_t s) {
    if (!a.n_leaves) return;
    if (kd.kind == KF_LAPLACE) launch_kind<KF_LAPLACE>(kd, a, lo, hi, sum, grid, s);
    else if (kd.kind == KF_TABLE) launch_kind<KF_TABLE>(kd, a, lo, hi, sum, grid, s);
    else launch_kind<KF_MODHELM>(kd, a, lo, hi, sum, grid, s);
}

}  // namespace kf
