// plan_kernels.h -- device kernels of the planned KIFMM evaluation (included by plan.hip).
//
// Every pass works on a batch of R right-hand sides (R a power of two <= 64).  Slabs are
// [rows][P] with row = slab_row(box) * R + r, so the R rows of a box are contiguous; a 64-row
// MFMA tile holds 64/R boxes x R right-hand sides and all of them share the tile's operators
// and source boxes.  Work lists of the MFMA passes carry slab rows; the VALU passes map a box
// to its row through rid[].  The VALU kernels (P2M, P2P, L2P) evaluate each kernel value once and apply it to a
// chunk of RC right-hand sides held in registers (grid.y runs over the R/RC chunks).
#pragma once
#include "kf_common.h"
#include "kf_kerlds.h"
#include "kf_mfma.h"
#include "kf_xcd.h"
#include "ops_host.h"

// Neighbouring leaves (and P2M boxes) per XCD run in the vector-ALU point passes (xcd_chunked);
// 0 keeps the dispatcher's round-robin order.
#ifndef KF_VALU_RUN
#define KF_VALU_RUN 0
#endif

namespace kf {

// Device addresses of the right-hand sides of one batch, passed by value in the kernel
// arguments (so a batch needs no host-to-device copy): the density vector r of the batch is
// in[r][0..n) in input order, its potential goes to out[r][0..n).
constexpr int KF_MAX_BATCH = 64;
struct RhsIn {
    const double* p[KF_MAX_BATCH];
};
struct RhsOut {
    double* p[KF_MAX_BATCH];
};


// tau_s[i*R + r] = in[r][perm[i]] for r < nvalid, 0 for the padding columns, a thread per
// (tree index i, r): the writes are coalesced, each read is an 8-byte load from a random place
// of in[r].  Used below KF_ROWS_MIN right-hand sides.
__global__ void k_gather_tau(RhsIn in, int nvalid, const int32_t* __restrict__ perm, int64_t n, int lgR,
                             double* __restrict__ out) {
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= (n << lgR)) return;
    int64_t i = e >> lgR;
    int r = (int)(e & ((1 << lgR) - 1));
    out[e] = r < nvalid ? in.p[r][perm[i]] : 0.0;
}

// The same for wide batches, by rows: workgroup k takes the input indices [JB k, JB k + JB)
// (JB = max(16, 256 / R)) for all R right-hand sides, so that each cache line of in[r] (16
// consecutive input indices) is read within one workgroup, and the contiguous tree-order rows
// tau_s[iperm[j]*R .. +R) it writes are completed there too (random input order, 64
// right-hand sides: the element-wise form fetches each input line once per element).
__global__ void __launch_bounds__(256) k_gather_rows(RhsIn in, int nvalid, const int32_t* __restrict__ iperm,
                                                     int64_t n, int lgR, double* __restrict__ out) {
    const int R = 1 << lgR, JB = R >= 16 ? 16 : 256 >> lgR;
    const int64_t j0 = (int64_t)blockIdx.x * JB;
    for (int e = threadIdx.x; e < JB * R; e += 256) {
        const int r = e / JB;
        const int64_t j = j0 + (e - r * JB);
        if (j < n) out[((int64_t)iperm[j] << lgR) + r] = r < nvalid ? in.p[r][j] : 0.0;
    }
}

// Potentials from tree-order rows back to input order, out[r][j] = rows[iperm[j]*R + r] for
// r < nvalid: the reverse of k_gather_rows (the wide-batch L2P leaves its totals in rows).
__global__ void __launch_bounds__(256) k_scatter_pot(RhsOut out, int nvalid, const int32_t* __restrict__ iperm,
                                                     int64_t n, int lgR, const double* __restrict__ rows) {
    const int R = 1 << lgR, JB = R >= 16 ? 16 : 256 >> lgR;
    const int64_t j0 = (int64_t)blockIdx.x * JB;
    for (int e = threadIdx.x; e < JB * R; e += 256) {
        const int r = e / JB;
        const int64_t j = j0 + (e - r * JB);
        if (j < n && r < nvalid) out.p[r][j] = rows[((int64_t)iperm[j] << lgR) + r];
    }
}

// iperm[perm[i]] = i.
__global__ void k_invert_perm(const int32_t* __restrict__ perm, int64_t n, int32_t* __restrict__ iperm) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) iperm[perm[i]] = (int32_t)i;
}

// P2M (reference class_fmm.py:57-71): one wave per (box, rhs chunk); lane j of group g takes
// check point j on the box's large circle and every G-th source (G = 64 / p lane groups, so
// small p keeps all lanes busy); sources staged through LDS 64 at a time.
// ST: the kernel values are read from the stored matrix (k_p2m_store: box i of the list at
// moff[i], source k's p values at k * p), bitwise the values computed here otherwise.
template <int KIND, int RC, bool ST>
__global__ void __launch_bounds__(64) k_p2m(const int32_t* __restrict__ list, const int32_t* __restrict__ rid,
                                            const int32_t* __restrict__ bop,
                                            const double* __restrict__ cx, const double* __restrict__ cy,
                                            const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                            const double* __restrict__ xs, const double* __restrict__ ys,
                                            const double* __restrict__ tau, const double* __restrict__ surf,
                                            int p, int P, int lgR, KerT<KIND> ker, const int64_t* __restrict__ moff,
                                            const double* __restrict__ M, double* __restrict__ U) {
    __shared__ double sx[64], sy[64], st[64 * RC], red[64 * RC];
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    stage_logt<KIND>(logt);  // published by the first staging barrier below
    pin_coef(ker);
    const int lane = threadIdx.x, R = 1 << lgR, r0 = blockIdx.y * RC;
    const int bi = KF_VALU_RUN > 0 ? xcd_chunked(blockIdx.x, gridDim.x, KF_VALU_RUN) : blockIdx.x;
    const int32_t b = list[bi];
    const int l = bop[b];
    const int G = 64 / p, g = lane / p, j = lane - g * p;
    const bool on = g < G;
    const double tx = on ? cx[b] + surf[(size_t)(2 * l) * P + j] : 0.0;
    const double ty = on ? cy[b] + surf[(size_t)(2 * l + 1) * P + j] : 0.0;
    double acc[RC];
#pragma unroll
    for (int q = 0; q < RC; ++q) acc[q] = 0.0;
    const int32_t b0 = bot[b], b1 = top[b];
    for (int32_t c = b0; c < b1; c += 64) {
        const int m = min(64, b1 - c);
        __syncthreads();
        if (!ST && lane < m) {
            sx[lane] = xs[c + lane];
            sy[lane] = ys[c + lane];
        }
        for (int e = lane; e < m * RC; e += 64) st[e] = tau[(size_t)(c + e / RC) * R + r0 + e % RC];
        __syncthreads();
        if (on) {
            if constexpr (ST) {
                const double* Mb = M + moff[bi] + (size_t)(c - b0) * p + j;
#pragma unroll 4
                for (int k = g; k < m; k += G) {
                    const double v = Mb[(size_t)k * p];
#pragma unroll
                    for (int q = 0; q < RC; ++q) acc[q] = fma(v, st[k * RC + q], acc[q]);
                }
            } else {
                constexpr int S = 4;  // sources per step (their kernel values side by side), taken in order
                int k = g;
                for (; k + (S - 1) * G < m; k += S * G) {
                    double r2[S], v[S];
#pragma unroll
                    for (int i = 0; i < S; ++i) {
                        const double dx = tx - sx[k + i * G], dy = ty - sy[k + i * G];
                        r2[i] = fma(dx, dx, dy * dy);
                    }
                    raw_group(ker, r2, logt, v);
#pragma unroll
                    for (int i = 0; i < S; ++i)
#pragma unroll
                        for (int q = 0; q < RC; ++q) acc[q] = fma(v[i], st[(k + i * G) * RC + q], acc[q]);
                }
                for (; k < m; k += G) {
                    double dx = tx - sx[k], dy = ty - sy[k];
                    double v = ker.raw(fma(dx, dx, dy * dy), logt);
#pragma unroll
                    for (int q = 0; q < RC; ++q) acc[q] = fma(v, st[k * RC + q], acc[q]);
                }
            }
        }
    }
    if (G > 1) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RC; ++q) red[lane * RC + q] = acc[q];
        __syncthreads();
        if (g == 0)
#pragma unroll
            for (int q = 0; q < RC; ++q) {
                double sum = 0.0;
                for (int gg = 0; gg < G; ++gg) sum += red[(gg * p + j) * RC + q];
                acc[q] = sum;
            }
    }
    if (lane < P)
#pragma unroll
        for (int q = 0; q < RC; ++q)
            U[((size_t)rid[b] * R + r0 + q) * P + lane] = lane < p ? ker.pref() * acc[q] : 0.0;
}

// Stored P2M values (the reference planner's stored upward leaf interactions, fmm.py:646-711):
// box i of the list at moff[i], p values per source, the arithmetic of k_p2m.
template <int KIND>
__global__ void __launch_bounds__(64) k_p2m_store(const int32_t* __restrict__ list, int n, const int32_t* __restrict__ bop,
                                                  const double* __restrict__ cx, const double* __restrict__ cy,
                                                  const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                                  const double* __restrict__ xs, const double* __restrict__ ys,
                                                  const double* __restrict__ surf, int p, int P, KerT<KIND> ker,
                                                  const int64_t* __restrict__ moff, double* __restrict__ M) {
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    stage_logt<KIND>(logt);
    __syncthreads();
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int32_t b = list[i], b0 = bot[b];
        const int l = bop[b];
        const int64_t cnt = (int64_t)(top[b] - b0) * p;
        for (int64_t e = threadIdx.x; e < cnt; e += 64) {
            const int k = (int)(e / p), j = (int)(e - (int64_t)k * p);
            const double tx = cx[b] + surf[(size_t)(2 * l) * P + j], ty = cy[b] + surf[(size_t)(2 * l + 1) * P + j];
            double dx = tx - xs[b0 + k], dy = ty - ys[b0 + k];
            M[moff[i] + e] = ker.raw(fma(dx, dx, dy * dy), logt);
        }
    }
}

// Upward pass of one level (reference class_fmm.py:309-319): U = M2M(children X) for boxes
// with children, U = P2M result for the others, then X = E^{-1} U = F diag(1/lambda) F^T U.
// Tile = 16 W / R boxes of the level list (boxes with children first) x R right-hand sides, W
// waves per workgroup (4; 8 for the large levels of wide batches, whose boxes share the level's
// operators: one staging per 128 rows); the list holds slab rows, crow the rows of each box's
// four children (-1: no sources, read as zero).  The next operator is fetched into registers
// while the current one is applied.
template <int NB, int W = 4>
__global__ void __launch_bounds__(64 * W) k_up(const int32_t* __restrict__ urow, const int32_t* __restrict__ crow,
                                            int n, int n_m2m, int lgR, const double* __restrict__ ops,
                                            const double* __restrict__ Ug, double* __restrict__ X,
                                            float* __restrict__ X32) {
    constexpr int P = 16 * NB;
    __shared__ __attribute__((aligned(32))) double Bs[P * P];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int bpt = (16 * W) >> lgR;
    const int j = 16 * wave + (lane & 15);
    const int i = blockIdx.x * bpt + (j >> lgR), r = j & ((1 << lgR) - 1);
    const int32_t b = i < n ? urow[i] : -1;
    const bool m2m = i < n_m2m;
    const bool any_m2m = blockIdx.x * bpt < n_m2m;
    const long row = b >= 0 ? ((long)b << lgR) + r : -1L;
    OpStage<NB, double, 64 * W> st;
    d4 acc[NB], R[NB], Rn[NB];
    rows_zero(acc);
    st.load(ops + (any_m2m ? LevelOps::m2mb(P, 0) : LevelOps::F(P)));
    if (any_m2m) {
        int32_t c = m2m ? crow[4 * (size_t)i] : -1;
        rows_load(R, X, c >= 0 ? ((long)c << lgR) + r : -1L);
        for (int q = 0; q < 4; ++q) {
            __syncthreads();
            st.store(Bs);
            __syncthreads();
            st.load(ops + (q < 3 ? LevelOps::m2mb(P, q + 1) : LevelOps::F(P)));
            if (q < 3) {
                c = m2m ? crow[4 * (size_t)i + q + 1] : -1;
                rows_load(Rn, X, c >= 0 ? ((long)c << lgR) + r : -1L);
            }
            if (__ballot(m2m)) rows_gemm(acc, R, Bs, P);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) R[nb] = Rn[nb];
        }
    }
    if (!m2m) rows_add(acc, Ug, row);
    __syncthreads();
    st.store(Bs);
    __syncthreads();
    st.load(ops + LevelOps::FT(P));
    rows_zero(R);
    rows_gemm(R, acc, Bs, P);
    rows_scale(R, ops + LevelOps::invlam(P));
    __syncthreads();
    st.store(Bs);
    __syncthreads();
    rows_zero(acc);
    rows_gemm(acc, R, Bs, P);
    rows_store(acc, X, row);
    if (X32 && row >= 0) {  // split precision: FP32 copy for the M2L pass
        const int kq = lane >> 4;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v) X32[row * P + rcol(0.0, nb, kq, v)] = (float)acc[nb][v];
    }
}

// The upward pass of one level in short tiles: 16 * RB rows (box x right-hand side, the
// global row g = 16 RB blockIdx.x + j holds box g >> lgR, rhs g & (R-1)) per workgroup of NB
// waves, wave w forming column block w of every product (the column-split form).  A level with
// few boxes then spreads over 4 / RB times as many workgroups as k_up's 64-row tiles, and each
// wave's dependent MFMA chain is NB times shorter; the whole rows a product needs come from
// the children's slab rows (M2M) or, between the products, from an LDS exchange.  Same
// arithmetic, in the same order, as k_up: the expansions are bitwise k_up's.
template <int NB, int RB>
__global__ void __launch_bounds__(256) k_up_cs(const int32_t* __restrict__ urow, const int32_t* __restrict__ crow,
                                               int n, int n_m2m, int lgR, const double* __restrict__ ops,
                                               const double* __restrict__ Ug, double* __restrict__ X,
                                               float* __restrict__ X32) {
    constexpr int P = 16 * NB, LD = cols_lds_ld<NB>();
    __shared__ __attribute__((aligned(16))) double Us[2][16 * RB * LD];
    __shared__ __attribute__((aligned(16))) double Ch[4][16 * RB * LD];  // the children's rows
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long g0 = (long)(16 * RB) * blockIdx.x;
    const int rmask = (1 << lgR) - 1;
    const bool any_m2m = (int)(g0 >> lgR) < n_m2m;  // boxes with children come first
    // one round of index loads: each row's list entry and its box's four children (one int4)
    int ib[RB];
    long row[RB];
    int4 ch[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const long g = g0 + 16 * rb + (lane & 15);
        ib[rb] = (int)(g >> lgR);
        const int32_t b = ib[rb] < n ? urow[ib[rb]] : -1;
        row[rb] = b >= 0 ? ((long)b << lgR) + (g & rmask) : -1L;
        ch[rb] = any_m2m && ib[rb] < n_m2m ? reinterpret_cast<const int4*>(crow)[ib[rb]] : int4{-1, -1, -1, -1};
    }
    d4 acc[RB], A[NB], An[NB], R[RB][NB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) acc[rb] = d4{0, 0, 0, 0};
    if (any_m2m) {
        // the operator slices of M2M 0 and 1 and every child row in one memory round: wave w
        // loads children w, w + NB, ... and the rows meet in LDS
        cols_load(A, ops + LevelOps::m2mb(P, 0), w);
        cols_load(An, ops + LevelOps::m2mb(P, 1), w);
        for (int q = w; q < 4; q += NB) {
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                const int32_t c = q == 0 ? ch[rb].x : q == 1 ? ch[rb].y : q == 2 ? ch[rb].z : ch[rb].w;
                rows_load(R[rb], X, c >= 0 ? ((long)c << lgR) + ((g0 + 16 * rb + (lane & 15)) & rmask) : -1L);
                rows_to_lds(Ch[q], R[rb], rb);
            }
        }
        __syncthreads();
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) rows_from_lds(R[rb], Ch[q], rb);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int v = 0; v < 4; ++v)
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb) acc[rb] = mfma_f64(A[nb][v], R[rb][nb][v], acc[rb]);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) A[nb] = An[nb];
            if (q < 3) cols_load(An, ops + (q < 2 ? LevelOps::m2mb(P, q + 2) : LevelOps::F(P)), w);
        }
    } else {
        cols_load(A, ops + LevelOps::F(P), w);
    }
    // X = ((U F) diag(1/lambda)) F^T, U = M2M sum or the P2M check potential
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        if (ib[rb] >= n_m2m) col_add<NB>(acc[rb], Ug, row[rb], w);
        cols_to_lds<NB>(Us[0], acc[rb], rb, w);
    }
    cols_load(An, ops + LevelOps::FT(P), w);
    __syncthreads();
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        rows_from_lds(R[rb], Us[0], rb);
        acc[rb] = d4{0, 0, 0, 0};
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) acc[rb] = mfma_f64(A[nb][v], R[rb][nb][v], acc[rb]);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        col_scale(acc[rb], ops + LevelOps::invlam(P), w);
        cols_to_lds<NB>(Us[1], acc[rb], rb, w);
    }
    __syncthreads();
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        rows_from_lds(R[rb], Us[1], rb);
        acc[rb] = d4{0, 0, 0, 0};
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) acc[rb] = mfma_f64(An[nb][v], R[rb][nb][v], acc[rb]);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        col_store<NB>(acc[rb], X, row[rb], w);
        if (X32) col_store<NB>(acc[rb], X32, row[rb], w);  // split precision: FP32 copy for the M2L
    }
}

// Index data of an M2L tile in LDS: the 27 interaction-list source rows of each of its targets
// and the operator of each slot, staged once at the start (a barrier follows) so that the slot
// loop issues no dependent index loads between its operator and row fetches.
struct M2lIndex {
    int32_t* src;  // [target in tile][27]
    int32_t* op;   // [27]
    __device__ __forceinline__ void stage(const int4& tile, const int32_t* __restrict__ vsrc,
                                          const int16_t* __restrict__ slot_op) {
        __shared__ int32_t s_src[64 * 27], s_op[27];
        src = s_src;
        op = s_op;
        for (int e = threadIdx.x; e < tile.y * 27; e += blockDim.x) s_src[e] = vsrc[(size_t)tile.x * 27 + e];
        if (threadIdx.x < 27) s_op[threadIdx.x] = slot_op[tile.z * 27 + threadIdx.x];
        __syncthreads();
    }
};

// s_waitcnt vmcnt(0): every vector-memory load of this wave has completed.  At the top of a
// pipelined loop it tells the compiler's wait analysis that nothing issued before is still in
// flight, so the loop body waits only for what it uses.
__device__ __forceinline__ void kf_wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// M2L of every level in one launch (reference class_fmm.py:336-343 with the collected
// operators split back into the 40 offsets): D = sum over the interaction-list slots of
// M2L_o X_source.  tgt and vsrc hold slab rows.  A tile holds up to 64/R targets of one level with the same child position q
// (so a slot means the same operator for all of them) x R right-hand sides; `mask` lists the
// slots any target uses.  Operator s+1 and the source rows of slot s+1 are fetched while slot s
// is applied.
#ifndef KF_M2L_WAVES
#define KF_M2L_WAVES 3  // waves per SIMD the FP64 M2L is compiled for (registers <= 512 / waves)
#endif
template <int NB>
__global__ void __launch_bounds__(256, KF_M2L_WAVES) k_m2l(const int4* __restrict__ tiles, const uint32_t* __restrict__ masks,
                                             const int32_t* __restrict__ tgt, const int32_t* __restrict__ vsrc,
                                             const int16_t* __restrict__ slot_op, const double* __restrict__ ops,
                                             size_t ops_stride, int lgR, const double* __restrict__ X,
                                             double* __restrict__ D) {
    constexpr int P = 16 * NB;
    __shared__ __attribute__((aligned(32))) double Bs[P * P];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ti = xcd_contiguous(blockIdx.x, gridDim.x);
    const int4 tile = tiles[ti];
    const uint32_t mask = masks[ti];
    const int q = tile.z;
    const double* opl = ops + (size_t)tile.w * ops_stride;
    const int j = 16 * wave + (lane & 15);
    const int k = j >> lgR, r = j & ((1 << lgR) - 1);
    const bool valid = k < tile.y;
    const int t = valid ? tile.x + k : -1;
    const int32_t b = valid ? tgt[t] : -1;
    M2lIndex ix;
    ix.stage(tile, vsrc, slot_op);
    OpStage<NB, double> st;
    d4 acc[NB], R[NB], Rn[NB];
    rows_zero(acc);
    int s = mask ? __ffs(mask) - 1 : -1;
    int32_t src = -1;
    if (s >= 0) {
        st.load(opl + LevelOps::m2lt(P, ix.op[s]));
        src = valid ? ix.src[k * 27 + s] : -1;
        rows_load(R, X, src >= 0 ? ((long)src << lgR) + r : -1L);
    }
    // slot s: its operator is in st and its source rows in Rc; fetch slot s+1's into st / Rx
    auto slot = [&](const d4 (&Rc)[NB], d4 (&Rx)[NB]) {
        kf_wait_vmem();  // slot s's operator and rows are in; nothing older stays pending
        __syncthreads();
        st.store(Bs);
        __syncthreads();
        const uint32_t rest = mask & ~((2u << s) - 1u);
        const int sn = rest ? __ffs(rest) - 1 : -1;
        int32_t srcn = -1;
        if (sn >= 0) {
            st.load(opl + LevelOps::m2lt(P, ix.op[sn]));
            srcn = valid ? ix.src[k * 27 + sn] : -1;
            rows_load(Rx, X, srcn >= 0 ? ((long)srcn << lgR) + r : -1L);
        }
        if (__ballot(src >= 0)) rows_gemm(acc, Rc, Bs, P);
        src = srcn;
        s = sn;
    };
    while (s >= 0) {  // two slots per trip, the row registers alternating (no copies)
        slot(R, Rn);
        if (s < 0) break;
        slot(Rn, R);
    }
    rows_store(acc, D, b >= 0 ? ((long)b << lgR) + r : -1L);
}

// Split-precision M2L: the same tiles and slots as k_m2l, with FP32 operators (the 40 blocks
// of each op -- level and size class -- scaled to unit max-norm; scale[op] restores them, op
// being the tile's w field) and FP32 densities X32.  Each
// slot's product is formed on FP32 MFMA (FP32 accumulation over its p terms) and the slots'
// products are summed in FP64; the result is written to D in FP64 (the L2L + C2E pass stays
// FP64).
template <int NB>
__global__ void __launch_bounds__(256) k_m2l_f32(const int4* __restrict__ tiles, const uint32_t* __restrict__ masks,
                                                 const int32_t* __restrict__ tgt, const int32_t* __restrict__ vsrc,
                                                 const int16_t* __restrict__ slot_op, const float* __restrict__ ops32,
                                                 const double* __restrict__ scale, int lgR,
                                                 const float* __restrict__ X32, double* __restrict__ D) {
    constexpr int P = 16 * NB;
    __shared__ __attribute__((aligned(32))) float Bs[P * P];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ti = xcd_contiguous(blockIdx.x, gridDim.x);
    const int4 tile = tiles[ti];
    const uint32_t mask = masks[ti];
    const int q = tile.z;
    const float* opl = ops32 + (size_t)tile.w * 40 * P * P;
    const int j = 16 * wave + (lane & 15);
    const int k = j >> lgR, r = j & ((1 << lgR) - 1);
    const bool valid = k < tile.y;
    const int t = valid ? tile.x + k : -1;
    const int32_t b = valid ? tgt[t] : -1;
    M2lIndex ix;
    ix.stage(tile, vsrc, slot_op);
    OpStage<NB, float> st;
    f4 part[NB], R[NB], Rn[NB];
    d4 acc[NB];  // FP64 sum of the slots' FP32 products
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = d4{0, 0, 0, 0};
    int s = mask ? __ffs(mask) - 1 : -1;
    int32_t src = -1;
    if (s >= 0) {
        st.load(opl + (size_t)ix.op[s] * P * P);
        src = valid ? ix.src[k * 27 + s] : -1;
        rows_load(R, X32, src >= 0 ? ((long)src << lgR) + r : -1L);
    }
    auto slot = [&](const f4 (&Rc)[NB], f4 (&Rx)[NB]) {
        kf_wait_vmem();  // slot s's operator and rows are in; nothing older stays pending
        __syncthreads();
        st.store(Bs);
        __syncthreads();
        const uint32_t rest = mask & ~((2u << s) - 1u);
        const int sn = rest ? __ffs(rest) - 1 : -1;
        int32_t srcn = -1;
        if (sn >= 0) {
            st.load(opl + (size_t)ix.op[sn] * P * P);
            srcn = valid ? ix.src[k * 27 + sn] : -1;
            rows_load(Rx, X32, srcn >= 0 ? ((long)srcn << lgR) + r : -1L);
        }
        if (__ballot(src >= 0)) {
            rows_zero(part);
            rows_gemm(part, Rc, Bs, P);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[nb][v] += (double)part[nb][v];
        }
        src = srcn;
        s = sn;
    };
    while (s >= 0) {  // two slots per trip, the row registers alternating (no copies)
        slot(R, Rn);
        if (s < 0) break;
        slot(Rn, R);
    }
    if (b >= 0) {
        const long row = ((long)b << lgR) + r;
        const double sc = scale[tile.w];
        const int kq = lane >> 4;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v) D[row * P + rcol(0.0f, nb, kq, v)] = sc * (double)acc[nb][v];
    }
}

// FP16-split precision: the M2L of k_m2l_f32 with each slot's product formed on
// v_mfma_f32_16x16x32_f16 as hi*hi + lo*hi + hi*lo of FP16 planes (the lo*lo term, below 2^-22
// of the product, is left out), FP32 accumulation within a slot, the slots summed in FP64.
// The operator planes come split (build_ops16_dev: hi = fp16(v), lo = fp16(v - hi) of the
// scaled block, transposed to [m][k]).  The FP32 source rows are split in registers: each row
// is scaled by a power of two s with max |x| / s in [1/2, 1) (a row's values sit in the four
// lanes l%16 + 16 q, so two swaps give the maximum), y = x / s, hi = fp16(y), lo = fp16(y - hi),
// so hi + lo carries y to 2^-24 (the matrix cores keep FP16 subnormals) and s multiplies the
// slot's product back.  Splitting in the M2L costs VALU time beside the matrix cores but no
// pass and no buffer of its own; the kernel is bound by the operators and rows it streams
// through L2 (both 4 bytes a value, as in k_m2l_f32).  Lane l of wave w holds tile row
// 16w + l%16; the B operand of K chunk c is its row's values 32c + 8(l/16) .. + 7 (values past
// P are zeros), the A operand the operator plane's row 16mb + l%16 at the same k (LDS, rows
// padded by 8 halves against bank conflicts); the output lands in the FP32 layout of kf_mfma.h.
// W waves per workgroup (64 W threads, 16 W tile rows): 8 for the tile pairs of wide batches,
// whose targets share the slot operators, so that one staging serves both.
template <int NB, int W>
__global__ void __launch_bounds__(64 * W) k_m2l_f16(const int4* __restrict__ tiles, const uint32_t* __restrict__ masks,
                                                 const int32_t* __restrict__ tgt, const int32_t* __restrict__ vsrc,
                                                 const int16_t* __restrict__ slot_op, const _Float16* __restrict__ ops16,
                                                 const double* __restrict__ scale, int lgR,
                                                 const float* __restrict__ X32, double* __restrict__ D) {
    typedef _Float16 h8 __attribute__((ext_vector_type(8)));
    typedef float f8 __attribute__((ext_vector_type(8)));
    constexpr int P = 16 * NB, PK = (P + 31) / 32 * 32, NC = PK / 32, LD = PK + 8;
    constexpr int OPV = 2 * P * PK / 8;                  // 16-byte vectors of one slot's two planes
    constexpr int TH = 64 * W, PER = (OPV + TH - 1) / TH;
    __shared__ __attribute__((aligned(16))) _Float16 Bs[2 * P * LD];  // hi plane, then lo plane
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, i16 = lane & 15, q4 = lane >> 4;
    const int ti = xcd_contiguous(blockIdx.x, gridDim.x);
    const int4 tile = tiles[ti];
    const uint32_t mask = masks[ti];
    const _Float16* opl = ops16 + (size_t)tile.w * 40 * 2 * P * PK;
    const int j = 16 * wave + i16;
    const int k = j >> lgR, r = j & ((1 << lgR) - 1);
    const bool valid = k < tile.y;
    const int t = valid ? tile.x + k : -1;
    const int32_t b = valid ? tgt[t] : -1;
    M2lIndex ix;
    ix.stage(tile, vsrc, slot_op);
    h8 opv[PER];
    auto op_load = [&](int o) {
        const h8* src = reinterpret_cast<const h8*>(opl + (size_t)o * 2 * P * PK);
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = threadIdx.x + TH * u;
            if (OPV % TH == 0 || i < OPV) opv[u] = src[i];
        }
    };
    auto op_store = [&]() {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = threadIdx.x + TH * u;
            if (OPV % TH == 0 || i < OPV) {
                const int pl = i / (P * PK / 8), rem = i - pl * (P * PK / 8);
                const int m = rem / (PK / 8), kk = (rem - m * (PK / 8)) * 8;
                *reinterpret_cast<h8*>(Bs + pl * P * LD + m * LD + kk) = opv[u];
            }
        }
    };
    struct Row {  // the lane's FP32 values of a source row: chunk c holds k = 32c + 8(l/16) .. + 7
        f8 v[NC];
    };
    auto row_load = [&](Row& R, int32_t src) {
        const long row = src >= 0 ? ((long)src << lgR) + r : -1L;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int k0 = 32 * c + 8 * q4;
            const bool on = row >= 0 && k0 < P;  // P is a multiple of 16: a chunk is all in or all out
            const f8 z = {0, 0, 0, 0, 0, 0, 0, 0};
            R.v[c] = on ? *reinterpret_cast<const f8*>(X32 + row * P + k0) : z;  // 32-byte aligned
        }
    };
    d4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = d4{0, 0, 0, 0};
    int s = mask ? __ffs(mask) - 1 : -1;
    int32_t src = -1;
    Row Ra, Rb;
    if (s >= 0) {
        op_load(ix.op[s]);
        src = valid ? ix.src[k * 27 + s] : -1;
        row_load(Ra, src);
    }
    auto slot = [&](const Row& Rc, Row& Rx) {
        kf_wait_vmem();  // slot s's operator and rows are in; nothing older stays pending
        __syncthreads();
        op_store();
        __syncthreads();
        const uint32_t rest = mask & ~((2u << s) - 1u);
        const int sn = rest ? __ffs(rest) - 1 : -1;
        int32_t srcn = -1;
        if (sn >= 0) {
            op_load(ix.op[sn]);
            srcn = valid ? ix.src[k * 27 + sn] : -1;
            row_load(Rx, srcn);
        }
        if (__ballot(src >= 0)) {  // wave-uniform: every lane takes part in the swaps
            float mx = 0.0f;
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int u = 0; u < 8; ++u) mx = fmaxf(mx, fabsf(Rc.v[c][u]));
            mx = fmaxf(mx, __shfl_xor(mx, 16));
            mx = fmaxf(mx, __shfl_xor(mx, 32));
            int e = 0;
            (void)frexpf(mx, &e);  // mx = f 2^e with f in [1/2, 1); mx = 0 gives e = 0
            const float inv = ldexpf(1.0f, -e);
            h8 xh[NC], xl[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const float yu = Rc.v[c][u] * inv;
                    xh[c][u] = (_Float16)yu;
                    xl[c][u] = (_Float16)(yu - (float)xh[c][u]);
                }
            }
            const double sc = ldexp(1.0, e);
#pragma unroll
            for (int mb = 0; mb < NB; ++mb) {
                f4 part = f4{0, 0, 0, 0};
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const _Float16* arow = Bs + (16 * mb + i16) * LD + 32 * c + 8 * q4;
                    const h8 ah = *reinterpret_cast<const h8*>(arow);
                    const h8 al = *reinterpret_cast<const h8*>(arow + P * LD);
                    part = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, xh[c], part, 0, 0, 0);
                    part = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, xh[c], part, 0, 0, 0);
                    part = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, xl[c], part, 0, 0, 0);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[mb][v] = fma(sc, (double)part[v], acc[mb][v]);
            }
        }
        src = srcn;
        s = sn;
    };
    while (s >= 0) {  // two slots per trip, the row registers alternating (no copies)
        slot(Ra, Rb);
        if (s < 0) break;
        slot(Rb, Ra);
    }
    if (b >= 0) {
        const long row = ((long)b << lgR) + r;
        const double sc = scale[tile.w];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v) D[row * P + rcol(0.0f, nb, q4, v)] = sc * acc[nb][v];
    }
}

// Downward pass of one level (reference class_fmm.py:328-334, 348-350): D += L2L(parent Psi)
// (levels >= 3; the level-1 expansions are zero), then Psi = E^{-1} D.  W waves per workgroup,
// 16 W rows of a tile (W = 8: the tile pairs of wide batches, one op and child position).
template <int NB, int W = 4>
__global__ void __launch_bounds__(64 * W) k_l2l_c2e(const int4* __restrict__ tiles, const int32_t* __restrict__ tgt,
                                                 const int32_t* __restrict__ par, const double* __restrict__ ops,
                                                 int do_l2l, int lgR, const double* __restrict__ D,
                                                 double* __restrict__ Psi) {
    constexpr int P = 16 * NB;
    __shared__ __attribute__((aligned(32))) double Bs[P * P];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int4 tile = tiles[blockIdx.x];
    const int j = 16 * wave + (lane & 15);
    const int k = j >> lgR, r = j & ((1 << lgR) - 1);
    const bool valid = k < tile.y;
    const int t = valid ? tile.x + k : -1;
    const int32_t b = valid ? tgt[t] : -1;
    const long row = b >= 0 ? ((long)b << lgR) + r : -1L;
    OpStage<NB, double, 64 * W> st;
    d4 acc[NB], R[NB];
    rows_load(acc, D, row);
    st.load(ops + (do_l2l ? LevelOps::l2lb(P, tile.z) : LevelOps::F(P)));
    if (do_l2l) {
        const int32_t pr = valid ? par[t] : -1;
        rows_load(R, Psi, pr >= 0 ? ((long)pr << lgR) + r : -1L);
        __syncthreads();
        st.store(Bs);
        __syncthreads();
        st.load(ops + LevelOps::F(P));
        rows_gemm(acc, R, Bs, P);
    }
    __syncthreads();
    st.store(Bs);
    __syncthreads();
    st.load(ops + LevelOps::FT(P));
    rows_zero(R);
    rows_gemm(R, acc, Bs, P);
    rows_scale(R, ops + LevelOps::invlam(P));
    __syncthreads();
    st.store(Bs);
    __syncthreads();
    rows_zero(acc);
    rows_gemm(acc, R, Bs, P);
    rows_store(acc, Psi, row);
}

// k_l2l_c2e in the column-split form of k_up_cs: each 64-row tile of the table is taken by
// 4 / RB workgroups of 16 * RB rows (sub-tiles past the tile's targets return at once), NB
// waves each forming one column block.  Bitwise k_l2l_c2e's expansions.
template <int NB, int RB>
__global__ void __launch_bounds__(256) k_l2l_c2e_cs(const int4* __restrict__ tiles, const int32_t* __restrict__ tgt,
                                                    const int32_t* __restrict__ par, const double* __restrict__ ops,
                                                    int do_l2l, int lgR, const double* __restrict__ D,
                                                    double* __restrict__ Psi) {
    constexpr int P = 16 * NB, LD = cols_lds_ld<NB>(), SUB = 4 / RB;
    __shared__ __attribute__((aligned(16))) double Us[2][16 * RB * LD];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int4 tile = tiles[blockIdx.x / SUB];
    const int j0 = 16 * RB * (blockIdx.x % SUB);
    if ((j0 >> lgR) >= tile.y) return;  // no target in this sub-tile (the whole workgroup)
    const int rmask = (1 << lgR) - 1;
    long row[RB], prow[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const int j = j0 + 16 * rb + (lane & 15);
        const int k = j >> lgR, r = j & rmask;
        const bool valid = k < tile.y;
        const int32_t b = valid ? tgt[tile.x + k] : -1;
        const int32_t pr = valid && do_l2l ? par[tile.x + k] : -1;
        row[rb] = b >= 0 ? ((long)b << lgR) + r : -1L;
        prow[rb] = pr >= 0 ? ((long)pr << lgR) + r : -1L;
    }
    d4 acc[RB], A[NB], An[NB], R[RB][NB];
    // the first two operator slices do not depend on the indices: in flight with them
    cols_load(A, ops + (do_l2l ? LevelOps::l2lb(P, tile.z) : LevelOps::F(P)), w);
    cols_load(An, ops + (do_l2l ? LevelOps::F(P) : LevelOps::FT(P)), w);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) col_load<NB>(acc[rb], D, row[rb], w);
    if (do_l2l) {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) rows_load(R[rb], Psi, prow[rb]);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) acc[rb] = mfma_f64(A[nb][v], R[rb][nb][v], acc[rb]);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) A[nb] = An[nb];
        cols_load(An, ops + LevelOps::FT(P), w);
    }
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) cols_to_lds<NB>(Us[0], acc[rb], rb, w);
    __syncthreads();
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        rows_from_lds(R[rb], Us[0], rb);
        acc[rb] = d4{0, 0, 0, 0};
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) acc[rb] = mfma_f64(A[nb][v], R[rb][nb][v], acc[rb]);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        col_scale(acc[rb], ops + LevelOps::invlam(P), w);
        cols_to_lds<NB>(Us[1], acc[rb], rb, w);
    }
    __syncthreads();
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        rows_from_lds(R[rb], Us[1], rb);
        acc[rb] = d4{0, 0, 0, 0};
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) acc[rb] = mfma_f64(An[nb][v], R[rb][nb][v], acc[rb]);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) col_store<NB>(acc[rb], Psi, row[rb], w);
}

// ---------------------------------------------------------------- persistent chains ----
// The upward and downward passes in one launch each instead of one per level.  Work items
// (16 rows of a level's list) are taken in order from a queue in the control block; the items
// are ordered so that every item depends only on earlier ones (upward: deepest level first;
// downward: top level first), so among the running workgroups there is always one whose inputs
// are complete, whatever the number of resident workgroups.  The rows one item writes and
// another reads (upward X, downward Psi) are stored and loaded as agent-scope relaxed atomics:
// on this device the stores write through the XCD's L2 and the loads miss it, so no fence has
// to write back or invalidate a whole L2 (an agent-scope release / acquire does both, and at
// one per item they cost more than the per-level launches save).  A finished item publishes its
// rows by writing the evaluation's epoch into their flags once its stores have completed (every
// wave waits for its memory counter, then the workgroup barrier); an item waits for the flags
// of the rows it reads with a bounded poll.  Same arithmetic in the same order as the per-level
// kernels k_up_cs / k_l2l_c2e_cs (RB = 1): the expansions are bitwise theirs.
#ifndef KF_CHAIN_SPIN_LIMIT
#define KF_CHAIN_SPIN_LIMIT (1 << 22)  // polls (each with a short sleep) before a wait gives up
#endif

// Control block: [0] evaluation epoch, [1] upward queue, [2] downward queue, [3] error bits
// (1: a wait gave up; the expansions of that evaluation are then wrong, and nothing hangs).
__global__ void k_chain_begin(int* __restrict__ ctl) {
    if (threadIdx.x == 0) {
        ctl[0] += 1;
        ctl[1] = 0;
        ctl[2] = 0;
    }
}

__device__ __forceinline__ void chain_wait(const int* flag, int epoch, int* err) {
    int spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
        if (++spins > KF_CHAIN_SPIN_LIMIT) {
            atomicOr(err, 1);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// The next item of a queue, for the whole workgroup (the barrier also ends the previous item's
// use of LDS).
__device__ __forceinline__ int chain_next(int* q) {
    __shared__ int item;
    __syncthreads();
    if (threadIdx.x == 0) item = atomicAdd(q, 1);
    __syncthreads();
    return item;
}

// Publishes the rows of the item (row[lane & 15] of the first wave) once every wave's
// write-through stores have completed.
__device__ __forceinline__ void chain_publish(int* done, long row, int epoch) {
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are acknowledged
    __syncthreads();
    if (threadIdx.x < 16 && row >= 0) __hip_atomic_store(done + row, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Row and column-block access for rows shared between the items of a chain (agent-scope
// relaxed atomics: write-through stores, loads that miss the XCD's L2; values unchanged).
template <int NB>
__device__ __forceinline__ void rows_load_coh(d4 (&R)[NB], const double* slab, long row) {
    const int kq = (threadIdx.x & 63) >> 4;
    constexpr int P = 16 * NB;
    if (row >= 0) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                R[nb][v] = __hip_atomic_load(slab + row * P + rcol(0.0, nb, kq, v), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    } else {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) R[nb] = d4{0, 0, 0, 0};
    }
}
template <int NB>
__device__ __forceinline__ void col_store_coh(const d4& a, double* slab, long row, int w) {
    const int kq = (threadIdx.x & 63) >> 4;
    constexpr int P = 16 * NB;
    if (row < 0) return;
#pragma unroll
    for (int v = 0; v < 4; ++v)
        __hip_atomic_store(slab + row * P + rcol(0.0, w, kq, v), a[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Upward pass of every level: items (op, 16-row tile of the op's list), deepest level first.
template <int NB>
__global__ void __launch_bounds__(256) k_up_chain(const int2* __restrict__ items, int n_items,
                                                  const int32_t* __restrict__ urow_all, const int32_t* __restrict__ crow_all,
                                                  const int* __restrict__ up_off, const int* __restrict__ up_n,
                                                  const int* __restrict__ up_nm2m, int lgR, const double* __restrict__ ops_all,
                                                  const double* __restrict__ Ug, double* __restrict__ X,
                                                  float* __restrict__ X32, int* __restrict__ done, int* __restrict__ ctl) {
    constexpr int P = 16 * NB, LD = cols_lds_ld<NB>();
    __shared__ __attribute__((aligned(16))) double Us[2][16 * LD];
    __shared__ __attribute__((aligned(16))) double Ch[4][16 * LD];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int epoch = ctl[0];
    const int rmask = (1 << lgR) - 1;
    for (;;) {
        const int it = chain_next(&ctl[1]);
        if (it >= n_items) break;
        const int2 I = items[it];
        const int op = I.x;
        const int32_t* urow = urow_all + up_off[op];
        const int32_t* crow = crow_all + 4 * (size_t)up_off[op];
        const int n = up_n[op], n_m2m = up_nm2m[op];
        const double* ops = ops_all + (size_t)op * LevelOps::size(P);
        const long g0 = 16L * I.y;
        const bool any_m2m = (int)(g0 >> lgR) < n_m2m;
        const long g = g0 + (lane & 15);
        const int ib = (int)(g >> lgR);
        const int32_t b = ib < n ? urow[ib] : -1;
        const long row = b >= 0 ? ((long)b << lgR) + (g & rmask) : -1L;
        const int4 ch = any_m2m && ib < n_m2m ? reinterpret_cast<const int4*>(crow)[ib] : int4{-1, -1, -1, -1};
        d4 acc = d4{0, 0, 0, 0}, A[NB], An[NB], R[NB];
        if (any_m2m) {
            cols_load(A, ops + LevelOps::m2mb(P, 0), w);
            cols_load(An, ops + LevelOps::m2mb(P, 1), w);
            for (int q = w; q < 4; q += NB) {
                const int32_t c = q == 0 ? ch.x : q == 1 ? ch.y : q == 2 ? ch.z : ch.w;
                const long cr = c >= 0 ? ((long)c << lgR) + (g & rmask) : -1L;
                if (cr >= 0) chain_wait(done + cr, epoch, ctl + 3);  // the child rows are read after their flags
                rows_load_coh(R, X, cr);
                rows_to_lds(Ch[q], R, 0);
            }
            __syncthreads();
            for (int q = 0; q < 4; ++q) {
                rows_from_lds(R, Ch[q], 0);
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc = mfma_f64(A[nb][v], R[nb][v], acc);
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) A[nb] = An[nb];
                if (q < 3) cols_load(An, ops + (q < 2 ? LevelOps::m2mb(P, q + 2) : LevelOps::F(P)), w);
            }
        } else {
            cols_load(A, ops + LevelOps::F(P), w);
        }
        if (ib >= n_m2m) col_add<NB>(acc, Ug, row, w);
        cols_to_lds<NB>(Us[0], acc, 0, w);
        cols_load(An, ops + LevelOps::FT(P), w);
        __syncthreads();
        rows_from_lds(R, Us[0], 0);
        acc = d4{0, 0, 0, 0};
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc = mfma_f64(A[nb][v], R[nb][v], acc);
        col_scale(acc, ops + LevelOps::invlam(P), w);
        cols_to_lds<NB>(Us[1], acc, 0, w);
        __syncthreads();
        rows_from_lds(R, Us[1], 0);
        acc = d4{0, 0, 0, 0};
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc = mfma_f64(An[nb][v], R[nb][v], acc);
        col_store_coh<NB>(acc, X, row, w);
        if (X32) col_store<NB>(acc, X32, row, w);  // read by the M2L launch only
        chain_publish(done, row, epoch);
    }
}

// Downward pass (L2L + C2E) of every level: items (tile of the downward tile table, 16-row
// quarter), top level first (the table lists the tiles level by level); a target waits for its
// parent's local expansion.  ops index: the tile's op field; L2L below level 2 only.
template <int NB>
__global__ void __launch_bounds__(256) k_down_chain(const int4* __restrict__ tiles, int n_tiles,
                                                    const int32_t* __restrict__ tgt, const int32_t* __restrict__ par,
                                                    const double* __restrict__ ops_all, int L, int lgR,
                                                    const double* __restrict__ D, double* __restrict__ Psi,
                                                    int* __restrict__ done, int* __restrict__ ctl) {
    constexpr int P = 16 * NB, LD = cols_lds_ld<NB>();
    __shared__ __attribute__((aligned(16))) double Us[2][16 * LD];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int epoch = ctl[0];
    const int rmask = (1 << lgR) - 1;
    for (;;) {
        const int it = chain_next(&ctl[2]);
        if (it >= 4 * n_tiles) break;
        const int4 tile = tiles[it >> 2];
        const int j0 = 16 * (it & 3);
        if ((j0 >> lgR) >= tile.y) continue;  // no target in this quarter (the whole workgroup)
        const int op = tile.w;
        const int do_l2l = op % L >= 3;
        const double* ops = ops_all + (size_t)op * LevelOps::size(P);
        const int j = j0 + (lane & 15);
        const int k = j >> lgR, r = j & rmask;
        const bool valid = k < tile.y;
        const int32_t bt = valid ? tgt[tile.x + k] : -1;
        const int32_t pr = valid && do_l2l ? par[tile.x + k] : -1;
        const long row = bt >= 0 ? ((long)bt << lgR) + r : -1L;
        const long prow = pr >= 0 ? ((long)pr << lgR) + r : -1L;
        d4 acc, A[NB], An[NB], R[NB];
        cols_load(A, ops + (do_l2l ? LevelOps::l2lb(P, tile.z) : LevelOps::F(P)), w);
        cols_load(An, ops + (do_l2l ? LevelOps::F(P) : LevelOps::FT(P)), w);
        col_load<NB>(acc, D, row, w);
        if (do_l2l) {
            if (prow >= 0) chain_wait(done + prow, epoch, ctl + 3);  // the parent's expansion is read after its flag
            rows_load_coh(R, Psi, prow);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc = mfma_f64(A[nb][v], R[nb][v], acc);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) A[nb] = An[nb];
            cols_load(An, ops + LevelOps::FT(P), w);
        }
        cols_to_lds<NB>(Us[0], acc, 0, w);
        __syncthreads();
        rows_from_lds(R, Us[0], 0);
        acc = d4{0, 0, 0, 0};
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc = mfma_f64(A[nb][v], R[nb][v], acc);
        col_scale(acc, ops + LevelOps::invlam(P), w);
        cols_to_lds<NB>(Us[1], acc, 0, w);
        __syncthreads();
        rows_from_lds(R, Us[1], 0);
        acc = d4{0, 0, 0, 0};
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc = mfma_f64(An[nb][v], R[nb][v], acc);
        col_store_coh<NB>(acc, Psi, row, w);
        chain_publish(done, row, epoch);
    }
}

// Near field (reference fmm.py:30-64): for every real leaf, the direct sum over the sources of
// its 9 colleagues with index self-exclusion, into near[t*R + r] (tree order).  One 64-lane
// workgroup per (leaf, rhs chunk); a block of cnt <= 64 targets is held one or two per lane
// (whichever leaves fewer lanes idle: 40 targets fill 60 lanes two per lane, 40 one per lane),
// and the 64 / ceil(cnt / T) lane groups split the sources.  The leaf's own sources come first (the target's own point is skipped without a
// branch: its squared distance is mapped to 1 and its kernel value to 0), then the sources of
// the other colleagues, concatenated in LDS up to p2p_stage<RC>() at a time so that a leaf needs
// few barriers and long inner loops (the stage shrinks with the RHS chunk to keep the LDS per
// workgroup, and with it the occupancy, bounded).  The arithmetic of one right-hand side does
// not depend on the chunk width, so a density gives bitwise the same potential in any batch.
#ifndef KF_P2P_STAGE1
#define KF_P2P_STAGE1 128  // sources staged per round for one right-hand side (LDS per workgroup,
                           // and with it the resident waves; 128 measured 3% faster than 256)
#endif
template <int RC>
constexpr int p2p_stage() {
    return RC == 1 ? KF_P2P_STAGE1 : (RC <= 4 ? 128 : 64);
}

// S consecutive sources of a lane group (k, k + G, ...) for the group's T targets per lane: the
// S T kernel values side by side (raw_group), then each accumulator takes the sources in order.
template <int S, int KIND, int RC, int T, bool SELF>
__device__ __forceinline__ void p2p_sources(const KerT<KIND>& ker, const double* logt, const double (&tx)[T],
                                            const double (&ty)[T], const int (&tloc)[T], const double* sx,
                                            const double* sy, const double* st, int k, int G,
                                            double (&acc)[T][RC]) {
    double r2[S * T], v[S * T];
    bool me[S * T];
#pragma unroll
    for (int i = 0; i < S; ++i) {
        const int ki = k + i * G;
        const double sxk = sx[ki], syk = sy[ki];
#pragma unroll
        for (int u = 0; u < T; ++u) {
            const double dx = tx[u] - sxk, dy = ty[u] - syk;
            me[i * T + u] = SELF && (ki == tloc[u]);
            r2[i * T + u] = me[i * T + u] ? 1.0 : fma(dx, dx, dy * dy);
        }
    }
    raw_group(ker, r2, logt, v);
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
        for (int q = 0; q < RC; ++q) {
            const double sq = st[(k + i * G) * RC + q];
#pragma unroll
            for (int u = 0; u < T; ++u) acc[u][q] = fma(sq, me[i * T + u] ? 0.0 : v[i * T + u], acc[u][q]);
        }
}

#ifndef KF_P2P_VALUES
#define KF_P2P_VALUES 4  // kernel values per lane and step of the near field (sources x targets)
#endif
template <int KIND, int RC, int T, bool SELF>
__device__ __forceinline__ void p2p_chunk(const KerT<KIND>& ker, const double* logt, const double (&tx)[T],
                                          const double (&ty)[T], const int (&tloc)[T], const double* sx,
                                          const double* sy, const double* st, int k0, int m, int G,
                                          double (&acc)[T][RC]) {
    // S sources per step, KF_P2P_VALUES kernel values in all, then the remaining sources one
    // at a time; each accumulator takes the sources in order
    constexpr int S = KF_P2P_VALUES / T > 1 ? KF_P2P_VALUES / T : 1;
    int k = k0;
    for (; k + (S - 1) * G < m; k += S * G)
        p2p_sources<S, KIND, RC, T, SELF>(ker, logt, tx, ty, tloc, sx, sy, st, k, G, acc);
    for (; k < m; k += G) p2p_sources<1, KIND, RC, T, SELF>(ker, logt, tx, ty, tloc, sx, sy, st, k, G, acc);
}

// One block of up to 64 targets of a leaf with T targets per lane: groups of ceil(cnt / T)
// lanes, lane tl of a group holding targets tl and tl + ceil(cnt / T); group g takes the
// sources whose index (in the leaf, or in the virtual sequence of the other colleagues) is
// g mod G.  T depends only on cnt, so a target's summation order is the same for every RHS
// chunk width.
template <int T>
__device__ __forceinline__ bool p2p_use(int cnt) {
    const int two = (64 / ((cnt + 1) / 2)) * cnt, one = 2 * (64 / cnt) * cnt;  // targets per 128 lane slots
    return T == 2 ? two > one : two <= one;
}

template <int KIND, int RC, int T, int STAGE>
__device__ __forceinline__ void p2p_block(const KerT<KIND>& ker, const double* logt, int32_t tb, int nt, int t0,
                                          int cnt, int n_other, const int32_t* cbeg, const int32_t* cend,
                                          const double* __restrict__ xs, const double* __restrict__ ys,
                                          const double* __restrict__ tau, int R, int r0, double* sx, double* sy,
                                          double* st, double* red, double* __restrict__ near) {
    const int lane = threadIdx.x;
    const int Lg = (cnt + T - 1) / T;
    const int G = 64 / Lg, g = lane / Lg, tl = lane - g * Lg;
    const bool act = g < G;
    double tx[T], ty[T];
    int32_t tg[T];
    bool ok[T];
#pragma unroll
    for (int u = 0; u < T; ++u) {
        const int i = tl + u * Lg;
        ok[u] = act && i < cnt;
        tg[u] = tb + t0 + (ok[u] ? i : tl);  // a missing second target repeats the first
        tx[u] = act ? xs[tg[u]] : 0.0;
        ty[u] = act ? ys[tg[u]] : 0.0;
    }
    double acc[T][RC];
#pragma unroll
    for (int u = 0; u < T; ++u)
#pragma unroll
        for (int q = 0; q < RC; ++q) acc[u][q] = 0.0;
    // the leaf's own sources, with self-exclusion
    for (int32_t s0 = tb; s0 < tb + nt; s0 += STAGE) {
        const int m = min(STAGE, tb + nt - s0);
        __syncthreads();
        for (int e = lane; e < m; e += 64) {
            sx[e] = xs[s0 + e];
            sy[e] = ys[s0 + e];
        }
        for (int e = lane; e < m * RC; e += 64) st[e] = tau[(size_t)(s0 + e / RC) * R + r0 + e % RC];
        __syncthreads();
        if (act) {
            const int k0 = (g - (s0 - tb) % G + G) % G;
            int tloc[T];
            bool self = false;
#pragma unroll
            for (int u = 0; u < T; ++u) {
                tloc[u] = ok[u] && tg[u] >= s0 && tg[u] < s0 + m ? tg[u] - s0 : -1;
                self = self || tloc[u] >= 0;
            }
            if (self)
                p2p_chunk<KIND, RC, T, true>(ker, logt, tx, ty, tloc, sx, sy, st, k0, m, G, acc);
            else
                p2p_chunk<KIND, RC, T, false>(ker, logt, tx, ty, tloc, sx, sy, st, k0, m, G, acc);
        }
    }
    // the other colleagues' sources
    for (int v0 = 0; v0 < n_other; v0 += STAGE) {
        const int m = min(STAGE, n_other - v0);
        __syncthreads();
        for (int e = lane; e < m; e += 64) {
            int v = v0 + e, j = 0;
            while (v >= cend[j] - cbeg[j]) {
                v -= cend[j] - cbeg[j];
                ++j;
            }
            const int32_t s = cbeg[j] + v;
            sx[e] = xs[s];
            sy[e] = ys[s];
#pragma unroll
            for (int q = 0; q < RC; ++q) st[e * RC + q] = tau[(size_t)s * R + r0 + q];
        }
        __syncthreads();
        if (act) {
            const int tloc[T] = {};
            p2p_chunk<KIND, RC, T, false>(ker, logt, tx, ty, tloc, sx, sy, st, (g - v0 % G + G) % G, m, G, acc);
        }
    }
    // the groups' partial sums, target set u at a time
#pragma unroll
    for (int u = 0; u < T; ++u) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RC; ++q) red[lane * RC + q] = acc[u][q];
        __syncthreads();
        if (g == 0 && ok[u])
#pragma unroll
            for (int q = 0; q < RC; ++q) {
                double sum = 0.0;
                for (int gg = 0; gg < G; ++gg) sum += red[(gg * Lg + tl) * RC + q];
                near[(size_t)tg[u] * R + r0 + q] = ker.pref() * sum;
            }
    }
}

template <int KIND, int RC>
__global__ void __launch_bounds__(64) k_p2p(const int32_t* __restrict__ leaves, int n_leaves,
                                            const int32_t* __restrict__ leaf_coll,
                                            const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                            const double* __restrict__ xs, const double* __restrict__ ys,
                                            const double* __restrict__ tau, int lgR, KerT<KIND> ker,
                                            double* __restrict__ near) {
    constexpr int STAGE = p2p_stage<RC>();
    __shared__ double sx[STAGE], sy[STAGE], st[STAGE * RC], red[64 * RC];
    __shared__ int32_t cbeg[9], cend[9];
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    stage_logt<KIND>(logt);  // published by the barrier that opens each leaf
    pin_coef(ker);
    const int lane = threadIdx.x, R = 1 << lgR, r0 = blockIdx.y * RC;
    // grid-stride over the leaves: a grid smaller than the leaf count keeps fewer workgroups
    // resident (the near field then shares the GPU with the overlapped passes)
    for (int it = blockIdx.x; it < n_leaves; it += gridDim.x) {
        const int li = KF_VALU_RUN > 0 && (int)gridDim.x == n_leaves ? xcd_chunked(it, n_leaves, KF_VALU_RUN) : it;
        const int32_t b = leaves[li];
        const int32_t tb = bot[b], nt = top[b] - tb;
        __syncthreads();
        // the other colleagues' source ranges, as one virtual sequence [0, n_other)
        if (lane < 9) {
            const int32_t c = leaf_coll[(size_t)li * 9 + lane];
            const bool use = c >= 0 && c != b;
            cbeg[lane] = use ? bot[c] : 0;
            cend[lane] = use ? top[c] : 0;
        }
        __syncthreads();
        int n_other = 0;
#pragma unroll
        for (int j = 0; j < 9; ++j) n_other += cend[j] - cbeg[j];
        for (int t0 = 0; t0 < nt; t0 += 64) {
            const int cnt = min(64, nt - t0);
            if (p2p_use<2>(cnt))
                p2p_block<KIND, RC, 2, STAGE>(ker, logt, tb, nt, t0, cnt, n_other, cbeg, cend, xs, ys, tau, R, r0, sx,
                                              sy, st, red, near);
            else
                p2p_block<KIND, RC, 1, STAGE>(ker, logt, tb, nt, t0, cnt, n_other, cbeg, cend, xs, ys, tau, R, r0, sx,
                                              sy, st, red, near);
        }
    }
}

// Stored near field (the reference's planned near field, fmm.py:646-711 building and
// fmm.py:750-758 applying it, kept here as dense per-leaf blocks instead of CSR): leaf li's block
// at off[li] holds G for its nt targets against its ns = nt + n_other sources -- its own points,
// then the other colleagues' in list order, as in k_p2p -- column-major (source s, target t at
// s * nt + t), prefactor included, the self pairs zero.
__device__ __forceinline__ int32_t near_source(int s, int32_t tb, int nt, const int32_t* cbeg, const int32_t* cend) {
    if (s < nt) return tb + s;
    int v = s - nt, j = 0;
    while (v >= cend[j] - cbeg[j]) {
        v -= cend[j] - cbeg[j];
        ++j;
    }
    return cbeg[j] + v;
}

// The colleague source ranges of leaf li other than the leaf itself, into cbeg/cend (LDS, 9
// each); returns the leaf's box.  Called by every lane of the workgroup, between barriers.
__device__ __forceinline__ int32_t near_ranges(int li, const int32_t* __restrict__ leaves,
                                               const int32_t* __restrict__ leaf_coll, const int32_t* __restrict__ bot,
                                               const int32_t* __restrict__ top, int32_t* cbeg, int32_t* cend) {
    const int32_t b = leaves[li];
    __syncthreads();
    if (threadIdx.x < 9) {
        const int32_t c = leaf_coll[(size_t)li * 9 + threadIdx.x];
        const bool use = c >= 0 && c != b;
        cbeg[threadIdx.x] = use ? bot[c] : 0;
        cend[threadIdx.x] = use ? top[c] : 0;
    }
    __syncthreads();
    return b;
}

template <int KIND>
__global__ void __launch_bounds__(64) k_p2p_store(const int32_t* __restrict__ leaves, int n_leaves,
                                                  const int32_t* __restrict__ leaf_coll,
                                                  const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                                  const double* __restrict__ xs, const double* __restrict__ ys,
                                                  KerT<KIND> ker, const int64_t* __restrict__ off,
                                                  double* __restrict__ M) {
    __shared__ int32_t cbeg[9], cend[9];
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    stage_logt<KIND>(logt);  // published by the barriers of near_ranges
    for (int li = blockIdx.x; li < n_leaves; li += gridDim.x) {
        const int32_t b = near_ranges(li, leaves, leaf_coll, bot, top, cbeg, cend);
        const int32_t tb = bot[b];
        const int nt = top[b] - tb;
        int n_other = 0;
#pragma unroll
        for (int j = 0; j < 9; ++j) n_other += cend[j] - cbeg[j];
        const int64_t n = (int64_t)nt * (nt + n_other);
        double* Ml = M + off[li];
        for (int64_t e = threadIdx.x; e < n; e += 64) {
            const int s = (int)(e / nt), t = (int)(e - (int64_t)s * nt);
            const int32_t src = near_source(s, tb, nt, cbeg, cend), tg = tb + t;
            const double dx = xs[tg] - xs[src], dy = ys[tg] - ys[src];
            const bool me = src == tg;
            const double v = ker.pref() * ker.raw(me ? 1.0 : fma(dx, dx, dy * dy), logt);
            Ml[e] = me ? 0.0 : v;
        }
    }
}

// Applies the stored blocks: one 64-lane workgroup per (leaf, RHS chunk), a block of cnt <= 64
// targets one per lane, the 64 / cnt lane groups taking every G-th source (group g the sources
// s = g mod G), so a wave's loads of one step cover G consecutive columns; the groups' sums are
// added in group order.  G depends only on cnt: a density gives bitwise the same potential in
// any batch.
template <int RC>
__global__ void __launch_bounds__(64) k_p2p_stored(const int32_t* __restrict__ leaves, int n_leaves,
                                                   const int32_t* __restrict__ leaf_coll,
                                                   const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                                   const double* __restrict__ tau, int lgR,
                                                   const int64_t* __restrict__ off, const double* __restrict__ M,
                                                   double* __restrict__ near) {
    constexpr int STAGE = p2p_stage<RC>();
    __shared__ double st[STAGE * RC], red[64 * RC];
    __shared__ int32_t cbeg[9], cend[9];
    const int lane = threadIdx.x, R = 1 << lgR, r0 = blockIdx.y * RC;
    for (int li = blockIdx.x; li < n_leaves; li += gridDim.x) {
        const int32_t b = near_ranges(li, leaves, leaf_coll, bot, top, cbeg, cend);
        const int32_t tb = bot[b];
        const int nt = top[b] - tb;
        int n_other = 0;
#pragma unroll
        for (int j = 0; j < 9; ++j) n_other += cend[j] - cbeg[j];
        const int ns = nt + n_other;
        const double* Ml = M + off[li];
        for (int t0 = 0; t0 < nt; t0 += 64) {
            const int cnt = min(64, nt - t0);
            const int G = 64 / cnt, g = lane / cnt, tl = lane - g * cnt;
            const bool act = g < G;
            double acc[RC];
#pragma unroll
            for (int q = 0; q < RC; ++q) acc[q] = 0.0;
            for (int s0 = 0; s0 < ns; s0 += STAGE) {
                const int m = min(STAGE, ns - s0);
                __syncthreads();
                // every lane's loads of the stage are issued before the first store
                constexpr int PER = (STAGE + 63) / 64;
                double tv[PER][RC];
#pragma unroll
                for (int i = 0; i < PER; ++i) {
                    const int e = lane + 64 * i;
                    const int32_t src = e < m ? near_source(s0 + e, tb, nt, cbeg, cend) : tb;
#pragma unroll
                    for (int q = 0; q < RC; ++q) tv[i][q] = tau[(size_t)src * R + r0 + q];
                }
#pragma unroll
                for (int i = 0; i < PER; ++i)
                    if (lane + 64 * i < m)
#pragma unroll
                        for (int q = 0; q < RC; ++q) st[(lane + 64 * i) * RC + q] = tv[i][q];
                __syncthreads();
                if (act) {
                    // a lane reads one column element per source: many loads in flight carry
                    // the bandwidth (a leaf of more than 32 targets has one group)
                    const double* col = Ml + (size_t)s0 * nt + t0 + tl;
#pragma unroll 16
                    for (int k = (g - s0 % G + G) % G; k < m; k += G) {
                        const double v = col[(size_t)k * nt];
#pragma unroll
                        for (int q = 0; q < RC; ++q) acc[q] = fma(v, st[k * RC + q], acc[q]);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < RC; ++q) red[lane * RC + q] = acc[q];
            __syncthreads();
            if (g == 0)
#pragma unroll
                for (int q = 0; q < RC; ++q) {
                    double sum = 0.0;
                    for (int gg = 0; gg < G; ++gg) sum += red[(gg * cnt + tl) * RC + q];
                    near[(size_t)(tb + t0 + tl) * R + r0 + q] = sum;
                }
        }
    }
}

// One block of up to 64 targets of L2P, one or two per lane as in the near field: group g of
// ceil(cnt / T) lanes takes every G-th equivalent point.
template <int KIND, int RC, int T, bool ST>
__device__ __forceinline__ void l2p_block(const KerT<KIND>& ker, const double* logt, bool far, int32_t tb, int t0,
                                          int cnt, double bx, double by, const double* __restrict__ xs,
                                          const double* __restrict__ ys, const int32_t* __restrict__ perm,
                                          const double* lsx, const double* lsy, const double* psi, int p, double* red,
                                          double* __restrict__ near, int R, int r0, int nvalid, int rows,
                                          const double* __restrict__ Ml, int nt, const RhsOut& out) {
    const int lane = threadIdx.x;
    const int Lg = (cnt + T - 1) / T;
    const int G = 64 / Lg, g = lane / Lg, tl = lane - g * Lg;
    const bool act = g < G;
    double rx[T], ry[T];
    int32_t tg[T];
    bool ok[T];
#pragma unroll
    for (int u = 0; u < T; ++u) {
        const int i = tl + u * Lg;
        ok[u] = act && i < cnt;
        tg[u] = tb + t0 + (ok[u] ? i : tl);  // a missing second target repeats the first
        rx[u] = act ? xs[tg[u]] - bx : 0.0;
        ry[u] = act ? ys[tg[u]] - by : 0.0;
    }
    __syncthreads();
    double acc[T][RC];
#pragma unroll
    for (int u = 0; u < T; ++u)
#pragma unroll
        for (int q = 0; q < RC; ++q) acc[u][q] = 0.0;
    // S equivalent points per step (with T targets per lane: S T kernel values side by side), each
    // accumulator taking the points in order; the stored form reads its values one point at a time
    constexpr int S = ST ? 1 : (4 / T > 1 ? 4 / T : 1);
    auto points = [&](int jj, auto ns) {
        constexpr int NS = decltype(ns)::value;
        double v[NS * T];
        if constexpr (ST) {
#pragma unroll
            for (int u = 0; u < T; ++u) v[u] = Ml[(size_t)jj * nt + (tg[u] - tb)];
        } else {
            double r2[NS * T];
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                const double ex = lsx[jj + i * G], ey = lsy[jj + i * G];
#pragma unroll
                for (int u = 0; u < T; ++u) {
                    const double dx = rx[u] - ex, dy = ry[u] - ey;
                    r2[i * T + u] = fma(dx, dx, dy * dy);
                }
            }
            raw_group(ker, r2, logt, v);
        }
#pragma unroll
        for (int i = 0; i < NS; ++i)
#pragma unroll
            for (int q = 0; q < RC; ++q) {
                const double w = psi[(jj + i * G) * RC + q];
#pragma unroll
                for (int u = 0; u < T; ++u) acc[u][q] = fma(v[i * T + u], w, acc[u][q]);
            }
    };
    if (far && act) {
        int jj = g;
        for (; jj + (S - 1) * G < p; jj += S * G) points(jj, std::integral_constant<int, S>{});
        for (; jj < p; jj += G) points(jj, std::integral_constant<int, 1>{});
    }
#pragma unroll
    for (int u = 0; u < T; ++u) {
        if (u > 0) __syncthreads();
#pragma unroll
        for (int q = 0; q < RC; ++q) red[lane * RC + q] = acc[u][q];
        __syncthreads();
        if (g == 0 && ok[u]) {
            const int32_t o = perm[tg[u]];
#pragma unroll
            for (int q = 0; q < RC; ++q) {
                if (r0 + q >= nvalid) break;
                double sum = 0.0;
                for (int gg = 0; gg < G; ++gg) sum += red[(gg * Lg + tl) * RC + q];
                const size_t nr = (size_t)tg[u] * R + r0 + q;
                const double pot = ker.pref() * sum + near[nr];
                if (rows)
                    near[nr] = pot;  // the row form: k_scatter_pot writes the input order
                else
                    out.p[r0 + q][o] = pot;
            }
        }
    }
}

// L2P (reference fmm.py:91-115): far field of every real leaf from its local equivalent
// density Psi (levels >= 2; level-1 expansions are zero), plus the near field, written in input
// order for the nvalid real right-hand sides (rows != 0: left in `near`, tree order, for
// k_scatter_pot).  Lane groups split the p terms for small leaves.
// ST: the kernel values are read from the stored matrix (k_l2p_store: leaf li at loff[li],
// equivalent point jj's nt values at jj * nt), bitwise the values computed here otherwise.
template <int KIND, int RC, bool ST>
__global__ void __launch_bounds__(64) k_l2p(const int32_t* __restrict__ leaves, const int32_t* __restrict__ rid,
                                            const int32_t* __restrict__ lvl, const int32_t* __restrict__ bop,
                                            const double* __restrict__ cx, const double* __restrict__ cy,
                                            const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                            const double* __restrict__ xs, const double* __restrict__ ys,
                                            const int32_t* __restrict__ perm, const double* __restrict__ surf,
                                            const double* __restrict__ Psi, double* __restrict__ near,
                                            int p, int P, int lgR, int nvalid, int rows, KerT<KIND> ker,
                                            const int64_t* __restrict__ loff, const double* __restrict__ M,
                                            RhsOut out) {
    __shared__ double psi[64 * RC], lsx[64], lsy[64], red[64 * RC];
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    stage_logt<KIND>(logt);  // published with the surface points below
    pin_coef(ker);
    const int lane = threadIdx.x, R = 1 << lgR, r0 = blockIdx.y * RC;
    const int li = KF_VALU_RUN > 0 ? xcd_chunked(blockIdx.x, gridDim.x, KF_VALU_RUN) : blockIdx.x;
    const int32_t b = leaves[li];
    const bool far = lvl[b] >= 2;
    const int ob = bop[b];
    const double bx = cx[b], by = cy[b];
    if (lane < p) {
        lsx[lane] = surf[(size_t)(2 * ob) * P + lane];
        lsy[lane] = surf[(size_t)(2 * ob + 1) * P + lane];
    }
    for (int e = lane; e < p * RC; e += 64) {
        int jj = e / RC, q = e % RC;
        psi[e] = far ? Psi[((size_t)rid[b] * R + r0 + q) * P + jj] : 0.0;
    }
    const int32_t tb = bot[b], nt = top[b] - tb;
    const double* Ml = ST ? M + loff[li] : nullptr;
    for (int t0 = 0; t0 < nt; t0 += 64) {
        const int cnt = min(64, nt - t0);
        if (p2p_use<2>(cnt))
            l2p_block<KIND, RC, 2, ST>(ker, logt, far, tb, t0, cnt, bx, by, xs, ys, perm, lsx, lsy, psi, p, red, near,
                                       R, r0, nvalid, rows, Ml, nt, out);
        else
            l2p_block<KIND, RC, 1, ST>(ker, logt, far, tb, t0, cnt, bx, by, xs, ys, perm, lsx, lsy, psi, p, red, near,
                                       R, r0, nvalid, rows, Ml, nt, out);
    }
}

// Stored L2P values (the reference planner's stored downward leaf interactions, fmm.py:646-711):
// leaf li at loff[li] (nothing for leaves above level 2, which have no far field), equivalent
// point jj's values for the leaf's nt targets at jj * nt, the arithmetic of l2p_block.
template <int KIND>
__global__ void __launch_bounds__(64) k_l2p_store(const int32_t* __restrict__ leaves, int n,
                                                  const int32_t* __restrict__ lvl, const int32_t* __restrict__ bop,
                                                  const double* __restrict__ cx, const double* __restrict__ cy,
                                                  const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                                  const double* __restrict__ xs, const double* __restrict__ ys,
                                                  const double* __restrict__ surf, int p, int P, KerT<KIND> ker,
                                                  const int64_t* __restrict__ loff, double* __restrict__ M) {
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    stage_logt<KIND>(logt);
    __syncthreads();
    for (int li = blockIdx.x; li < n; li += gridDim.x) {
        const int32_t b = leaves[li];
        if (lvl[b] < 2) continue;
        const int ob = bop[b];
        const double bx = cx[b], by = cy[b];
        const int32_t tb = bot[b];
        const int nt = top[b] - tb;
        for (int64_t e = threadIdx.x; e < (int64_t)nt * p; e += 64) {
            const int jj = (int)(e / nt), i = (int)(e - (int64_t)jj * nt);
            const double rx = xs[tb + i] - bx, ry = ys[tb + i] - by;
            const double dx = rx - surf[(size_t)(2 * ob) * P + jj], dy = ry - surf[(size_t)(2 * ob + 1) * P + jj];
            M[loff[li] + e] = ker.raw(fma(dx, dx, dy * dy), logt);
        }
    }
}

// Target locate (reference tree.py:712-727): descend by strict ">" against box centres.
__global__ void k_locate(const double* __restrict__ tx, const double* __restrict__ ty, int64_t m,
                         const double* __restrict__ cx, const double* __restrict__ cy,
                         const int32_t* __restrict__ child0, const uint8_t* __restrict__ leafflag,
                         int32_t* __restrict__ out) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    double x = tx[i], y = ty[i];
    int32_t b = 0;
    while (!leafflag[b]) b = child0[b] + 2 * (x > cx[b]) + (y > cy[b]);
    out[i] = b;
}

// Target evaluation (reference class_fmm.py:73-111) for right-hand side 0 of the last
// expansion: L2P from the located leaf plus the colleague direct sum; check_self skips
// sources that coincide with the target.  Right-hand side r = blockIdx.y of the last
// expansion goes to pot[r*m + i].
template <int KIND>
__global__ void __launch_bounds__(256) k_targets(const double* __restrict__ tx, const double* __restrict__ ty,
                                                 int64_t m, const int32_t* __restrict__ loc,
                                                 const int32_t* __restrict__ rid, const int32_t* __restrict__ lvl,
                                                 const int32_t* __restrict__ bop, const int32_t* __restrict__ coll,
                                                 const double* __restrict__ cx, const double* __restrict__ cy,
                                                 const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                                 const double* __restrict__ xs, const double* __restrict__ ys,
                                                 const double* __restrict__ tau, const double* __restrict__ surf,
                                                 const double* __restrict__ Psi, int p, int P, int lgR,
                                                 KerT<KIND> ker, int check, double* __restrict__ pot) {
    __shared__ __attribute__((aligned(16))) double logt[logt_lds<KIND>()];
    stage_logt<KIND>(logt);
    pin_coef(ker);
    __syncthreads();
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int r = blockIdx.y;
    const double x = tx[i], y = ty[i];
    const int32_t b = loc[i];
    const int l = bop[b];
    double acc = 0.0;
    if (lvl[b] >= 2) {
        const double rx = x - cx[b], ry = y - cy[b];
        const double* ps = Psi + (((size_t)rid[b] << lgR) + r) * P;
        for (int j = 0; j < p; ++j) {
            double dx = rx - surf[(size_t)(2 * l) * P + j], dy = ry - surf[(size_t)(2 * l + 1) * P + j];
            acc += ker.g(fma(dx, dx, dy * dy), logt) * ps[j];
        }
    }
    double nearv = 0.0;
    for (int slot = 0; slot < 9; ++slot) {
        const int32_t c = coll[(size_t)b * 9 + slot];
        if (c < 0) continue;
        for (int32_t s = bot[c]; s < top[c]; ++s) {
            double dx = x - xs[s], dy = y - ys[s];
            if (check && dx == 0.0 && dy == 0.0) continue;
            nearv += ker.g(fma(dx, dx, dy * dy), logt) * tau[((size_t)s << lgR) + r];
        }
    }
    pot[(size_t)r * m + i] = acc + nearv;
}

}  // namespace kf
