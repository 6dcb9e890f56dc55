// plan.hip -- planned KIFMM evaluation on one MI355X: work lists, HBM slabs and the passes.
//
// Data layout in HBM (one plan):
//   points    xs, ys (tree order, owned by the tree), tau (tree order, per evaluation)
//   boxes     cx, cy, bot, top, level, child0, leaf flag        (global box id, all levels)
//   slabs     U, X, Psi: [boxes][P] FP64, P = p rounded up to a multiple of 16 (zero padded)
//             U   upward check potential of leaf boxes (P2M)
//             X   upward equivalent density            (reference Equiv_Densities)
//             Psi downward equivalent density          (reference Local_Expansions)
//   operators per level, see ops_host.h
// Passes (reference class_fmm.py:299-352, fmm.py:350-408):
//   k_gather_tau -> k_p2m (all levels) -> k_up per level, bottom-up (M2M + C2E)
//   -> k_m2l (all levels, 27-slot M2L into D) -> k_l2l_c2e per level, top-down (L2L + C2E)
//   -> k_l2p (far field + near field, input order).
// The near field k_p2p depends only on the gathered densities; in kf_plan_eval it runs on a
// second stream of the plan, overlapped with the level-by-level passes, and joins before L2P.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "kf_internal.h"
#include "kf_mfma.h"
#include "ops_host.h"

using namespace kf;

struct kf_plan {
    const kf_tree* tree = nullptr;
    KfKernel ker{};
    int p = 0, P = 0, NB = 0, prec = 0, L = 0;
    int64_t N = 0, B = 0;
    DevBuf<double> cx, cy, U, X, Psi, tau, ops, surf;
    DevBuf<int32_t> bot, top, lvl, child0, coll;
    DevBuf<uint8_t> leafflag;
    DevBuf<int32_t> p2m, up, dn_tgt, dn_par, dn_vsrc, leaves, leaf_coll, loc;
    DevBuf<int4> dn_tiles;
    DevBuf<uint32_t> dn_masks;
    int n_tiles = 0;
    DevBuf<int16_t> slot_op;
    int n_p2m = 0, n_leaves = 0;
    std::vector<int> up_off, up_n, up_nm2m, tile_off, tile_n;
    kf_plan_stats stats{};
    int64_t aliased_pairs = 0;
    bool expanded = false;
    bool near_pending = false;         // near field of the last expansion in flight on `side`
    DevBuf<double> nearb;              // near-field potentials, tree order
    hipStream_t side = nullptr;        // stream for the overlapped near field
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // per-pass profiling: one set of events per evaluation in a ring; totals are summed when
    // a set is reused or when the caller reads them.  Events 0-5 bracket the passes on the
    // evaluation stream, 6-7 the near field.
    static constexpr int RING = 64, NEV = 8, NPASS = 6;
    bool profile = false;
    int cur = -1;
    int64_t n_rec = 0, n_done = 0, n_acc = 0;
    double prof_ms[NPASS] = {0, 0, 0, 0, 0, 0};
    hipEvent_t ev[RING][NEV] = {};
    ~kf_plan() {
        for (auto& set : ev)
            for (auto& e : set)
                if (e) (void)hipEventDestroy(e);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (side) (void)hipStreamDestroy(side);
    }
};

namespace kf {

__global__ void k_gather_tau(const double* __restrict__ tau, const int32_t* __restrict__ perm,
                             int64_t n, double* __restrict__ out) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = tau[perm[i]];
}

// P2M (reference class_fmm.py:57-71): one wave per box, lane j = check point j on the box's
// large circle; sources staged through LDS 64 at a time.
template <int KIND>
__global__ void __launch_bounds__(64) k_p2m(const int32_t* __restrict__ list, const int32_t* __restrict__ lvl,
                                            const double* __restrict__ cx, const double* __restrict__ cy,
                                            const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                            const double* __restrict__ xs, const double* __restrict__ ys,
                                            const double* __restrict__ tau, const double* __restrict__ surf,
                                            int p, int P, KerT<KIND> ker, double* __restrict__ U) {
    __shared__ double sx[64], sy[64], st[64];
    const int lane = threadIdx.x;
    const int32_t b = list[blockIdx.x];
    const int l = lvl[b];
    const bool on = lane < p;
    const double tx = on ? cx[b] + surf[(size_t)(2 * l) * P + lane] : 0.0;
    const double ty = on ? cy[b] + surf[(size_t)(2 * l + 1) * P + lane] : 0.0;
    double acc = 0.0;
    const int32_t b0 = bot[b], b1 = top[b];
    for (int32_t c = b0; c < b1; c += 64) {
        int m = min(64, b1 - c);
        __syncthreads();
        if (lane < m) {
            sx[lane] = xs[c + lane];
            sy[lane] = ys[c + lane];
            st[lane] = tau[c + lane];
        }
        __syncthreads();
        if (on)
            for (int k = 0; k < m; ++k) {
                double dx = tx - sx[k], dy = ty - sy[k];
                acc += ker.raw(fma(dx, dx, dy * dy)) * st[k];
            }
    }
    if (lane < P) U[(size_t)b * P + lane] = on ? ker.pref() * acc : 0.0;
}

// Upward pass of one level (reference class_fmm.py:309-319): U = M2M(children X) for boxes
// with children, U = P2M result for the others, then X = E^{-1} U = F diag(1/lambda) F^T U.
// Tiles of 64 boxes (16 per wave); boxes with children come first in the list.  The next
// operator is fetched into registers while the current one is applied.
template <int NB>
__global__ void __launch_bounds__(256) k_up(const int32_t* __restrict__ list, int n, int n_m2m,
                                            const int32_t* __restrict__ child0,
                                            const double* __restrict__ ops,
                                            const double* __restrict__ Ug, double* __restrict__ X) {
    constexpr int P = 16 * NB;
    __shared__ __attribute__((aligned(32))) double Bs[P * P];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tile0 = blockIdx.x * 64;
    const int i = tile0 + 16 * wave + (lane & 15);
    const int32_t b = i < n ? list[i] : -1;
    const bool m2m = i < n_m2m;
    const bool any_m2m = tile0 < n_m2m;
    OpStage<NB, double> st;
    d4 acc[NB], R[NB];
    rows_zero(acc);
    st.load(ops + (any_m2m ? LevelOps::m2mb(P, 0) : LevelOps::F(P)));
    if (any_m2m) {
        for (int q = 0; q < 4; ++q) {
            __syncthreads();
            st.store(Bs);
            __syncthreads();
            st.load(ops + (q < 3 ? LevelOps::m2mb(P, q + 1) : LevelOps::F(P)));
            rows_load(R, X, m2m ? (long)child0[b] + q : -1L);
            if (__ballot(m2m)) rows_gemm(acc, R, Bs, P);
        }
    }
    if (!m2m && b >= 0) rows_add(acc, Ug, (long)b);
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
    rows_store(acc, X, (long)b);
}

// M2L of every level in one launch (reference class_fmm.py:336-343 with the collected
// operators split back into the 40 offsets): D = sum over the interaction-list slots of
// M2L_o X_source.  A tile holds up to 64 targets of one level with the same child position q,
// so a slot means the same operator for all of them; `mask` lists the slots any target uses.
// Operator s+1 and the source rows of slot s+1 are fetched while slot s is applied.
template <int NB>
__global__ void __launch_bounds__(256) k_m2l(const int4* __restrict__ tiles, const uint32_t* __restrict__ masks,
                                             const int32_t* __restrict__ tgt, const int32_t* __restrict__ vsrc,
                                             const int16_t* __restrict__ slot_op, const double* __restrict__ ops,
                                             size_t ops_stride, const double* __restrict__ X,
                                             double* __restrict__ D) {
    constexpr int P = 16 * NB;
    __shared__ __attribute__((aligned(32))) double Bs[P * P];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int4 tile = tiles[blockIdx.x];
    const uint32_t mask = masks[blockIdx.x];
    const int q = tile.z;
    const double* opl = ops + (size_t)tile.w * ops_stride;
    const int i = 16 * wave + (lane & 15);
    const bool valid = i < tile.y;
    const int t = valid ? tile.x + i : -1;
    const int32_t b = valid ? tgt[t] : -1;
    OpStage<NB, double> st;
    d4 acc[NB], R[NB], Rn[NB];
    rows_zero(acc);
    int s = mask ? __ffs(mask) - 1 : -1;
    int32_t src = -1;
    if (s >= 0) {
        st.load(opl + LevelOps::m2lt(P, slot_op[q * 27 + s]));
        src = valid ? vsrc[(size_t)t * 27 + s] : -1;
        rows_load(R, X, (long)src);
    }
    while (s >= 0) {
        __syncthreads();
        st.store(Bs);
        __syncthreads();
        const uint32_t rest = mask & ~((2u << s) - 1u);
        const int sn = rest ? __ffs(rest) - 1 : -1;
        int32_t srcn = -1;
        if (sn >= 0) {
            st.load(opl + LevelOps::m2lt(P, slot_op[q * 27 + sn]));
            srcn = valid ? vsrc[(size_t)t * 27 + sn] : -1;
            rows_load(Rn, X, (long)srcn);
        }
        if (__ballot(src >= 0)) rows_gemm(acc, R, Bs, P);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) R[nb] = Rn[nb];
        src = srcn;
        s = sn;
    }
    rows_store(acc, D, (long)b);
}

// Downward pass of one level (reference class_fmm.py:328-334, 348-350): D += L2L(parent Psi)
// (levels >= 3; the level-1 expansions are zero), then Psi = E^{-1} D.
template <int NB>
__global__ void __launch_bounds__(256) k_l2l_c2e(const int4* __restrict__ tiles, const int32_t* __restrict__ tgt,
                                                 const int32_t* __restrict__ par, const double* __restrict__ ops,
                                                 int do_l2l, const double* __restrict__ D,
                                                 double* __restrict__ Psi) {
    constexpr int P = 16 * NB;
    __shared__ __attribute__((aligned(32))) double Bs[P * P];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int4 tile = tiles[blockIdx.x];
    const int i = 16 * wave + (lane & 15);
    const bool valid = i < tile.y;
    const int t = valid ? tile.x + i : -1;
    const int32_t b = valid ? tgt[t] : -1;
    OpStage<NB, double> st;
    d4 acc[NB], R[NB];
    rows_load(acc, D, (long)b);
    st.load(ops + (do_l2l ? LevelOps::l2lb(P, tile.z) : LevelOps::F(P)));
    if (do_l2l) {
        rows_load(R, Psi, valid ? (long)par[t] : -1L);
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
    rows_store(acc, Psi, (long)b);
}

// Near field (reference fmm.py:30-64): for every real leaf, the direct sum over the sources of
// its 9 colleagues with index self-exclusion, into near[] (tree order).  One 64-lane
// workgroup per leaf; with T <= 64 targets, 64/T lane groups split each 64-source chunk.
template <int KIND, bool SELF>
__device__ __forceinline__ void p2p_chunk(const KerT<KIND>& ker, double tx, double ty, int t_local,
                                          const double* sx, const double* sy, const double* st, int k0, int m,
                                          int G, double& acc, double& acc_e) {
    // SELF: the chunk is the target's own leaf; the pair k == t_local is skipped by mapping
    // its squared distance to 1 (log 1 = 0, so it adds nothing), without a branch
    for (int k = k0; k < m; k += G) {
        double dx = tx - sx[k], dy = ty - sy[k];
        double r2 = fma(dx, dx, dy * dy);
        double w = st[k];
        if (SELF) {
            bool me = (k == t_local);
            r2 = me ? 1.0 : r2;
            w = me ? 0.0 : w;
        }
        if constexpr (KIND == KF_LAPLACE) {
            double lm, e;
            kf_log_split(r2, lm, e);
            acc = fma(w, lm, acc);
            acc_e = fma(w, e, acc_e);
        } else {
            acc = fma(w, ker.raw(r2), acc);
        }
    }
}

template <int KIND>
__global__ void __launch_bounds__(64) k_p2p(const int32_t* __restrict__ leaves, const int32_t* __restrict__ leaf_coll,
                                            const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                            const double* __restrict__ xs, const double* __restrict__ ys,
                                            const double* __restrict__ tau, KerT<KIND> ker,
                                            double* __restrict__ near) {
    __shared__ double sx[64], sy[64], st[64], red[64];
    const int lane = threadIdx.x;
    const int32_t b = leaves[blockIdx.x];
    const int32_t tb = bot[b], nt = top[b] - tb;
    for (int t0 = 0; t0 < nt; t0 += 64) {
        const int cnt = min(64, nt - t0);
        int T = 1;
        while (T < cnt) T <<= 1;
        const int G = 64 / T, tl = lane & (T - 1), g = lane / T;
        const bool act = tl < cnt;
        const int32_t t = tb + t0 + tl;
        const double tx = act ? xs[t] : 0.0, ty = act ? ys[t] : 0.0;
        double acc = 0.0, acc_e = 0.0;
        for (int slot = 0; slot < 9; ++slot) {
            const int32_t c = leaf_coll[(size_t)blockIdx.x * 9 + slot];
            if (c < 0) continue;
            const int32_t c0 = bot[c], c1 = top[c];
            for (int32_t s0 = c0; s0 < c1; s0 += 64) {
                const int m = min(64, c1 - s0);
                __syncthreads();
                if (lane < m) {
                    sx[lane] = xs[s0 + lane];
                    sy[lane] = ys[s0 + lane];
                    st[lane] = tau[s0 + lane];
                }
                __syncthreads();
                if (!act) continue;
                // the target's own point lies in this chunk iff c == b and s0 <= t < s0 + 64
                if (c == b && t >= s0 && t < s0 + 64)
                    p2p_chunk<KIND, true>(ker, tx, ty, t - s0, sx, sy, st, g, m, G, acc, acc_e);
                else
                    p2p_chunk<KIND, false>(ker, tx, ty, -1, sx, sy, st, g, m, G, acc, acc_e);
            }
        }
        __syncthreads();
        red[lane] = fma(acc_e, KF_LN2, acc);
        __syncthreads();
        if (g == 0 && act) {
            double sum = 0.0;
            for (int gg = 0; gg < G; ++gg) sum += red[gg * T + tl];
            near[t] = ker.pref() * sum;
        }
    }
}

// L2P (reference fmm.py:91-115): far field of every real leaf from its local equivalent
// density Psi (levels >= 2; level-1 expansions are zero), plus the near field, written in
// input order.  Lane groups split the p terms for small leaves.
template <int KIND>
__global__ void __launch_bounds__(64) k_l2p(const int32_t* __restrict__ leaves, const int32_t* __restrict__ lvl,
                                            const double* __restrict__ cx, const double* __restrict__ cy,
                                            const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                            const double* __restrict__ xs, const double* __restrict__ ys,
                                            const int32_t* __restrict__ perm, const double* __restrict__ surf,
                                            const double* __restrict__ Psi, const double* __restrict__ near,
                                            int p, int P, KerT<KIND> ker, double* __restrict__ pot) {
    __shared__ double psi[64], lsx[64], lsy[64], red[64];
    const int lane = threadIdx.x;
    const int32_t b = leaves[blockIdx.x];
    const int l = lvl[b];
    const bool far = l >= 2;
    const double bx = cx[b], by = cy[b];
    if (lane < p) {
        psi[lane] = far ? Psi[(size_t)b * P + lane] : 0.0;
        lsx[lane] = surf[(size_t)(2 * l) * P + lane];
        lsy[lane] = surf[(size_t)(2 * l + 1) * P + lane];
    }
    const int32_t tb = bot[b], nt = top[b] - tb;
    for (int t0 = 0; t0 < nt; t0 += 64) {
        const int cnt = min(64, nt - t0);
        int T = 1;
        while (T < cnt) T <<= 1;
        const int G = 64 / T, tl = lane & (T - 1), g = lane / T;
        const bool act = tl < cnt;
        const int32_t t = tb + t0 + tl;
        const double rx = act ? xs[t] - bx : 0.0, ry = act ? ys[t] - by : 0.0;
        __syncthreads();
        double acc = 0.0;
        if (far && act)
            for (int j = g; j < p; j += G) {
                double dx = rx - lsx[j], dy = ry - lsy[j];
                acc += ker.raw(fma(dx, dx, dy * dy)) * psi[j];
            }
        red[lane] = acc;
        __syncthreads();
        if (g == 0 && act) {
            double sum = 0.0;
            for (int gg = 0; gg < G; ++gg) sum += red[gg * T + tl];
            pot[perm[t]] = ker.pref() * sum + near[t];
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

// Target evaluation (reference class_fmm.py:73-111): L2P from the located leaf plus the
// colleague direct sum; check_self skips sources that coincide with the target.
template <int KIND>
__global__ void __launch_bounds__(256) k_targets(const double* __restrict__ tx, const double* __restrict__ ty,
                                                 int64_t m, const int32_t* __restrict__ loc,
                                                 const int32_t* __restrict__ lvl, const int32_t* __restrict__ coll,
                                                 const double* __restrict__ cx, const double* __restrict__ cy,
                                                 const int32_t* __restrict__ bot, const int32_t* __restrict__ top,
                                                 const double* __restrict__ xs, const double* __restrict__ ys,
                                                 const double* __restrict__ tau, const double* __restrict__ surf,
                                                 const double* __restrict__ Psi, int p, int P, KerT<KIND> ker,
                                                 int check, double* __restrict__ pot) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double x = tx[i], y = ty[i];
    const int32_t b = loc[i];
    const int l = lvl[b];
    double acc = 0.0;
    if (l >= 2) {
        const double rx = x - cx[b], ry = y - cy[b];
        for (int j = 0; j < p; ++j) {
            double dx = rx - surf[(size_t)(2 * l) * P + j], dy = ry - surf[(size_t)(2 * l + 1) * P + j];
            acc += ker.g(fma(dx, dx, dy * dy)) * Psi[(size_t)b * P + j];
        }
    }
    double nearv = 0.0;
    for (int slot = 0; slot < 9; ++slot) {
        const int32_t c = coll[(size_t)b * 9 + slot];
        if (c < 0) continue;
        for (int32_t s = bot[c]; s < top[c]; ++s) {
            double dx = x - xs[s], dy = y - ys[s];
            if (check && dx == 0.0 && dy == 0.0) continue;
            nearv += ker.g(fma(dx, dx, dy * dy)) * tau[s];
        }
    }
    pot[i] = acc + nearv;
}

// ----------------------------------------------------------------- host side ------------

static void build_plan(kf_plan* pl, const kf_tree* tr, const kf_kernel_desc* kd, int p, int prec,
                       hipStream_t s) {
    KF_REQUIRE(tr && kd, KF_ERR_ARG, "null argument");
    const HostTree& h = tr->h;
    KF_REQUIRE(tr->perm.p, KF_ERR_STATE, "tree has no device points");
    KF_REQUIRE(h.nlevels() >= 2, KF_ERR_ARG, "plan needs a tree with at least two levels");
    KF_REQUIRE(p >= 4 && p <= 64, KF_ERR_ARG, "p (Nequiv) must be in [4, 64]");
    KF_REQUIRE(kd->kind == KF_LAPLACE || kd->kind == KF_MODHELM, KF_ERR_ARG, "unknown kernel kind");
    KF_REQUIRE(kd->kind != KF_MODHELM || kd->k > 0.0, KF_ERR_ARG, "modified Helmholtz needs k > 0");
    KF_REQUIRE(prec == KF_PREC_FP64, KF_ERR_ARG, "precision mode not available in this build");
    pl->tree = tr;
    pl->ker = KfKernel{kd->kind, kd->k, kd->scale};
    pl->p = p;
    pl->NB = (p + 15) / 16;
    pl->P = 16 * pl->NB;
    pl->prec = prec;
    pl->L = h.nlevels();
    pl->N = h.n;
    pl->B = (int64_t)h.lvl.size();
    const int P = pl->P, L = pl->L;
    const int64_t B = pl->B;
    // per-box arrays
    std::vector<double> cx(B), cy(B);
    std::vector<int32_t> bot(B), top(B), lv(B), ch(B), coll((size_t)B * 9, -1);
    std::vector<uint8_t> lf(B);
    for (int64_t b = 0; b < B; ++b) {
        cx[b] = h.xmid((int32_t)b);
        cy[b] = h.ymid((int32_t)b);
        bot[b] = (int32_t)h.bot[b];
        top[b] = (int32_t)h.top[b];
        lv[b] = h.lvl[b];
        ch[b] = h.child0[b];
        lf[b] = h.leaf[b];
        const auto& C = h.coll[h.lvl[b]];
        for (int j = 0; j < 9; ++j) coll[(size_t)b * 9 + j] = C[(size_t)h.pos[b] * 9 + j];
    }
    // circles and operators
    std::vector<double> surf((size_t)L * 2 * P, 0.0);
    std::vector<double> ops((size_t)L * LevelOps::size(P), 0.0);
    for (int l = 0; l < L; ++l) {
        std::vector<double> a, c, lx, ly;
        level_circles(h.width(l), p, a, c, lx, ly);
        for (int j = 0; j < p; ++j) {
            surf[(size_t)(2 * l) * P + j] = lx[j];
            surf[(size_t)(2 * l + 1) * P + j] = ly[j];
        }
        if (l >= 1)
            build_level_ops(pl->ker, h.root_width, l, p, P, l >= 2, l <= L - 2,
                            ops.data() + (size_t)l * LevelOps::size(P));
    }
    // upward lists: per level, boxes with sources; boxes with children first
    std::vector<int32_t> p2m, up;
    pl->up_off.assign(L, 0);
    pl->up_n.assign(L, 0);
    pl->up_nm2m.assign(L, 0);
    kf_plan_stats& st = pl->stats;
    st = kf_plan_stats{};
    st.p = p;
    st.p_padded = P;
    for (int l = 1; l < L; ++l) {
        pl->up_off[l] = (int)up.size();
        for (int32_t b : h.levels[l])
            if (h.ns(b) > 0 && h.child0[b] >= 0) up.push_back(b);
        pl->up_nm2m[l] = (int)up.size() - pl->up_off[l];
        for (int32_t b : h.levels[l])
            if (h.ns(b) > 0 && h.child0[b] < 0) {
                up.push_back(b);
                p2m.push_back(b);
                st.p2m_evals += (int64_t)p * h.ns(b);
            }
        pl->up_n[l] = (int)up.size() - pl->up_off[l];
        st.m2m_boxes += pl->up_nm2m[l];
        st.c2e_up_rows += pl->up_n[l];
    }
    // downward lists: every real box of levels >= 2, grouped by child position
    const VSlots& vs = vslots();
    std::vector<int32_t> tg, pa, vsrc;
    std::vector<int4> tiles;
    std::vector<uint32_t> masks;
    pl->tile_off.assign(L, 0);
    pl->tile_n.assign(L, 0);
    for (int l = 2; l < L; ++l) {
        pl->tile_off[l] = (int)tiles.size();
        const int64_t side = (int64_t)1 << l;
        for (int q = 0; q < 4; ++q) {
            int start = (int)tg.size();
            for (int32_t t : h.levels[l]) {
                if (h.fake[t] || 2 * (h.ix[t] & 1) + (h.iy[t] & 1) != q) continue;
                int32_t par = h.parent[t];
                tg.push_back(t);
                pa.push_back(par);
                for (int sl = 0; sl < 27; ++sl) {
                    int64_t gx = h.ix[t] + vs.ox[q][sl], gy = h.iy[t] + vs.oy[q][sl];
                    int32_t src = -1;
                    if (gx >= 0 && gy >= 0 && gx < side && gy < side) {
                        const auto& C = h.coll[l - 1];
                        for (int j = 0; j < 9; ++j) {
                            int32_t c = C[(size_t)h.pos[par] * 9 + j];
                            if (c < 0 || h.ix[c] != (gx >> 1) || h.iy[c] != (gy >> 1)) continue;
                            if (h.child0[c] < 0) {
                                ++pl->aliased_pairs;
                                break;
                            }
                            int32_t cand = h.child0[c] + (int32_t)(2 * (gx & 1) + (gy & 1));
                            if (h.ns(cand) > 0) src = cand;
                            break;
                        }
                    }
                    vsrc.push_back(src);
                    st.m2l_pairs += src >= 0;
                }
                if (l >= 3) st.l2l_boxes += 1;
            }
            int cnt = (int)tg.size() - start;
            for (int o = 0; o < cnt; o += 64) {
                int m = std::min(64, cnt - o);
                tiles.push_back(make_int4(start + o, m, q, l));
                uint32_t mk = 0;
                for (int k = 0; k < m; ++k)
                    for (int sl = 0; sl < 27; ++sl)
                        if (vsrc[(size_t)(start + o + k) * 27 + sl] >= 0) mk |= 1u << sl;
                masks.push_back(mk);
            }
        }
        pl->tile_n[l] = (int)tiles.size() - pl->tile_off[l];
    }
    st.c2e_down_rows += (int64_t)tg.size();
    // leaves: real leaves with points, all levels >= 1
    std::vector<int32_t> leaves, lcoll;
    for (int l = 1; l < L; ++l)
        for (int32_t b : h.levels[l]) {
            if (!h.leaf[b] || h.ns(b) <= 0) continue;
            leaves.push_back(b);
            int64_t src = 0;
            for (int j = 0; j < 9; ++j) {
                int32_t c = h.coll[l][(size_t)h.pos[b] * 9 + j];
                lcoll.push_back(c);
                if (c >= 0) src += h.ns(c);
            }
            st.p2p_pairs += h.ns(b) * src - h.ns(b);
            if (l >= 2) st.l2p_evals += (int64_t)p * h.ns(b);
            st.leaves += 1;
        }
    std::vector<int16_t> sop(4 * 27);
    for (int q = 0; q < 4; ++q)
        for (int sl = 0; sl < 27; ++sl) sop[q * 27 + sl] = (int16_t)vs.op[q][sl];
    // upload
    pl->cx.upload(cx, s);
    pl->cy.upload(cy, s);
    pl->bot.upload(bot, s);
    pl->top.upload(top, s);
    pl->lvl.upload(lv, s);
    pl->child0.upload(ch, s);
    pl->leafflag.upload(lf, s);
    pl->coll.upload(coll, s);
    pl->surf.upload(surf, s);
    pl->ops.upload(ops, s);
    pl->p2m.upload(p2m, s);
    pl->n_p2m = (int)p2m.size();
    pl->up.upload(up, s);
    pl->dn_tgt.upload(tg, s);
    pl->dn_par.upload(pa, s);
    pl->dn_vsrc.upload(vsrc, s);
    pl->dn_tiles.upload(tiles, s);
    pl->dn_masks.upload(masks, s);
    pl->n_tiles = (int)tiles.size();
    pl->slot_op.upload(sop, s);
    pl->leaves.upload(leaves, s);
    pl->leaf_coll.upload(lcoll, s);
    pl->n_leaves = (int)leaves.size();
    pl->U.alloc((size_t)B * P);
    pl->X.alloc((size_t)B * P);
    pl->Psi.alloc((size_t)B * P);
    pl->tau.alloc((size_t)pl->N);
    pl->nearb.alloc((size_t)pl->N);
    pl->U.zero(s);
    pl->X.zero(s);
    pl->Psi.zero(s);
    KF_HIP_CHECK(hipStreamCreateWithFlags(&pl->side, hipStreamNonBlocking));
    KF_HIP_CHECK(hipEventCreateWithFlags(&pl->ev_fork, hipEventDisableTiming));
    KF_HIP_CHECK(hipEventCreateWithFlags(&pl->ev_join, hipEventDisableTiming));
    KF_HIP_CHECK(hipStreamSynchronize(s));
}

template <int NB>
static void launch_up(kf_plan* pl, int l, hipStream_t s) {
    int n = pl->up_n[l];
    if (!n) return;
    k_up<NB><<<(n + 63) / 64, 256, 0, s>>>(pl->up.p + pl->up_off[l], n, pl->up_nm2m[l], pl->child0.p,
                                           pl->ops.p + (size_t)l * LevelOps::size(pl->P), pl->U.p, pl->X.p);
}

template <int NB>
static void launch_m2l(kf_plan* pl, hipStream_t s) {
    if (!pl->n_tiles) return;
    k_m2l<NB><<<pl->n_tiles, 256, 0, s>>>(pl->dn_tiles.p, pl->dn_masks.p, pl->dn_tgt.p, pl->dn_vsrc.p,
                                          pl->slot_op.p, pl->ops.p, LevelOps::size(pl->P), pl->X.p, pl->U.p);
}

template <int NB>
static void launch_down(kf_plan* pl, int l, hipStream_t s) {
    int n = pl->tile_n[l];
    if (!n) return;
    k_l2l_c2e<NB><<<n, 256, 0, s>>>(pl->dn_tiles.p + pl->tile_off[l], pl->dn_tgt.p, pl->dn_par.p,
                                    pl->ops.p + (size_t)l * LevelOps::size(pl->P), l >= 3 ? 1 : 0, pl->U.p,
                                    pl->Psi.p);
}

static void accumulate(kf_plan* pl, int set) {
    // a set whose evaluation did not reach the self pass (expand only) contributes nothing
    if (hipEventSynchronize(pl->ev[set][5]) != hipSuccess) return;
    for (int i = 0; i < 5; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pl->ev[set][i], pl->ev[set][i + 1]) == hipSuccess) pl->prof_ms[i] += ms;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pl->ev[set][6], pl->ev[set][7]) == hipSuccess) pl->prof_ms[5] += ms;
    pl->n_acc += 1;
}

static void begin_profile(kf_plan* pl) {
    if (!pl->profile) return;
    if (pl->n_rec - pl->n_done >= kf_plan::RING) accumulate(pl, (int)(pl->n_done++ % kf_plan::RING));
    pl->cur = (int)(pl->n_rec % kf_plan::RING);
    pl->n_rec += 1;
}

static void mark(kf_plan* pl, int i, hipStream_t s) {
    if (pl->profile && pl->cur >= 0) KF_HIP_CHECK(hipEventRecord(pl->ev[pl->cur][i], s));
}

template <int KIND>
static void launch_p2m(kf_plan* pl, hipStream_t s) {
    const kf_tree* tr = pl->tree;
    if (pl->n_p2m)
        k_p2m<KIND><<<pl->n_p2m, 64, 0, s>>>(pl->p2m.p, pl->lvl.p, pl->cx.p, pl->cy.p, pl->bot.p, pl->top.p,
                                             tr->xs.p, tr->ys.p, pl->tau.p, pl->surf.p, pl->p, pl->P,
                                             KerT<KIND>{pl->ker.k, pl->ker.scale}, pl->U.p);
}

template <int KIND>
static void launch_p2p(kf_plan* pl, hipStream_t s) {
    const kf_tree* tr = pl->tree;
    if (pl->n_leaves)
        k_p2p<KIND><<<pl->n_leaves, 64, 0, s>>>(pl->leaves.p, pl->leaf_coll.p, pl->bot.p, pl->top.p, tr->xs.p,
                                                tr->ys.p, pl->tau.p, KerT<KIND>{pl->ker.k, pl->ker.scale},
                                                pl->nearb.p);
}

template <int KIND>
static void launch_l2p(kf_plan* pl, double* d_pot, hipStream_t s) {
    const kf_tree* tr = pl->tree;
    if (pl->n_leaves)
        k_l2p<KIND><<<pl->n_leaves, 64, 0, s>>>(pl->leaves.p, pl->lvl.p, pl->cx.p, pl->cy.p, pl->bot.p, pl->top.p,
                                                tr->xs.p, tr->ys.p, tr->perm.p, pl->surf.p, pl->Psi.p, pl->nearb.p,
                                                pl->p, pl->P, KerT<KIND>{pl->ker.k, pl->ker.scale}, d_pot);
}

static void run_p2p(kf_plan* pl, hipStream_t s) {
    mark(pl, 6, s);
    if (pl->ker.kind == KF_LAPLACE) launch_p2p<KF_LAPLACE>(pl, s);
    else launch_p2p<KF_MODHELM>(pl, s);
    mark(pl, 7, s);
}

// Upward and downward passes for one density.  With overlap_near, the near field of the
// same density is launched on the plan's side stream right after the gather; self_one()
// joins it.
static void expand_one(kf_plan* pl, const double* d_tau, hipStream_t s, bool overlap_near) {
    const kf_tree* tr = pl->tree;
    int64_t N = pl->N;
    begin_profile(pl);
    mark(pl, 0, s);
    k_gather_tau<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(d_tau, tr->perm.p, N, pl->tau.p);
    mark(pl, 1, s);
    if (overlap_near) {
        KF_HIP_CHECK(hipEventRecord(pl->ev_fork, s));
        KF_HIP_CHECK(hipStreamWaitEvent(pl->side, pl->ev_fork, 0));
        run_p2p(pl, pl->side);
        KF_HIP_CHECK(hipEventRecord(pl->ev_join, pl->side));
    }
    if (pl->ker.kind == KF_LAPLACE) launch_p2m<KF_LAPLACE>(pl, s);
    else launch_p2m<KF_MODHELM>(pl, s);
    mark(pl, 2, s);
    for (int l = pl->L - 1; l >= 1; --l) {
        switch (pl->NB) {
            case 1: launch_up<1>(pl, l, s); break;
            case 2: launch_up<2>(pl, l, s); break;
            case 3: launch_up<3>(pl, l, s); break;
            default: launch_up<4>(pl, l, s); break;
        }
    }
    mark(pl, 3, s);
    // M2L of all levels writes the downward check potentials D into the U slab (U is free
    // once the upward pass is done); then L2L + C2E level by level
    switch (pl->NB) {
        case 1: launch_m2l<1>(pl, s); break;
        case 2: launch_m2l<2>(pl, s); break;
        case 3: launch_m2l<3>(pl, s); break;
        default: launch_m2l<4>(pl, s); break;
    }
    for (int l = 2; l < pl->L; ++l) {
        switch (pl->NB) {
            case 1: launch_down<1>(pl, l, s); break;
            case 2: launch_down<2>(pl, l, s); break;
            case 3: launch_down<3>(pl, l, s); break;
            default: launch_down<4>(pl, l, s); break;
        }
    }
    mark(pl, 4, s);
    KF_HIP_CHECK(hipGetLastError());
    pl->expanded = true;
    pl->near_pending = overlap_near;
}

// Far field + near field at the sources, input order.
static void self_one(kf_plan* pl, double* d_pot, hipStream_t s) {
    if (pl->near_pending) KF_HIP_CHECK(hipStreamWaitEvent(s, pl->ev_join, 0));
    else run_p2p(pl, s);
    pl->near_pending = false;
    if (pl->ker.kind == KF_LAPLACE) launch_l2p<KF_LAPLACE>(pl, d_pot, s);
    else launch_l2p<KF_MODHELM>(pl, d_pot, s);
    mark(pl, 5, s);
    KF_HIP_CHECK(hipGetLastError());
}

static void finish_profile(kf_plan* pl) {
    // the self pass ends the evaluation; events are read lazily (no synchronisation here)
    pl->cur = -1;
}

static void drain_profile(kf_plan* pl) {
    while (pl->n_done < pl->n_rec) accumulate(pl, (int)(pl->n_done++ % kf_plan::RING));
}

}  // namespace kf

extern "C" {

int kf_plan_create(const kf_tree* tree, const kf_kernel_desc* kernel, int32_t p, int32_t precision,
                   void* stream, kf_plan** out) {
    return guarded([&] {
        KF_REQUIRE(out, KF_ERR_ARG, "null output handle");
        *out = nullptr;
        kf_plan* pl = new kf_plan();
        try {
            build_plan(pl, tree, kernel, p, precision, (hipStream_t)stream);
        } catch (...) {
            delete pl;
            throw;
        }
        *out = pl;
    });
}

int kf_plan_stats_get(const kf_plan* pl, kf_plan_stats* st) {
    return guarded([&] {
        KF_REQUIRE(pl && st, KF_ERR_ARG, "null argument");
        *st = pl->stats;
    });
}

int kf_plan_expand(kf_plan* pl, const double* d_tau, int64_t ld_tau, int32_t nrhs, void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl && d_tau, KF_ERR_ARG, "null argument");
        KF_REQUIRE(nrhs == 1, KF_ERR_ARG, "kf_plan_expand keeps one right-hand side; use kf_plan_eval");
        (void)ld_tau;
        expand_one(pl, d_tau, (hipStream_t)stream, false);
    });
}

int kf_plan_self(kf_plan* pl, double* d_pot, int64_t ld_pot, void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl && d_pot, KF_ERR_ARG, "null argument");
        KF_REQUIRE(pl->expanded, KF_ERR_STATE, "kf_plan_expand must run first");
        (void)ld_pot;
        self_one(pl, d_pot, (hipStream_t)stream);
        finish_profile(pl);
    });
}

int kf_plan_eval(kf_plan* pl, const double* d_tau, int64_t ld_tau, int32_t nrhs, double* d_pot,
                 int64_t ld_pot, void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl && d_tau && d_pot, KF_ERR_ARG, "null argument");
        KF_REQUIRE(nrhs >= 1, KF_ERR_ARG, "nrhs must be >= 1");
        KF_REQUIRE(nrhs == 1 || (ld_tau >= pl->N && ld_pot >= pl->N), KF_ERR_ARG, "leading dimension < n");
        hipStream_t s = (hipStream_t)stream;
        for (int r = 0; r < nrhs; ++r) {
            expand_one(pl, d_tau + (size_t)r * ld_tau, s, true);
            self_one(pl, d_pot + (size_t)r * ld_pot, s);
        }
        finish_profile(pl);
    });
}

int kf_plan_targets(kf_plan* pl, const double* d_tx, const double* d_ty, int64_t m, int32_t check_self,
                    double* d_pot, void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl && d_tx && d_ty && d_pot, KF_ERR_ARG, "null argument");
        KF_REQUIRE(pl->expanded, KF_ERR_STATE, "kf_plan_expand must run first");
        KF_REQUIRE(m >= 0, KF_ERR_ARG, "negative target count");
        if (m == 0) return;
        hipStream_t s = (hipStream_t)stream;
        const kf_tree* tr = pl->tree;
        if ((int64_t)pl->loc.n < m) pl->loc.alloc((size_t)m);
        unsigned nb = (unsigned)((m + 255) / 256);
        k_locate<<<nb, 256, 0, s>>>(d_tx, d_ty, m, pl->cx.p, pl->cy.p, pl->child0.p, pl->leafflag.p, pl->loc.p);
        if (pl->ker.kind == KF_LAPLACE)
            k_targets<KF_LAPLACE><<<nb, 256, 0, s>>>(d_tx, d_ty, m, pl->loc.p, pl->lvl.p, pl->coll.p, pl->cx.p,
                                                     pl->cy.p, pl->bot.p, pl->top.p, tr->xs.p, tr->ys.p, pl->tau.p,
                                                     pl->surf.p, pl->Psi.p, pl->p, pl->P,
                                                     KerT<KF_LAPLACE>{pl->ker.k, pl->ker.scale},
                                                     check_self ? 1 : 0, d_pot);
        else
            k_targets<KF_MODHELM><<<nb, 256, 0, s>>>(d_tx, d_ty, m, pl->loc.p, pl->lvl.p, pl->coll.p, pl->cx.p,
                                                     pl->cy.p, pl->bot.p, pl->top.p, tr->xs.p, tr->ys.p, pl->tau.p,
                                                     pl->surf.p, pl->Psi.p, pl->p, pl->P,
                                                     KerT<KF_MODHELM>{pl->ker.k, pl->ker.scale},
                                                     check_self ? 1 : 0, d_pot);
        KF_HIP_CHECK(hipGetLastError());
    });
}

int kf_plan_profile(kf_plan* pl, int32_t enable, double* ms, int32_t cap) {
    int written = 0;
    int rc = guarded([&] {
        KF_REQUIRE(pl, KF_ERR_ARG, "null plan");
        if (pl->profile) drain_profile(pl);
        double out[kf_plan::NPASS + 1];
        for (int i = 0; i < kf_plan::NPASS; ++i) out[i] = pl->prof_ms[i];
        out[kf_plan::NPASS] = (double)pl->n_acc;
        for (int i = 0; ms && i < cap && i < kf_plan::NPASS + 1; ++i) ms[written++] = out[i];
        if (enable)
            for (auto& set : pl->ev)
                for (auto& e : set)
                    if (!e) KF_HIP_CHECK(hipEventCreate(&e));
        pl->profile = enable != 0;
        pl->n_rec = pl->n_done = pl->n_acc = 0;
        pl->cur = -1;
        for (double& v : pl->prof_ms) v = 0.0;
    });
    return rc == KF_OK ? written : rc;
}

void kf_plan_destroy(kf_plan* pl) { delete pl; }

}  // extern "C"
