// plan.hip -- planned KIFMM evaluation on one MI355X: work lists, HBM slabs and the passes.
//
// Data layout in HBM (one plan), for a batch of R right-hand sides (R a power of two <= 64):
//   points    xs, ys (tree order, owned by the tree), tau_s[N][R] and near[N][R] (tree order)
//   boxes     cx, cy, bot, top, level, child0, leaf flag, colleagues (global box id)
//   slabs     U, X, Psi: [boxes][R][P] FP64, P = p rounded up to a multiple of 16 (zero padded)
//             U   upward check potential of leaf boxes (P2M), then the downward check potential
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
#include "plan_kernels.h"

using namespace kf;

struct kf_plan {
    const kf_tree* tree = nullptr;
    KfKernel ker{};
    int p = 0, P = 0, NB = 0, prec = 0, L = 0;
    int64_t N = 0, B = 0;
    DevBuf<double> cx, cy, U, X, Psi, tau, ops, surf;
    DevBuf<float> X32, ops32;          // split precision: FP32 densities and M2L operators
    DevBuf<double> m2l_scale;          // split precision: per-level operator scale
    DevBuf<int32_t> bot, top, lvl, child0, coll;
    DevBuf<uint8_t> leafflag;
    DevBuf<int32_t> p2m, up, dn_tgt, dn_par, dn_vsrc, leaves, leaf_coll, loc;
    DevBuf<int4> dn_tiles;
    DevBuf<uint32_t> dn_masks;
    int n_tiles = 0;
    DevBuf<int16_t> slot_op;
    int n_p2m = 0, n_leaves = 0;
    std::vector<int> up_off, up_n, up_nm2m, tile_off, tile_n;
    // batch of right-hand sides the slabs and tiles are laid out for
    int R = 0, lgR = 0, Rcap = 0, nvalid = 0;
    std::vector<int4> h_groups;       // (first target, count, child position q, level)
    std::vector<int32_t> h_vsrc;      // 27 interaction-list slots per target
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

// ----------------------------------------------------------------- host side ------------

// Lays the slabs and downward tiles out for a batch of R right-hand sides (a power of two
// <= 64): slab rows are box * R + r, a tile holds 64/R targets of one (level, child position)
// group.  Slabs grow to the largest R used; X is cleared whenever the layout changes (rows of
// boxes without sources are read as zero children and never written).
static void set_batch(kf_plan* pl, int R, hipStream_t s) {
    if (R == pl->R) return;
    int lg = 0;
    while ((1 << lg) < R) ++lg;
    KF_REQUIRE((1 << lg) == R && R <= 64, KF_ERR_ARG, "batch width must be a power of two <= 64");
    const int bpt = 64 / R;
    std::vector<int4> tiles;
    std::vector<uint32_t> masks;
    pl->tile_off.assign(pl->L, 0);
    pl->tile_n.assign(pl->L, 0);
    int cur_level = -1;
    for (const int4& g : pl->h_groups) {
        if (g.w != cur_level) {
            cur_level = g.w;
            pl->tile_off[cur_level] = (int)tiles.size();
        }
        for (int o = 0; o < g.y; o += bpt) {
            int m = std::min(bpt, g.y - o);
            tiles.push_back(make_int4(g.x + o, m, g.z, g.w));
            uint32_t mk = 0;
            for (int k = 0; k < m; ++k)
                for (int sl = 0; sl < 27; ++sl)
                    if (pl->h_vsrc[(size_t)(g.x + o + k) * 27 + sl] >= 0) mk |= 1u << sl;
            masks.push_back(mk);
        }
        pl->tile_n[cur_level] = (int)tiles.size() - pl->tile_off[cur_level];
    }
    pl->dn_tiles.upload(tiles, s);
    pl->dn_masks.upload(masks, s);
    pl->n_tiles = (int)tiles.size();
    if (R > pl->Rcap) {
        const size_t rows = (size_t)pl->B * R;
        pl->U.alloc(rows * pl->P);
        pl->X.alloc(rows * pl->P);
        pl->Psi.alloc(rows * pl->P);
        pl->tau.alloc((size_t)pl->N * R);
        pl->nearb.alloc((size_t)pl->N * R);
        if (pl->prec == KF_PREC_SPLIT) pl->X32.alloc(rows * pl->P);
        pl->U.zero(s);
        pl->Psi.zero(s);
        pl->Rcap = R;
    }
    pl->X.zero(s);
    if (pl->prec == KF_PREC_SPLIT) KF_HIP_CHECK(hipMemsetAsync(pl->X32.p, 0, pl->X32.n * sizeof(float), s));
    pl->R = R;
    pl->lgR = lg;
}

static void build_plan(kf_plan* pl, const kf_tree* tr, const kf_kernel_desc* kd, int p, int prec,
                       hipStream_t s) {
    KF_REQUIRE(tr && kd, KF_ERR_ARG, "null argument");
    const HostTree& h = tr->h;
    KF_REQUIRE(tr->perm.p, KF_ERR_STATE, "tree has no device points");
    KF_REQUIRE(h.nlevels() >= 2, KF_ERR_ARG, "plan needs a tree with at least two levels");
    KF_REQUIRE(p >= 4 && p <= 64, KF_ERR_ARG, "p (Nequiv) must be in [4, 64]");
    KF_REQUIRE(kd->kind == KF_LAPLACE || kd->kind == KF_MODHELM, KF_ERR_ARG, "unknown kernel kind");
    KF_REQUIRE(kd->kind != KF_MODHELM || kd->k > 0.0, KF_ERR_ARG, "modified Helmholtz needs k > 0");
    KF_REQUIRE(prec == KF_PREC_FP64 || prec == KF_PREC_SPLIT, KF_ERR_ARG, "unknown precision mode");
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
    for (int l = 2; l < L; ++l) {
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
            pl->h_groups.push_back(make_int4(start, (int)tg.size() - start, q, l));
        }
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
    if (prec == KF_PREC_SPLIT) {
        // FP32 M2L operators, each level's 40 blocks scaled to unit max-norm
        std::vector<float> o32((size_t)L * 40 * P * P, 0.0f);
        std::vector<double> sc(L, 1.0);
        for (int l = 1; l < L; ++l) {
            const double* src = ops.data() + (size_t)l * LevelOps::size(P);
            double mx = 0.0;
            for (size_t e = 0; e < (size_t)40 * P * P; ++e) mx = std::max(mx, std::fabs(src[e]));
            sc[l] = mx > 0.0 ? mx : 1.0;
            for (size_t e = 0; e < (size_t)40 * P * P; ++e)
                o32[(size_t)l * 40 * P * P + e] = (float)(src[e] / sc[l]);
        }
        pl->ops32.upload(o32, s);
        pl->m2l_scale.upload(sc, s);
    }
    pl->p2m.upload(p2m, s);
    pl->n_p2m = (int)p2m.size();
    pl->up.upload(up, s);
    pl->dn_tgt.upload(tg, s);
    pl->dn_par.upload(pa, s);
    pl->dn_vsrc.upload(vsrc, s);
    pl->h_vsrc = vsrc;
    pl->slot_op.upload(sop, s);
    pl->leaves.upload(leaves, s);
    pl->leaf_coll.upload(lcoll, s);
    pl->n_leaves = (int)leaves.size();
    KF_HIP_CHECK(hipStreamCreateWithFlags(&pl->side, hipStreamNonBlocking));
    KF_HIP_CHECK(hipEventCreateWithFlags(&pl->ev_fork, hipEventDisableTiming));
    KF_HIP_CHECK(hipEventCreateWithFlags(&pl->ev_join, hipEventDisableTiming));
    set_batch(pl, 1, s);
    KF_HIP_CHECK(hipStreamSynchronize(s));
}

template <int NB>
static void launch_up(kf_plan* pl, int l, hipStream_t s) {
    int n = pl->up_n[l];
    if (!n) return;
    int bpt = 64 >> pl->lgR;
    k_up<NB><<<(n + bpt - 1) / bpt, 256, 0, s>>>(pl->up.p + pl->up_off[l], n, pl->up_nm2m[l], pl->lgR,
                                                 pl->child0.p, pl->ops.p + (size_t)l * LevelOps::size(pl->P),
                                                 pl->U.p, pl->X.p,
                                                 pl->prec == KF_PREC_SPLIT ? pl->X32.p : nullptr);
}

template <int NB>
static void launch_m2l(kf_plan* pl, hipStream_t s) {
    if (!pl->n_tiles) return;
    if (pl->prec == KF_PREC_SPLIT)
        k_m2l_f32<NB><<<pl->n_tiles, 256, 0, s>>>(pl->dn_tiles.p, pl->dn_masks.p, pl->dn_tgt.p, pl->dn_vsrc.p,
                                                  pl->slot_op.p, pl->ops32.p, pl->m2l_scale.p, pl->lgR, pl->X32.p,
                                                  pl->U.p);
    else
        k_m2l<NB><<<pl->n_tiles, 256, 0, s>>>(pl->dn_tiles.p, pl->dn_masks.p, pl->dn_tgt.p, pl->dn_vsrc.p,
                                              pl->slot_op.p, pl->ops.p, LevelOps::size(pl->P), pl->lgR, pl->X.p,
                                              pl->U.p);
}

template <int NB>
static void launch_down(kf_plan* pl, int l, hipStream_t s) {
    int n = pl->tile_n[l];
    if (!n) return;
    k_l2l_c2e<NB><<<n, 256, 0, s>>>(pl->dn_tiles.p + pl->tile_off[l], pl->dn_tgt.p, pl->dn_par.p,
                                    pl->ops.p + (size_t)l * LevelOps::size(pl->P), l >= 3 ? 1 : 0, pl->lgR,
                                    pl->U.p, pl->Psi.p);
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

// Right-hand-side chunk of the VALU kernels for the current batch width.
static int rhs_chunk(int R) { return R < 16 ? R : 16; }

template <int KIND, int RC, bool F32>
static void launch_valu(kf_plan* pl, int which, double* d_pot, int64_t ld_pot, hipStream_t s) {
    const kf_tree* tr = pl->tree;
    const KerT<KIND> ker{pl->ker.k, pl->ker.scale};
    const unsigned ny = (unsigned)(pl->R / RC);
    if (which == 0 && pl->n_p2m)
        k_p2m<KIND, RC, F32><<<dim3(pl->n_p2m, ny), 64, 0, s>>>(pl->p2m.p, pl->lvl.p, pl->cx.p, pl->cy.p, pl->bot.p,
                                                           pl->top.p, tr->xs.p, tr->ys.p, pl->tau.p, pl->surf.p,
                                                           pl->p, pl->P, pl->lgR, ker, pl->U.p);
    if (which == 1 && pl->n_leaves)
        k_p2p<KIND, RC><<<dim3(pl->n_leaves, ny), 64, 0, s>>>(pl->leaves.p, pl->leaf_coll.p, pl->bot.p, pl->top.p,
                                                              tr->xs.p, tr->ys.p, pl->tau.p, pl->lgR, ker,
                                                              pl->nearb.p);
    if (which == 2 && pl->n_leaves)
        k_l2p<KIND, RC, F32><<<dim3(pl->n_leaves, ny), 64, 0, s>>>(pl->leaves.p, pl->lvl.p, pl->cx.p, pl->cy.p,
                                                              pl->bot.p, pl->top.p, tr->xs.p, tr->ys.p, tr->perm.p,
                                                              pl->surf.p, pl->Psi.p, pl->nearb.p, pl->p, pl->P,
                                                              pl->lgR, pl->nvalid, ker, d_pot, ld_pot);
}

template <int KIND, bool F32>
static void valu_rc(kf_plan* pl, int which, double* d_pot, int64_t ld_pot, hipStream_t s) {
    switch (rhs_chunk(pl->R)) {
        case 1: launch_valu<KIND, 1, F32>(pl, which, d_pot, ld_pot, s); break;
        case 2: launch_valu<KIND, 2, F32>(pl, which, d_pot, ld_pot, s); break;
        case 4: launch_valu<KIND, 4, F32>(pl, which, d_pot, ld_pot, s); break;
        case 8: launch_valu<KIND, 8, F32>(pl, which, d_pot, ld_pot, s); break;
        default: launch_valu<KIND, 16, F32>(pl, which, d_pot, ld_pot, s); break;
    }
}

template <int KIND>
static void valu_kind(kf_plan* pl, int which, double* d_pot, int64_t ld_pot, hipStream_t s) {
    // P2M, P2P and L2P are FP64 in both modes.  P2M feeds the check-to-equivalent solve,
    // which amplifies rounding in the high Fourier modes by up to 1/lambda_min (~1e9 at
    // p = 64); FP32 M2L leaves ~1e-7 noise in the downward check potentials that the solve
    // amplifies into Psi, so an FP32 L2P on that Psi would lose ~1e-4 (measured on the oracle).
    valu_rc<KIND, false>(pl, which, d_pot, ld_pot, s);
}

// which: 0 = P2M, 1 = P2P, 2 = L2P
static void run_valu(kf_plan* pl, int which, double* d_pot, int64_t ld_pot, hipStream_t s) {
    if (pl->ker.kind == KF_LAPLACE) valu_kind<KF_LAPLACE>(pl, which, d_pot, ld_pot, s);
    else valu_kind<KF_MODHELM>(pl, which, d_pot, ld_pot, s);
}

static void run_p2p(kf_plan* pl, hipStream_t s) {
    mark(pl, 6, s);
    run_valu(pl, 1, nullptr, 0, s);
    mark(pl, 7, s);
}

// Upward and downward passes for a batch of nvalid densities d_tau[r*ld_tau + i] (the batch
// width R must be set).  With overlap_near, the near field of the same densities is launched
// on the plan's side stream right after the gather; self_batch() joins it.
static void expand_batch(kf_plan* pl, const double* d_tau, int64_t ld_tau, int nvalid, hipStream_t s,
                         bool overlap_near) {
    const kf_tree* tr = pl->tree;
    const int64_t tot = pl->N << pl->lgR;
    pl->nvalid = nvalid;
    begin_profile(pl);
    mark(pl, 0, s);
    k_gather_tau<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(d_tau, ld_tau, nvalid, tr->perm.p, pl->N, pl->lgR,
                                                               pl->tau.p);
    mark(pl, 1, s);
    if (overlap_near) {
        KF_HIP_CHECK(hipEventRecord(pl->ev_fork, s));
        KF_HIP_CHECK(hipStreamWaitEvent(pl->side, pl->ev_fork, 0));
        run_p2p(pl, pl->side);
        KF_HIP_CHECK(hipEventRecord(pl->ev_join, pl->side));
    }
    run_valu(pl, 0, nullptr, 0, s);
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

// Far field + near field at the sources for the batch, input order.
static void self_batch(kf_plan* pl, double* d_pot, int64_t ld_pot, hipStream_t s) {
    if (pl->near_pending) KF_HIP_CHECK(hipStreamWaitEvent(s, pl->ev_join, 0));
    else run_p2p(pl, s);
    pl->near_pending = false;
    run_valu(pl, 2, d_pot, ld_pot, s);
    mark(pl, 5, s);
    KF_HIP_CHECK(hipGetLastError());
}

static int batch_width(int n) {
    int R = 1;
    while (R < n && R < 64) R <<= 1;
    return R;
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
        KF_REQUIRE(nrhs >= 1 && nrhs <= 64, KF_ERR_ARG, "kf_plan_expand keeps 1..64 right-hand sides");
        KF_REQUIRE(nrhs == 1 || ld_tau >= pl->N, KF_ERR_ARG, "leading dimension < n");
        hipStream_t s = (hipStream_t)stream;
        set_batch(pl, batch_width(nrhs), s);
        expand_batch(pl, d_tau, nrhs == 1 ? pl->N : ld_tau, nrhs, s, false);
    });
}

int kf_plan_self(kf_plan* pl, double* d_pot, int64_t ld_pot, void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl && d_pot, KF_ERR_ARG, "null argument");
        KF_REQUIRE(pl->expanded, KF_ERR_STATE, "kf_plan_expand must run first");
        KF_REQUIRE(pl->nvalid == 1 || ld_pot >= pl->N, KF_ERR_ARG, "leading dimension < n");
        self_batch(pl, d_pot, pl->nvalid == 1 ? pl->N : ld_pot, (hipStream_t)stream);
        finish_profile(pl);
    });
}

int kf_plan_eval(kf_plan* pl, const double* d_tau, int64_t ld_tau, int32_t nrhs, double* d_pot,
                 int64_t ld_pot, void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl && d_tau && d_pot, KF_ERR_ARG, "null argument");
        KF_REQUIRE(nrhs >= 1, KF_ERR_ARG, "nrhs must be >= 1");
        KF_REQUIRE(nrhs == 1 || (ld_tau >= pl->N && ld_pot >= pl->N), KF_ERR_ARG, "leading dimension < n");
        if (nrhs == 1) ld_tau = ld_pot = pl->N;
        hipStream_t s = (hipStream_t)stream;
        for (int r0 = 0; r0 < nrhs; r0 += 64) {
            const int nb = std::min(64, nrhs - r0);
            set_batch(pl, batch_width(nb), s);
            expand_batch(pl, d_tau + (size_t)r0 * ld_tau, ld_tau, nb, s, true);
            self_batch(pl, d_pot + (size_t)r0 * ld_pot, ld_pot, s);
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
                                                     pl->surf.p, pl->Psi.p, pl->p, pl->P, pl->lgR,
                                                     KerT<KF_LAPLACE>{pl->ker.k, pl->ker.scale},
                                                     check_self ? 1 : 0, d_pot);
        else
            k_targets<KF_MODHELM><<<nb, 256, 0, s>>>(d_tx, d_ty, m, pl->loc.p, pl->lvl.p, pl->coll.p, pl->cx.p,
                                                     pl->cy.p, pl->bot.p, pl->top.p, tr->xs.p, tr->ys.p, pl->tau.p,
                                                     pl->surf.p, pl->Psi.p, pl->p, pl->P, pl->lgR,
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
