// plan.hip -- planned KIFMM evaluation on one MI355X: work lists, HBM slabs and the passes.
//
// Data layout in HBM (one plan), for a batch of R right-hand sides (R a power of two <= 64):
//   points    xs, ys (tree order, owned by the tree), tau_s[N][R] and near[N][R] (tree order)
//   boxes     cx, cy, bot, top, level, child0, leaf flag, colleagues, slab row (global box id)
//   slabs     U, X, Psi: [rows][R][P] FP64, P = p rounded up to a multiple of 16 (zero padded)
//             U   upward check potential of leaf boxes (P2M), then the downward check potential
//             X   upward equivalent density            (reference Equiv_Densities)
//             Psi downward equivalent density          (reference Local_Expansions)
//   operators per level, see ops_host.h
// Slab rows exist only for boxes whose expansions are used: boxes with sources (levels >= 1)
// take rows [0, n_src); the real boxes without sources at levels >= 2 follow, up to n_all.
// Potentials at the sources need only the first group (a box without sources has no source
// descendants, and its local expansion reaches no source point); expansions for arbitrary
// targets (kf_plan_expand) use all of them.  On trees of points on curves most boxes are empty.
// Passes (reference class_fmm.py:299-352, fmm.py:350-408):
//   k_gather_tau -> k_p2m (all levels) -> k_up per level, bottom-up (M2M + C2E)
//   -> k_m2l (all levels, 27-slot M2L into D) -> k_l2l_c2e per level, top-down (L2L + C2E)
//   -> k_l2p (far field + near field, input order).
// The near field k_p2p depends only on the gathered densities; in kf_plan_eval it runs on a
// second stream of the plan, overlapped with the level-by-level passes, and joins before L2P.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <tuple>
#include <type_traits>
#include <utility>

#include "batch_mfma.h"
#include "cplx.h"
#include "kf_internal.h"
#include "kf_mfma.h"
#include "near_sym.h"
#include "ops_host.h"
#include "ops_dev.h"
#include "plan_dev.h"
#include "plan_kernels.h"

using namespace kf;

// Downward tiles for one batch width R and target set: a tile holds up to 64/R targets of one
// group; masks[t] lists the interaction-list slots any of its targets uses.
#ifndef KF_M2L_SIBLING_ORDER
#define KF_M2L_SIBLING_ORDER 1  // interleave the child positions of each level's tiles
#endif
#ifndef KF_M2L16_GROUP
#define KF_M2L16_GROUP 2  // tiles per workgroup of the FP16-split M2L for batches of 32 and 64
#endif
struct TileSet {
    DevBuf<int4> tiles;
    DevBuf<uint32_t> masks;
    // batches of 32 and 64 right-hand sides: pairs of consecutive tiles of one op group (the
    // FP16-split M2L takes a pair per 512-thread workgroup, one operator staging for both)
    DevBuf<int4> tiles2;
    DevBuf<uint32_t> masks2;
    std::vector<int4> h_tiles2;
    std::vector<int> off2, n2;       // per op: first pair, pair count
    int count2 = 0;
    DevBuf<int2> up_items;           // persistent upward chain: (op, 16-row tile), deepest level first
    std::vector<int2> h_up;
    int n_up_items = 0;
    std::vector<int4> h_tiles;       // host copies: the source of the (possibly still running)
                                     // asynchronous upload stays valid
    std::vector<int> off, n;         // per level: first tile, tile count
    int count = 0;
    bool built = false;
};

// A captured evaluation of one batch width (source targets): the graph of its launches and the
// two nodes whose arguments change from call to call (the gather reads the densities, the L2P
// writes the potentials).
struct GraphSet {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t gather = nullptr, l2p = nullptr, scatter = nullptr;
    hipKernelNodeParams gp{}, lp{}, sp{};  // their launch configurations
    hipEvent_t done = nullptr;       // completion of the last replay
    bool launched = false;
};

// Expansion slabs shared by plans that never evaluate at the same time (kf_workspace_create):
// each buffer grows to the largest need of the plans that use it; `gen` counts the moves, and
// `last` is the plan whose evaluation wrote it last (another plan clears the rows it reads).
struct kf_workspace {
    DevBuf<double> U, Psi, X, tau, nearb;
    DevBuf<float> X32;
    uint64_t gen = 1;
    const kf_plan* last = nullptr;
    std::vector<kf_plan*> users;
};

struct kf_plan {
    std::vector<const kf_tree*> trees; // one tree, or the trees of a forest
    KfKernel ker{};
    kf_ktable ktab;                    // KF_TABLE: the plan's copy of the tabulated kernel
    kf_ktable ktab_im;                 // a complex kernel: the copy of its imaginary part's table
    bool cplx = false;                 // complex kernel: slab rows [re | im] of P / 2 each
    // stored kernel matrices (kf_plan_store_matrices): near field per leaf, P2M per p2m box,
    // L2P per leaf, each with its block offsets
    DevBuf<double> nearM, p2mM, l2pM;
    DevBuf<int64_t> nearM_off, p2mM_off, l2pM_off;
    bool stored = false;
    int p = 0, P = 0, NB = 0, prec = 0, L = 0, n_classes = 1;
    int64_t N = 0, B = 0;
    int64_t n_src = 0, n_all = 0;      // slab rows: boxes with sources, all used boxes
    // points in tree order: the tree's own arrays, or (forest) the trees' arrays concatenated
    // with the permutation offset to the concatenated input
    const double* xs = nullptr;
    const double* ys = nullptr;
    const int32_t* perm = nullptr;
    DevBuf<double> fxs, fys;
    DevBuf<int32_t> fperm;
    DevBuf<int32_t> iperm;             // inverse of perm (input index -> tree index)
    DevLists dl;                       // work lists, built on the device (plan_dev.h)
    DevBuf<double>&cx = dl.cx, &cy = dl.cy;
    DevBuf<double> U, X, Psi, tau, ops, surf;
    DevBuf<float> X32, ops32;          // split precision: FP32 densities and M2L operators
    DevBuf<double> m2l_scale;          // split precision: per-op operator scale
    DevBuf<_Float16> ops16;            // FP16-split: hi and lo planes of the M2L operators
    DevBuf<int32_t>&bot = dl.bot, &top = dl.top, &lvl = dl.lvl, &bop = dl.bop, &child0 = dl.child0, &coll = dl.coll,
                    &rid = dl.rid;
    DevBuf<uint8_t>& leafflag = dl.leaf;
    DevBuf<int32_t>&p2m = dl.p2m, &up_row = dl.urow, &up_crow = dl.crow, &dn_row = dl.trow, &dn_prow = dl.prow,
                    &dn_vsrc = dl.vsrc, &leaves = dl.leaves, &leaf_coll = dl.lcoll;
    DevBuf<int32_t> loc;
    // symmetric near field of one right-hand side (near_sym.h): per leaf and colleague slot, the
    // mutual colleague's leaf index and back slot; usable when no leaf has more than 64 points
    DevBuf<int32_t> pinfo;
    bool sym_ok = false;
    DevBuf<int16_t> slot_op;
    int n_p2m = 0, n_leaves = 0;
    std::vector<int> up_off, up_n, up_nm2m;
    // batch of right-hand sides the slabs and tiles are laid out for; tile tables are kept for
    // every (target set, width) used, slabs are sized for the largest width (Rcap) and row
    // count (rows_cap) used
    int R = 0, lgR = 0, Rcap = 0, nvalid = 0, mode = -1;
    int64_t rows_cap = 0;
    // kf_plan_attach_workspace: U, Psi, X, X32, tau and nearb are views of the workspace's
    // buffers (as of its generation ws_gen) instead of the plan's own
    kf_workspace* ws = nullptr;
    uint64_t ws_gen = 0;
    TileSet ts[2][7];
    const TileSet* tset = nullptr;
    PlanLists lists;                   // host copies of the work lists
    kf_plan_stats stats{};
    int expanded = -1;                 // target set of the last expansion, -1 none
    bool near_pending = false;         // near field of the last expansion in flight on `side`
    DevBuf<double> nearb;              // near-field potentials, tree order
    hipStream_t side = nullptr;        // stream for the overlapped near field
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_defer = nullptr;
    // per-pass profiling: one set of events per evaluation in a ring; totals are summed when
    // a set is reused or when the caller reads them.  Events 0-5 bracket the passes on the
    // evaluation stream, 6-7 the near field (its first share when split), 8 ends the M2L launch
    // inside the downward pass, 9-10 the deferred share of a split near field.
    static constexpr int RING = 64, NEV = 12, NPASS = 6;
    bool ev_split[RING] = {};
    bool ev_serial[RING] = {};         // the near field ran on the evaluation's stream (event 11: after it)
    bool profile = false;
    int cur = -1;
    int64_t n_rec = 0, n_done = 0, n_acc = 0;
    double prof_ms[NPASS] = {0, 0, 0, 0, 0, 0};
    double prof_m2l = 0;
    hipEvent_t ev[RING][NEV] = {};
    // launch-structure tunables (kf_plan_set_tuning / kf_plan_autotune)
    kf_tuning tune{0, 0, 0, 0, 0, 0, 0, 0};
    // persistent chains (kf_tuning.chain): control block (epoch, queues, error bits), per-row
    // publication flags of the upward (X) and downward (Psi) rows, the upward ranges per op
    DevBuf<int> chain_ctl, done_up, done_dn, d_up_off, d_up_n, d_up_nm2m;
    hipStream_t m2ls = nullptr;        // per-level M2L behind the upward pass (m2l_pipeline)
    std::vector<hipEvent_t> ev_up, ev_m2l;
    // graph replay (kf_plan_set_graphs): one captured evaluation per batch width, captured on
    // the plan's own stream `cap` and launched on the caller's stream
    bool use_graphs = false;
    bool capturing = false;
    void* l2p_func = nullptr;          // the L2P instantiation the current width launches
    GraphSet gs[7];
    hipStream_t cap = nullptr;
    void drop_graphs() {
        for (GraphSet& G : gs) {
            if (G.done) {
                if (G.launched) (void)hipEventSynchronize(G.done);
                (void)hipEventDestroy(G.done);
            }
            if (G.exec) (void)hipGraphExecDestroy(G.exec);
            if (G.graph) (void)hipGraphDestroy(G.graph);
            G = GraphSet{};
        }
    }
    ~kf_plan() {
        drop_graphs();
        if (ws) {
            if (ws->last == this) ws->last = nullptr;
            ws->users.erase(std::remove(ws->users.begin(), ws->users.end(), this), ws->users.end());
        }
        if (cap) (void)hipStreamDestroy(cap);
        if (m2ls) (void)hipStreamDestroy(m2ls);
        for (hipEvent_t e : ev_up) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev_m2l) (void)hipEventDestroy(e);
        for (auto& set : ev)
            for (auto& e : set)
                if (e) (void)hipEventDestroy(e);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (ev_defer) (void)hipEventDestroy(ev_defer);
        if (side) (void)hipStreamDestroy(side);
    }
};

namespace kf {

// ----------------------------------------------------------------- host side ------------

// Within each op's range [off, off + n) of a tile array, the child positions' tiles interleaved
// (the i-th tile of each of the four position groups in turn): siblings, whose interaction
// lists share most sources, then run close together in the dispatch order and on one XCD (L2
// reuse of the source rows).  At one right-hand side a tile holds 64 targets of one position,
// and the i-th tiles of the four positions hold children of roughly the same parents.
static void interleave_siblings(std::vector<int4>& v, const std::vector<int>& off, const std::vector<int>& n) {
    for (size_t op = 0; op < off.size(); ++op) {
        if (n[op] < 2 || !KF_M2L_SIBLING_ORDER) continue;
        const int4* first = v.data() + off[op];
        std::vector<int4> byq[4];
        for (int i = 0; i < n[op]; ++i) byq[first[i].z & 3].push_back(first[i]);
        std::vector<int4> merged;
        merged.reserve(n[op]);
        for (size_t i = 0; merged.size() < (size_t)n[op]; ++i)
            for (int q = 0; q < 4; ++q)
                if (i < byq[q].size()) merged.push_back(byq[q][i]);
        std::copy(merged.begin(), merged.end(), v.begin() + off[op]);
    }
}

static void build_tiles(kf_plan* pl, int lg, int mode, TileSet& T, hipStream_t s) {
    std::vector<Tile> tiles;
    build_tile_table(pl->lists, lg, mode, tiles, nullptr, T.off, T.n);
    T.h_tiles.resize(tiles.size());
    for (size_t i = 0; i < tiles.size(); ++i)
        T.h_tiles[i] = make_int4(tiles[i].first, tiles[i].count, tiles[i].q, tiles[i].op);
    T.h_tiles2.clear();
    T.off2.assign(T.off.size(), 0);
    T.n2.assign(T.n.size(), 0);
    if (lg >= 5) {
        // batches of 32 and 64: pairs of consecutive tiles of one op group (the list order)
        for (size_t i = 0; i < T.h_tiles.size();) {  // up to KF_M2L16_GROUP consecutive tiles of one op group
            int4 g = T.h_tiles[i++];
            if (T.n2[g.w]++ == 0) T.off2[g.w] = (int)T.h_tiles2.size();  // pairs never cross ops
            for (int k = 1; k < KF_M2L16_GROUP && i < T.h_tiles.size(); ++k, ++i) {
                const int4 b = T.h_tiles[i];
                if (!(b.z == g.z && b.w == g.w && b.x == g.x + g.y && g.y + b.y <= ((64 * KF_M2L16_GROUP) >> lg)))
                    break;
                g.y += b.y;
            }
            T.h_tiles2.push_back(g);
        }
        interleave_siblings(T.h_tiles2, T.off2, T.n2);
        T.tiles2.upload(T.h_tiles2, s);
        T.masks2.alloc(std::max<size_t>(1, T.h_tiles2.size()));
        dev_tile_masks(T.tiles2.p, (int)T.h_tiles2.size(), pl->dn_vsrc.p, T.masks2.p, s);
    }
    interleave_siblings(T.h_tiles, T.off, T.n);  // after the pairs, which join consecutive list tiles
    T.tiles.upload(T.h_tiles, s);
    T.count = (int)T.h_tiles.size();
    T.masks.alloc(std::max(1, T.count));
    dev_tile_masks(T.tiles.p, T.count, pl->dn_vsrc.p, T.masks.p, s);  // from the device interaction lists
    T.count2 = (int)T.h_tiles2.size();
    // upward chain items: 16-row tiles of each op's list, levels from the deepest up
    T.h_up.clear();
    for (int l = pl->L - 1; l >= 1; --l)
        for (int c = 0; c < pl->n_classes; ++c) {
            const int op = c * pl->L + l;
            const long rows = (long)pl->up_n[op] << lg;
            for (long t = 0; 16 * t < rows; ++t) T.h_up.push_back(make_int2(op, (int)t));
        }
    T.n_up_items = (int)T.h_up.size();
    T.up_items.upload(T.h_up, s);
    T.built = true;
}

// Doubles of the near-field buffer for R right-hand sides (complex kernels: re and im).
// Near-field potentials of R right-hand sides; one right-hand side in the symmetric form also
// keeps KF_SYM_SLOTS partial sums per point behind them.
static size_t near_len(const kf_plan* pl, int64_t R) {
    const size_t n = (size_t)pl->N * R * (pl->cplx ? 2 : 1);
    return pl->sym_ok ? std::max(n, (size_t)pl->N * (1 + KF_SYM_SLOTS)) : n;
}

// The plan's slabs as views of its workspace for R right-hand sides over `rows` rows; the
// workspace's buffers grow (and its generation advances) where this plan needs more.  Rows
// are cleared by claim_workspace() before an evaluation, not here.
static void bind_workspace(kf_plan* pl, int R, int64_t rows, hipStream_t s) {
    kf_workspace* w = pl->ws;
    const size_t P = (size_t)pl->P;
    const size_t nU = (size_t)rows * R * P, nX = (size_t)pl->n_src * R * P, nT = (size_t)pl->N * R;
    const size_t nN = near_len(pl, R), n32 = pl->prec != KF_PREC_FP64 ? nX : 0;
    bool grew = false;
    auto fit = [&](auto& buf, size_t need) {
        if (buf.n < need) {
            buf.alloc(need);
            grew = true;
        }
    };
    fit(w->U, nU);
    fit(w->Psi, nU);
    fit(w->X, nX);
    fit(w->tau, nT);
    fit(w->nearb, nN);
    fit(w->X32, n32);
    if (grew) {
        ++w->gen;
        w->last = nullptr;  // fresh memory: every user clears before it reads
    }
    (void)s;
    pl->U.view(w->U.p, nU);
    pl->Psi.view(w->Psi.p, nU);
    pl->X.view(w->X.p, nX);
    pl->tau.view(w->tau.p, nT);
    pl->nearb.view(w->nearb.p, nN);
    pl->X32.view(w->X32.p, n32);
    pl->done_up.release();  // persistent chains are off on a workspace
    pl->done_dn.release();
    pl->ws_gen = w->gen;
    pl->Rcap = R;
    pl->rows_cap = rows;
}

// Slabs for R right-hand sides over `rows` rows (U and Psi; X holds only boxes with sources).
static void grow_slabs(kf_plan* pl, int R, int64_t rows, hipStream_t s) {
    pl->drop_graphs();  // captured evaluations hold the old slab addresses
    R = std::max(R, pl->Rcap);
    rows = std::max(rows, pl->rows_cap);
    if (pl->ws) return bind_workspace(pl, R, rows, s);
    const size_t P = (size_t)pl->P;
    pl->U.alloc((size_t)rows * R * P);
    pl->Psi.alloc((size_t)rows * R * P);
    pl->done_dn.alloc((size_t)rows * R);
    pl->done_dn.zero(s);  // flags below every epoch (epochs start at 1)
    if (R > pl->Rcap || !pl->X.p) {
        pl->X.alloc((size_t)pl->n_src * R * P);
        pl->done_up.alloc((size_t)pl->n_src * R);
        pl->done_up.zero(s);
        if (pl->prec != KF_PREC_FP64) pl->X32.alloc((size_t)pl->n_src * R * P);
        pl->tau.alloc((size_t)pl->N * R);
        pl->nearb.alloc(near_len(pl, R));
    }
    // every slab row a pass reads is written earlier in the same evaluation (boxes without
    // sources are skipped as children and as interaction-list sources); cleared once here
    pl->U.zero(s);
    pl->X.zero(s);
    pl->Psi.zero(s);
    if (pl->prec != KF_PREC_FP64) KF_HIP_CHECK(hipMemsetAsync(pl->X32.p, 0, pl->X32.n * sizeof(float), s));
    pl->Rcap = R;
    pl->rows_cap = rows;
}

static int log2_width(int R) {
    int lg = 0;
    while ((1 << lg) < R) ++lg;
    KF_REQUIRE((1 << lg) == R && R <= 64, KF_ERR_ARG, "batch width must be a power of two <= 64");
    return lg;
}

static int64_t mode_rows(const kf_plan* pl, int mode) { return mode == KF_TGT_ALL ? pl->n_all : pl->n_src; }

// Lays the evaluation out for a batch of R right-hand sides (a power of two <= 64) and a
// target set: slab rows are row * R + r.  Tile tables are built on first use; slabs only
// grow.  After reserve_batch(n), source-target evaluations of width <= n neither allocate
// nor copy.
// Before an evaluation on a workspace: follow its buffers if another plan moved them.  Nothing
// is cleared: every slab element an evaluation reads is written earlier in the same evaluation
// (P2M, M2M + C2E, M2L and L2L + C2E write whole rows, padding columns included, for every
// box they cover, and the passes read only rows of boxes they cover), which the tests check
// by filling the workspace with NaN between evaluations (kf_workspace_fill).
static void claim_workspace(kf_plan* pl, hipStream_t s) {
    kf_workspace* w = pl->ws;
    if (!w) return;
    if (pl->ws_gen != w->gen) grow_slabs(pl, pl->Rcap, pl->rows_cap, s);
    if (w->last != pl) {
        w->last = pl;
        pl->expanded = -1;
    }
}

__global__ void k_fill_f64(double* __restrict__ p, size_t n, double v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f32(float* __restrict__ p, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

static void set_batch(kf_plan* pl, int R, int mode, hipStream_t s) {
    if (R != pl->R || mode != pl->mode) {
        const int lg = log2_width(R);
        TileSet& T = pl->ts[mode][lg];
        if (!T.built) build_tiles(pl, lg, mode, T, s);
        if (R > pl->Rcap || mode_rows(pl, mode) > pl->rows_cap) grow_slabs(pl, R, mode_rows(pl, mode), s);
        pl->R = R;
        pl->lgR = lg;
        pl->mode = mode;
        pl->tset = &T;
    }
    claim_workspace(pl, s);
}

static void reserve_batch(kf_plan* pl, int R, hipStream_t s) {
    const int lg = log2_width(R);
    for (int k = 0; k <= lg; ++k)
        if (!pl->ts[KF_TGT_SOURCES][k].built) build_tiles(pl, k, KF_TGT_SOURCES, pl->ts[KF_TGT_SOURCES][k], s);
    if (R > pl->Rcap || pl->n_src > pl->rows_cap) grow_slabs(pl, R, pl->n_src, s);
    if (pl->ws && pl->ws_gen != pl->ws->gen) grow_slabs(pl, pl->Rcap, pl->rows_cap, s);
}

// Device bytes of the plan and its tree when laid out for source-target batches of width R:
// Bytes of the expansion slabs for R right-hand sides over `rows` rows.
static int64_t slab_bytes(const kf_plan* pl, int64_t R, int64_t rows) {
    int64_t b = 2 * rows * R * pl->P * 8 + pl->n_src * R * pl->P * 8 + pl->N * R * 8 + (int64_t)near_len(pl, R) * 8;
    if (pl->prec != KF_PREC_FP64) b += pl->n_src * R * pl->P * 4;
    return b;
}

// what reserve_batch(R) would leave allocated (exact; nothing shrinks).  A plan on a workspace
// counts none of the slabs: they are the workspace's (kf_plan_slab_bytes).
static int64_t footprint(const kf_plan* pl, int R) {
    const int lg = log2_width(R);
    int64_t b = 0;
    std::vector<const TreePoints*> seen;
    for (const kf_tree* tr : pl->trees) {
        const TreePoints* tp = tr->pts.get();
        if (!tp || std::find(seen.begin(), seen.end(), tp) != seen.end()) continue;
        seen.push_back(tp);
        b += (int64_t)(tp->xs.n + tp->ys.n) * 8 + (int64_t)(tp->perm.n + tp->gperm.n) * 4;
    }
    b += (int64_t)(pl->fxs.n + pl->fys.n) * 8 + (int64_t)(pl->fperm.n + pl->iperm.n) * 4;
    b += (int64_t)pl->ktab.dc.n * 8;
    b += (int64_t)(pl->nearM.n + pl->nearM_off.n + pl->p2mM.n + pl->p2mM_off.n + pl->l2pM.n + pl->l2pM_off.n) * 8;
    b += (int64_t)(pl->cx.n + pl->cy.n + pl->ops.n + pl->surf.n + pl->m2l_scale.n) * 8;
    b += (int64_t)pl->ops32.n * 4 + (int64_t)pl->ops16.n * 2;
    b += (int64_t)(pl->bot.n + pl->top.n + pl->lvl.n + pl->bop.n + pl->child0.n + pl->coll.n + pl->rid.n) * 4;
    b += (int64_t)(pl->p2m.n + pl->up_row.n + pl->up_crow.n + pl->dn_row.n + pl->dn_prow.n + pl->dn_vsrc.n +
                   pl->leaves.n + pl->leaf_coll.n + pl->loc.n) * 4;
    b += (int64_t)pl->leafflag.n + (int64_t)pl->slot_op.n * 2;
    for (int m = 0; m < 2; ++m)
        for (int k = 0; k < 7; ++k) {
            const bool will = m == KF_TGT_SOURCES && k <= lg;
            if (!will && !pl->ts[m][k].built) continue;
            b += tile_count(pl->lists, k, m) * (int64_t)(sizeof(int4) + sizeof(uint32_t));
        }
    if (!pl->ws) b += slab_bytes(pl, std::max(R, pl->Rcap), std::max<int64_t>(pl->n_src, pl->rows_cap));
    return b;
}

// Points of a forest in one tree-ordered array: tree i's points at [pt_off[i], pt_off[i+1]),
// its permutation shifted to index the concatenated input.
__global__ void k_concat_points(const double* __restrict__ xs, const double* __restrict__ ys,
                                const int32_t* __restrict__ perm, int64_t n, int64_t off,
                                double* __restrict__ oxs, double* __restrict__ oys, int32_t* __restrict__ operm) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    oxs[off + i] = xs[i];
    oys[off + i] = ys[i];
    operm[off + i] = (int32_t)(perm[i] + off);
}

// Operators (real form, build_level_ops_complex) and large circles of every level of a complex
// kernel's plan, built on the host and uploaded; the circles at row stride P as build_ops_dev.
static void build_ops_complex(kf_plan* pl, hipStream_t s) {
    const int P = pl->P, p = pl->p, L = pl->L, Ph = P / 2;
    const double W = pl->lists.class_width[0];
    const size_t sz = LevelOps::size(P);
    std::vector<double> h((size_t)L * sz, 0.0), hs((size_t)L * 2 * P, 0.0);
    for (int l = 0; l < L; ++l) {
        if (l >= 1) build_level_ops_complex(pl->ker, W, l, p, Ph, l >= 2, l <= L - 2, h.data() + (size_t)l * sz);
        std::vector<double> a, b, lx, ly;
        level_circles(std::ldexp(W, -l), p, a, b, lx, ly);
        for (int j = 0; j < p; ++j) {
            hs[(size_t)(2 * l) * P + j] = lx[j];
            hs[(size_t)(2 * l + 1) * P + j] = ly[j];
        }
    }
    pl->ops.upload(h, s);
    pl->surf.upload(hs, s);
    KF_HIP_CHECK(hipStreamSynchronize(s));  // the host vectors are the upload sources
}

static void build_plan(kf_plan* pl, const kf_tree* const* trees, int n_trees, const kf_kernel_desc* kd, int p,
                       int prec, hipStream_t s) {
    KF_REQUIRE(trees && kd && n_trees >= 1, KF_ERR_ARG, "null argument");
    KF_REQUIRE(p >= 4 && p <= 64, KF_ERR_ARG, "p (Nequiv) must be in [4, 64]");
    KF_REQUIRE(kd->kind == KF_LAPLACE || kd->kind == KF_MODHELM || kd->kind == KF_TABLE, KF_ERR_ARG,
               "unknown kernel kind");
    KF_REQUIRE(kd->kind != KF_MODHELM || kd->k > 0.0, KF_ERR_ARG, "modified Helmholtz needs k > 0");
    KF_REQUIRE(kd->kind != KF_TABLE || kd->table, KF_ERR_ARG, "a tabulated kernel needs its table");
    KF_REQUIRE(prec == KF_PREC_FP64 || prec == KF_PREC_SPLIT || prec == KF_PREC_SPLIT16, KF_ERR_ARG,
               "unknown precision mode");
    const bool cplx = kd->table_im != nullptr;
    if (cplx) {
        KF_REQUIRE(kd->kind == KF_TABLE && kd->table, KF_ERR_ARG, "a complex kernel is a pair of tables");
        KF_REQUIRE(p <= 32, KF_ERR_ARG, "complex kernels take p (Nequiv) <= 32 (slab rows of 2 p doubles <= 64)");
        KF_REQUIRE(prec == KF_PREC_FP64, KF_ERR_ARG, "complex kernels run in FP64");
        KF_REQUIRE(n_trees == 1, KF_ERR_ARG, "complex kernels take single-tree plans");
    }
    int64_t npts = 0;
    for (int i = 0; i < n_trees; ++i) {
        const kf_tree* tr = trees[i];
        KF_REQUIRE(tr, KF_ERR_ARG, "null tree");
        KF_REQUIRE(tr->perm && tr->dev, KF_ERR_STATE, "tree has no device points");
        KF_REQUIRE(tr->levels() >= 2, KF_ERR_ARG, "plan needs trees with at least two levels");
        pl->trees.push_back(tr);
        npts += tr->h.n;
    }
    KF_REQUIRE(npts < ((int64_t)1 << 31), KF_ERR_ARG, "a plan holds fewer than 2^31 points");
    pl->ker = KfKernel{kd->kind, kd->k, kd->scale};
    if (kd->kind == KF_TABLE) {  // the plan's own copy of the table
        const kf_ktable& t = *kd->table;
        pl->ktab.c = t.c;
        pl->ktab.dc.upload(pl->ktab.c, s);
        pl->ktab.nseg = t.nseg;
        pl->ktab.deg = t.deg;
        pl->ktab.s0 = t.s0;
        pl->ktab.ld = t.ld;
        pl->ktab.du = t.du;
        pl->ktab.g0 = t.g0;
        pl->ker.th = pl->ktab.view(false);
        pl->ker.td = pl->ktab.view(true);
    }
    if (cplx) {  // the imaginary part's table, copied the same way
        const kf_ktable& t = *kd->table_im;
        pl->ktab_im.c = t.c;
        pl->ktab_im.dc.upload(pl->ktab_im.c, s);
        pl->ktab_im.nseg = t.nseg;
        pl->ktab_im.deg = t.deg;
        pl->ktab_im.s0 = t.s0;
        pl->ktab_im.ld = t.ld;
        pl->ktab_im.du = t.du;
        pl->ktab_im.g0 = t.g0;
        pl->ker.cplx = 1;
        pl->ker.th_im = pl->ktab_im.view(false);
        pl->ker.td_im = pl->ktab_im.view(true);
        pl->cplx = true;
    }
    pl->p = p;
    pl->NB = (p + 15) / 16 * (cplx ? 2 : 1);  // complex: rows [re | im] of 16 ceil(p / 16) each
    pl->P = 16 * pl->NB;
    pl->prec = prec;
    const int P = pl->P;
    PlanLists& ls = pl->lists;
    build_dev_lists(pl->trees, p, P, ls, pl->dl, s);
    if (pl->ker.kind == KF_TABLE) {  // operator circles and point pairs lie within 2 root widths
        double wmax = 0.0;
        for (double w : ls.class_width) wmax = std::max(wmax, w);
        KF_REQUIRE((double)(pl->ktab.s0 + pl->ktab.nseg) * pl->ktab.du >= std::log2(4.0 * wmax * wmax), KF_ERR_ARG,
                   "the kernel table must reach r^2 = (2 W)^2, W the root box width");
        KF_REQUIRE(!cplx || (double)(pl->ktab_im.s0 + pl->ktab_im.nseg) * pl->ktab_im.du >=
                                std::log2(4.0 * wmax * wmax),
                   KF_ERR_ARG, "the kernel table must reach r^2 = (2 W)^2, W the root box width");
    }
    pl->L = ls.L;
    pl->n_classes = ls.n_classes;
    pl->N = npts;
    pl->B = ls.box_count;
    pl->n_src = ls.n_src;
    pl->n_all = ls.n_all;
    pl->up_off = ls.up_off;
    pl->up_n = ls.up_n;
    pl->up_nm2m = ls.up_nm2m;
    pl->stats = ls.stats;
    if (!cplx && ls.leaf_count > 0) {
        int max_leaf = 0;
        pl->pinfo.alloc((size_t)(9 * ls.leaf_count));
        build_leaf_pairs(pl->leaves.p, (int)ls.leaf_count, pl->leaf_coll.p, pl->bot.p, pl->top.p, ls.box_count,
                         pl->pinfo.p, &max_leaf, s);
        pl->sym_ok = max_leaf <= KF_SYM_MAX_LEAF;
    }
    const int L = pl->L, n_ops = pl->n_classes * L;
    // circles and operators of every (class, level), on the device (complex kernels: on the host,
    // in real form)
    if (cplx) build_ops_complex(pl, s);
    else build_ops_dev(pl->ker, ls.class_width, L, p, P, pl->ops, pl->surf, s);
    if (prec != KF_PREC_FP64) build_ops32_dev(pl->ops, n_ops, L, P, pl->ops32, pl->m2l_scale, s);
    if (prec == KF_PREC_SPLIT16) build_ops16_dev(pl->ops, n_ops, L, P, pl->m2l_scale, pl->ops16, s);
    const VSlots& vs = vslots();
    std::vector<int16_t> sop(4 * 27);
    for (int q = 0; q < 4; ++q)
        for (int sl = 0; sl < 27; ++sl) sop[q * 27 + sl] = (int16_t)vs.op[q][sl];
    // points: one tree's own arrays; the whole batch of one batch build in its order (already
    // concatenated, with the permutation into the concatenated input); else a copy
    bool whole_batch = n_trees > 1 && trees[0]->pts && trees[0]->pts->n_trees == n_trees && trees[0]->pts->gperm.p;
    for (int i = 0; whole_batch && i < n_trees; ++i)
        whole_batch = trees[i]->pts == trees[0]->pts && trees[i]->slot == i;
    if (n_trees == 1) {
        pl->xs = trees[0]->xs;
        pl->ys = trees[0]->ys;
        pl->perm = trees[0]->perm;
    } else if (whole_batch) {
        pl->xs = trees[0]->pts->xs.p;
        pl->ys = trees[0]->pts->ys.p;
        pl->perm = trees[0]->pts->gperm.p;
    } else {
        pl->fxs.alloc((size_t)npts);
        pl->fys.alloc((size_t)npts);
        pl->fperm.alloc((size_t)npts);
        for (int i = 0; i < n_trees; ++i) {
            const kf_tree* tr = trees[i];
            const int64_t n = tr->h.n;
            k_concat_points<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(tr->xs, tr->ys, tr->perm, n,
                                                                        ls.pt_off[i], pl->fxs.p, pl->fys.p,
                                                                        pl->fperm.p);
        }
        KF_HIP_CHECK(hipGetLastError());
        pl->xs = pl->fxs.p;
        pl->ys = pl->fys.p;
        pl->perm = pl->fperm.p;
    }
    pl->iperm.alloc((size_t)npts);
    k_invert_perm<<<(unsigned)((npts + 255) / 256), 256, 0, s>>>(pl->perm, npts, pl->iperm.p);
    KF_HIP_CHECK(hipGetLastError());
    pl->n_p2m = (int)ls.p2m_count;
    pl->slot_op.upload(sop, s);
    pl->n_leaves = (int)ls.leaf_count;
    KF_HIP_CHECK(hipStreamCreateWithFlags(&pl->side, hipStreamNonBlocking));
    KF_HIP_CHECK(hipEventCreateWithFlags(&pl->ev_fork, hipEventDisableTiming));
    KF_HIP_CHECK(hipEventCreateWithFlags(&pl->ev_join, hipEventDisableTiming));
    KF_HIP_CHECK(hipEventCreateWithFlags(&pl->ev_defer, hipEventDisableTiming));
    pl->d_up_off.upload(pl->up_off, s);
    pl->d_up_n.upload(pl->up_n, s);
    pl->d_up_nm2m.upload(pl->up_nm2m, s);
    pl->chain_ctl.alloc(4);
    pl->chain_ctl.zero(s);
    set_batch(pl, 1, KF_TGT_SOURCES, s);
    KF_HIP_CHECK(hipStreamSynchronize(s));  // the host vectors above are the upload sources
}

// Device operators against ops_host.cpp's for every (class, level): blocks to 16 units in the
// last place of the block's largest entry, the basis exactly, lambda to 1e-13 of the largest
// (E's entries carry one rounding each, which moves the smallest eigenvalues by that much).
static std::string compare_ops_with_host(const kf_plan* pl, hipStream_t s) {
    const int P = pl->P, p = pl->p, L = pl->L;
    const size_t sz = LevelOps::size(P);
    std::vector<double> dev(pl->ops.n), host(sz);
    KF_HIP_CHECK(hipMemcpyAsync(dev.data(), pl->ops.p, dev.size() * 8, hipMemcpyDeviceToHost, s));
    KF_HIP_CHECK(hipStreamSynchronize(s));
    for (int c = 0; c < pl->n_classes; ++c)
        for (int l = 1; l < L; ++l) {
            const int op = c * L + l;
            if (pl->cplx)
                build_level_ops_complex(pl->ker, pl->lists.class_width[c], l, p, P / 2, l >= 2, l <= L - 2,
                                        host.data());
            else
                build_level_ops(pl->ker, pl->lists.class_width[c], l, p, P, l >= 2, l <= L - 2, host.data());
            const double* d = dev.data() + (size_t)op * sz;
            const std::string where = " (class " + std::to_string(c) + ", level " + std::to_string(l) + ")";
            for (int b = 0; b < 48; ++b) {
                double mx = 0.0, diff = 0.0;
                for (size_t k = 0; k < (size_t)P * P; ++k) {
                    mx = std::max(mx, std::fabs(host[(size_t)b * P * P + k]));
                    diff = std::max(diff, std::fabs(host[(size_t)b * P * P + k] - d[(size_t)b * P * P + k]));
                }
                if (diff > 16 * 2.220446049250313e-16 * mx)
                    return "block " + std::to_string(b) + where + ": difference " + std::to_string(diff / mx);
            }
            for (size_t k = LevelOps::F(P); k < LevelOps::invlam(P); ++k)
                if (host[k] != d[k]) return "Fourier basis" + where;
            double lmax = 0.0, ldiff = 0.0;
            for (int k = 0; k < p; ++k) lmax = std::max(lmax, std::fabs(1.0 / host[LevelOps::invlam(P) + k]));
            for (int k = 0; k < p; ++k)
                ldiff = std::max(ldiff, std::fabs(1.0 / host[LevelOps::invlam(P) + k] - 1.0 / d[LevelOps::invlam(P) + k]));
            if (ldiff > 1e-13 * lmax) return "eigenvalues" + where + ": difference " + std::to_string(ldiff / lmax);
        }
    return "";
}

// Rows per workgroup of a level launch over `rows` rows: the tuning's choice, else 16 (the
// column-split form: the most workgroups and the shortest dependent MFMA chains) while the
// level has fewer than 24K rows (1.5K such workgroups), 64 (each workgroup loads an operator
// into LDS once for 64 rows) above.  At one right-hand side every level of the 1M-point
// clustered tree (at most 16K boxes) takes the column-split form.
static constexpr long KF_LEVEL_ROWS_SPLIT = 24L * 1024;
static int level_rows(const kf_plan* pl, long rows) {
    if (pl->tune.level_rows) return pl->tune.level_rows;
    if (rows < KF_LEVEL_ROWS_SPLIT) return 16;
    return pl->lgR >= 5 ? 128 : 64;  // wide batches: two boxes per workgroup share the operator staging
}

template <int NB>
static void launch_up(kf_plan* pl, int op, hipStream_t s) {
    int n = pl->up_n[op];
    if (!n) return;
    const int32_t* urow = pl->up_row.p + pl->up_off[op];
    const int32_t* crow = pl->up_crow.p + 4 * (size_t)pl->up_off[op];
    const double* ops = pl->ops.p + (size_t)op * LevelOps::size(pl->P);
    float* x32 = pl->prec != KF_PREC_FP64 ? pl->X32.p : nullptr;
    const long rows = (long)n << pl->lgR;
    switch (level_rows(pl, rows)) {
        case 16:
            k_up_cs<NB, 1><<<(int)((rows + 15) / 16), 64 * NB, 0, s>>>(urow, crow, n, pl->up_nm2m[op], pl->lgR, ops,
                                                                        pl->U.p, pl->X.p, x32);
            break;
        case 32:
            k_up_cs<NB, 2><<<(int)((rows + 31) / 32), 64 * NB, 0, s>>>(urow, crow, n, pl->up_nm2m[op], pl->lgR, ops,
                                                                        pl->U.p, pl->X.p, x32);
            break;
        case 128:
            k_up<NB, 8><<<(int)((rows + 127) / 128), 512, 0, s>>>(urow, crow, n, pl->up_nm2m[op], pl->lgR, ops,
                                                                   pl->U.p, pl->X.p, x32);
            break;
        default:
            k_up<NB><<<(int)((rows + 63) / 64), 256, 0, s>>>(urow, crow, n, pl->up_nm2m[op], pl->lgR, ops, pl->U.p,
                                                             pl->X.p, x32);
    }
}

// Persistent chains: one launch per pass, at most KF_CHAIN_GRID workgroups taking items from
// the queue (whatever fraction of them is resident, the in-order queue keeps making progress).
#ifndef KF_CHAIN_GRID
#define KF_CHAIN_GRID 512
#endif
static bool use_chain(const kf_plan* pl) {
    if (pl->ws || pl->cplx) return false;  // flags are the plan's own; complex plans: per-level launches
    if (pl->tune.chain == 1) return false;
    if (pl->tune.chain == 2) return true;
    return pl->R <= 4;  // narrow batches: the per-level launches are latency-bound
}

// Where an overlapped near field starts (KF_FORK_*): near_fork, or by default after the P2M
// (the P2M then runs alone, and the near field overlaps the latency-bound upward pass);
// KF_FORK_NONE: no overlap, the near field runs on the evaluation's stream before the L2P (each
// pass then has the device to itself, as for timing the passes one by one).
enum { KF_FORK_AFTER_UP = 1, KF_FORK_AFTER_P2M = 2, KF_FORK_WITH_P2M = 3, KF_FORK_NONE = 4 };
static int fork_point(const kf_plan* pl) { return pl->tune.near_fork == 0 ? KF_FORK_AFTER_P2M : pl->tune.near_fork; }

template <int NB>
static void launch_up_chain(kf_plan* pl, hipStream_t s) {
    const TileSet& T = *pl->tset;
    if (!T.n_up_items) return;
    float* x32 = pl->prec != KF_PREC_FP64 ? pl->X32.p : nullptr;
    k_up_chain<NB><<<std::min(T.n_up_items, KF_CHAIN_GRID), 64 * NB, 0, s>>>(
        T.up_items.p, T.n_up_items, pl->up_row.p, pl->up_crow.p, pl->d_up_off.p, pl->d_up_n.p, pl->d_up_nm2m.p,
        pl->lgR, pl->ops.p, pl->U.p, pl->X.p, x32, pl->done_up.p, pl->chain_ctl.p);
}

template <int NB>
static void launch_down_chain(kf_plan* pl, hipStream_t s) {
    const TileSet& T = *pl->tset;
    if (!T.count) return;
    k_down_chain<NB><<<std::min(4 * T.count, KF_CHAIN_GRID), 64 * NB, 0, s>>>(
        T.tiles.p, T.count, pl->dn_row.p, pl->dn_prow.p, pl->ops.p, pl->L, pl->lgR, pl->U.p, pl->Psi.p,
        pl->done_dn.p, pl->chain_ctl.p);
}

template <int NB>
static void launch_m2l(kf_plan* pl, hipStream_t s) {
    const TileSet& T = *pl->tset;
    if (!T.count) return;
    if (pl->prec == KF_PREC_SPLIT16) {
        constexpr int WG = 4 * KF_M2L16_GROUP;  // waves per workgroup over a group of tiles
        if (T.count2 > 0)  // groups of tiles, one operator staging for all
            k_m2l_f16<NB, WG><<<T.count2, 64 * WG, 0, s>>>(T.tiles2.p, T.masks2.p, pl->dn_row.p, pl->dn_vsrc.p,
                                                           pl->slot_op.p, pl->ops16.p, pl->m2l_scale.p, pl->lgR,
                                                           pl->X32.p, pl->U.p);
        else
            k_m2l_f16<NB, 4><<<T.count, 256, 0, s>>>(T.tiles.p, T.masks.p, pl->dn_row.p, pl->dn_vsrc.p,
                                                     pl->slot_op.p, pl->ops16.p, pl->m2l_scale.p, pl->lgR, pl->X32.p,
                                                     pl->U.p);
    } else if (pl->prec == KF_PREC_SPLIT)
        k_m2l_f32<NB><<<T.count, 256, 0, s>>>(T.tiles.p, T.masks.p, pl->dn_row.p, pl->dn_vsrc.p,
                                                  pl->slot_op.p, pl->ops32.p, pl->m2l_scale.p, pl->lgR, pl->X32.p,
                                                  pl->U.p);
    else
        k_m2l<NB><<<T.count, 256, 0, s>>>(T.tiles.p, T.masks.p, pl->dn_row.p, pl->dn_vsrc.p,
                                              pl->slot_op.p, pl->ops.p, LevelOps::size(pl->P), pl->lgR, pl->X.p,
                                              pl->U.p);
}

// M2L of the tiles of one op index (the pipelined form: one launch per level and class).
template <int NB>
static void launch_m2l_op(kf_plan* pl, int op, hipStream_t s) {
    const TileSet& T = *pl->tset;
    const int n = T.n[op], o = T.off[op];
    if (!n) return;
    if (pl->prec == KF_PREC_SPLIT)
        k_m2l_f32<NB><<<n, 256, 0, s>>>(T.tiles.p + o, T.masks.p + o, pl->dn_row.p, pl->dn_vsrc.p, pl->slot_op.p,
                                        pl->ops32.p, pl->m2l_scale.p, pl->lgR, pl->X32.p, pl->U.p);
    else
        k_m2l<NB><<<n, 256, 0, s>>>(T.tiles.p + o, T.masks.p + o, pl->dn_row.p, pl->dn_vsrc.p, pl->slot_op.p,
                                    pl->ops.p, LevelOps::size(pl->P), pl->lgR, pl->X.p, pl->U.p);
}

template <int NB>
static void launch_down(kf_plan* pl, int op, hipStream_t s) {
    const TileSet& T = *pl->tset;
    int n = T.n[op];
    if (!n) return;
    const int l = op % pl->L;
    const int4* tiles = T.tiles.p + T.off[op];
    const double* ops = pl->ops.p + (size_t)op * LevelOps::size(pl->P);
    const int l2l = l >= 3 ? 1 : 0;
    switch (level_rows(pl, 64L * n)) {
        case 16:
            k_l2l_c2e_cs<NB, 1><<<4 * n, 64 * NB, 0, s>>>(tiles, pl->dn_row.p, pl->dn_prow.p, ops, l2l, pl->lgR,
                                                          pl->U.p, pl->Psi.p);
            break;
        case 32:
            k_l2l_c2e_cs<NB, 2><<<2 * n, 64 * NB, 0, s>>>(tiles, pl->dn_row.p, pl->dn_prow.p, ops, l2l, pl->lgR,
                                                          pl->U.p, pl->Psi.p);
            break;
        case 128:
            if (T.count2 > 0) {  // tile pairs (batches of 32 and 64)
                k_l2l_c2e<NB, 8><<<T.n2[op], 512, 0, s>>>(T.tiles2.p + T.off2[op], pl->dn_row.p, pl->dn_prow.p, ops,
                                                          l2l, pl->lgR, pl->U.p, pl->Psi.p);
                break;
            }
            [[fallthrough]];
        default:
            k_l2l_c2e<NB><<<n, 256, 0, s>>>(tiles, pl->dn_row.p, pl->dn_prow.p, ops, l2l, pl->lgR, pl->U.p, pl->Psi.p);
    }
}

static void accumulate(kf_plan* pl, int set) {
    // a set whose evaluation did not reach the self pass (expand only) contributes nothing
    if (hipEventSynchronize(pl->ev[set][5]) != hipSuccess) return;
    for (int i = 0; i < 5; ++i) {
        float ms = 0.f;
        // the L2P pass starts after an unoverlapped near field
        const int from = i == 4 && pl->ev_serial[set] ? 11 : i;
        if (hipEventElapsedTime(&ms, pl->ev[set][from], pl->ev[set][i + 1]) == hipSuccess) pl->prof_ms[i] += ms;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pl->ev[set][6], pl->ev[set][7]) == hipSuccess) pl->prof_ms[5] += ms;
    if (pl->ev_split[set] && hipEventElapsedTime(&ms, pl->ev[set][9], pl->ev[set][10]) == hipSuccess)
        pl->prof_ms[5] += ms;
    if (hipEventElapsedTime(&ms, pl->ev[set][3], pl->ev[set][8]) == hipSuccess) pl->prof_m2l += ms;
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

// Batches of at least this many right-hand sides move densities and potentials between input
// and tree order by rows (k_gather_tau with iperm, L2P into the near-field slab, then
// k_scatter_pot): R consecutive doubles per point instead of one 8-byte access per value.
static constexpr int KF_ROWS_MIN = 8;
static bool by_rows(const kf_plan* pl) { return pl->R >= KF_ROWS_MIN; }
// workgroups of k_gather_rows / k_scatter_pot: JB = max(16, 256 / R) input indices each
static unsigned rows_grid(const kf_plan* pl) {
    const int64_t jb = pl->R >= 16 ? 16 : 256 / pl->R;
    return (unsigned)((pl->N + jb - 1) / jb);
}

// which: 0 = P2M, 1 = P2P over the leaves [lo, hi), 2 = L2P (out: the potentials of the batch)
template <int KIND, int RC>
static void launch_valu(kf_plan* pl, int which, const RhsOut* out, hipStream_t s, int lo, int hi) {
    const KerT<KIND> ker = make_ker<KIND>(pl->ker);
    const unsigned ny = (unsigned)(pl->R / RC);
    if (which == 0 && pl->n_p2m) {
        auto p2m = pl->stored ? k_p2m<KIND, RC, true> : k_p2m<KIND, RC, false>;
        p2m<<<dim3(pl->n_p2m, ny), 64, 0, s>>>(pl->p2m.p, pl->rid.p, pl->bop.p, pl->cx.p, pl->cy.p, pl->bot.p,
                                               pl->top.p, pl->xs, pl->ys, pl->tau.p, pl->surf.p, pl->p, pl->P,
                                               pl->lgR, ker, pl->p2mM_off.p, pl->p2mM.p, pl->U.p);
    }
    if (which == 1 && hi > lo) {
        const int n = hi - lo;
        const dim3 grid(pl->tune.near_grid > 0 ? std::min(pl->tune.near_grid, n) : n, ny);
        if (pl->stored)
            k_p2p_stored<RC><<<grid, 64, 0, s>>>(pl->leaves.p + lo, n, pl->leaf_coll.p + (size_t)lo * 9, pl->bot.p,
                                                 pl->top.p, pl->tau.p, pl->lgR, pl->nearM_off.p + lo, pl->nearM.p,
                                                 pl->nearb.p);
        else
            k_p2p<KIND, RC><<<grid, 64, 0, s>>>(pl->leaves.p + lo, n, pl->leaf_coll.p + (size_t)lo * 9, pl->bot.p,
                                                pl->top.p, pl->xs, pl->ys, pl->tau.p, pl->lgR, ker, pl->nearb.p);
    }
    auto l2p = pl->stored ? k_l2p<KIND, RC, true> : k_l2p<KIND, RC, false>;
    if (which == 2) pl->l2p_func = (void*)l2p;
    if (which == 2 && pl->n_leaves)
        l2p<<<dim3(pl->n_leaves, ny), 64, 0, s>>>(pl->leaves.p, pl->rid.p, pl->lvl.p, pl->bop.p, pl->cx.p, pl->cy.p,
                                                  pl->bot.p, pl->top.p, pl->xs, pl->ys, pl->perm, pl->surf.p,
                                                  pl->Psi.p, pl->nearb.p, pl->p, pl->P, pl->lgR, pl->nvalid,
                                                  by_rows(pl) ? 1 : 0, ker, pl->l2pM_off.p, pl->l2pM.p, *out);
}

// Batches of 16, 32 or 64 right-hand sides whose kernel values are formed on the fly: near
// field, P2M and L2P on the FP64 MFMA (k_p2m_mfma, k_p2p_mfma, k_l2p_mfma; the L2P in the rows
// form, which all such widths use) unless batch_form asks for the vector forms.
static bool use_mfma(const kf_plan* pl) {
    return pl->R >= 16 && !pl->stored && !pl->cplx && pl->tune.batch_form != 1;
}

static void launch_mfma(kf_plan* pl, int which, hipStream_t s, int lo, int hi) {
    const BatchMfmaArgs a{pl->p2m.p, pl->rid.p,    pl->bop.p, pl->bot.p, pl->top.p, pl->lvl.p, pl->leaves.p,
                          pl->leaf_coll.p, pl->cx.p, pl->cy.p, pl->xs, pl->ys, pl->tau.p, pl->surf.p,
                          pl->Psi.p,   pl->U.p,     pl->nearb.p, (int)pl->n_p2m, (int)pl->n_leaves, pl->p, pl->P,
                          pl->R};
    if (which == 2) pl->l2p_func = nullptr;  // no output addresses in its arguments
    launch_batch_mfma(pl->ker, a, which, lo, hi, pl->tune.near_grid, s);
}

// P2M, P2P and L2P are FP64 in both precision modes.  P2M feeds the check-to-equivalent
// solve, which amplifies rounding in the high Fourier modes by up to 1/lambda_min (~1e9 at
// p = 64); FP32 M2L leaves ~1e-7 noise in the downward check potentials that the solve
// amplifies into Psi, so an FP32 L2P on that Psi would lose ~1e-4 (measured on the oracle).
template <int KIND>
static void valu_kind(kf_plan* pl, int which, const RhsOut* out, hipStream_t s, int lo, int hi) {
    switch (rhs_chunk(pl->R)) {
        case 1: launch_valu<KIND, 1>(pl, which, out, s, lo, hi); break;
        case 2: launch_valu<KIND, 2>(pl, which, out, s, lo, hi); break;
        case 4: launch_valu<KIND, 4>(pl, which, out, s, lo, hi); break;
        case 8: launch_valu<KIND, 8>(pl, which, out, s, lo, hi); break;
        default: launch_valu<KIND, 16>(pl, which, out, s, lo, hi); break;
    }
}

// The point passes of a complex kernel (cplx.hip); L2P writes the complex potentials.
static CplxArgs cplx_args(const kf_plan* pl) {
    return CplxArgs{pl->p2m.p,  pl->rid.p, pl->bop.p, pl->bot.p,  pl->top.p,  pl->lvl.p,  pl->leaves.p,
                    pl->leaf_coll.p, pl->perm, pl->cx.p,  pl->cy.p,   pl->xs,     pl->ys,     pl->tau.p,
                    pl->surf.p, pl->Psi.p, pl->U.p,   pl->nearb.p, (int)pl->n_p2m, (int)pl->n_leaves, pl->p,
                    pl->P,      pl->R,     pl->nvalid};
}

static void run_valu(kf_plan* pl, int which, const RhsOut* out, hipStream_t s, int lo = 0, int hi = 0) {
    if (pl->cplx) return launch_cplx(pl->ker, cplx_args(pl), which, lo, hi, out ? out->p : nullptr, s);
    if (use_mfma(pl)) return launch_mfma(pl, which, s, lo, hi);
    if (pl->ker.kind == KF_LAPLACE) valu_kind<KF_LAPLACE>(pl, which, out, s, lo, hi);
    else if (pl->ker.kind == KF_TABLE) valu_kind<KF_TABLE>(pl, which, out, s, lo, hi);
    else valu_kind<KF_MODHELM>(pl, which, out, s, lo, hi);
}

// Near field of the leaves [lo, hi), bracketed by profiling marks m0, m1.
// The symmetric near field (near_sym.h) for one right-hand side, where it applies: by default
// for tabulated kernels, whose evaluations are the costly part (1M points, one right-hand side:
// Laplace as a table 1.566 -> 1.515 ms, K0 as a table 1.985 -> 1.806 ms), not for the built-in
// kernels, whose near field it slows beside the upward pass (Laplace 0.957 -> 1.017 ms, K0
// 1.289 -> 1.319 ms); near_form 2 asks for it everywhere, 1 never.
static bool use_sym(const kf_plan* pl) {
    if (!pl->sym_ok || pl->R != 1 || pl->stored || pl->cplx) return false;
    return pl->tune.near_form == 2 || (pl->tune.near_form == 0 && pl->ker.kind == KF_TABLE);
}

static void run_p2p(kf_plan* pl, hipStream_t s, int lo, int hi, int m0, int m1) {
    mark(pl, m0, s);
    if (use_sym(pl)) {  // the slots are summed after the last range (a pair's owner may be in either)
        const NearSymArgs a{pl->leaves.p, pl->leaf_coll.p, pl->pinfo.p, pl->bot.p, pl->top.p, pl->xs, pl->ys,
                            pl->tau.p, pl->nearb.p, pl->N, pl->n_leaves};
        launch_near_sym(pl->ker, a, lo, hi, hi == pl->n_leaves, pl->tune.near_grid, s);
    } else {
        run_valu(pl, 1, nullptr, s, lo, hi);
    }
    mark(pl, m1, s);
}

// Leaves whose near field an overlapped evaluation defers until the M2L has finished (the
// tail of the leaf list): near_split per mille, or by default 10% for batches of up to 4
// right-hand sides when the near field starts before the upward pass (then the downward pass
// leaves CUs idle: 1M clustered points, one right-hand side, per-level launches 0.970 ->
// 0.946 ms FP64 and 0.834 -> 0.812 ms split precision; persistent chains 0.88 -> 0.85 ms, the
// autotuner's choice); none with the late fork or for wider batches.
static constexpr int KF_NEAR_SPLIT_DEFAULT = 100;
static int near_deferred(const kf_plan* pl) {
    int pm = pl->tune.near_split;
    if (pm == 0) pm = fork_point(pl) != KF_FORK_AFTER_UP && pl->R <= 4 ? KF_NEAR_SPLIT_DEFAULT : 0;
    return pm < 0 ? 0 : (int)((int64_t)pl->n_leaves * pm / 1000);
}

// Upward and downward passes for a batch of nvalid densities in.p[r][0..n) (the batch width R
// must be set).  With overlap_near, the near field of the same densities is launched on the
// plan's side stream right after the gather; self_batch() joins it.
static void expand_batch(kf_plan* pl, const RhsIn& in, int nvalid, hipStream_t s, bool overlap_near) {
    if (fork_point(pl) == KF_FORK_NONE) overlap_near = false;
    const int64_t tot = pl->N << pl->lgR;
    pl->nvalid = nvalid;
    begin_profile(pl);
    mark(pl, 0, s);
    const bool chain = use_chain(pl);
    if (chain) k_chain_begin<<<1, 64, 0, s>>>(pl->chain_ctl.p);  // new epoch, empty queues
    if (by_rows(pl))
        k_gather_rows<<<rows_grid(pl), 256, 0, s>>>(in, nvalid, pl->iperm.p, pl->N, pl->lgR, pl->tau.p);
    else
        k_gather_tau<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(in, nvalid, pl->perm, pl->N, pl->lgR, pl->tau.p);
    mark(pl, 1, s);
    const int n_defer = overlap_near ? near_deferred(pl) : 0, n_first = pl->n_leaves - n_defer;
    if (pl->profile && pl->cur >= 0) {
        pl->ev_split[pl->cur] = n_defer > 0;
        pl->ev_serial[pl->cur] = !overlap_near;
    }
    auto fork_near = [&] {
        KF_HIP_CHECK(hipEventRecord(pl->ev_fork, s));
        KF_HIP_CHECK(hipStreamWaitEvent(pl->side, pl->ev_fork, 0));
        run_p2p(pl, pl->side, 0, n_first, 6, 7);
        KF_HIP_CHECK(hipEventRecord(pl->ev_join, pl->side));
    };
    const int fork = fork_point(pl);
    if (overlap_near && fork == KF_FORK_WITH_P2M) fork_near();
    run_valu(pl, 0, nullptr, s);
    mark(pl, 2, s);
    if (overlap_near && fork == KF_FORK_AFTER_P2M) fork_near();
    // the FP16-split M2L has no per-level launch form (it runs as one launch)
    const bool pipe = pl->tune.m2l_pipeline != 0 && pl->L > 2 && !chain && pl->prec != KF_PREC_SPLIT16;
    if (pipe) {
        if (!pl->m2ls) KF_HIP_CHECK(hipStreamCreateWithFlags(&pl->m2ls, hipStreamNonBlocking));
        while ((int)pl->ev_up.size() < pl->L) {
            hipEvent_t a, b;
            KF_HIP_CHECK(hipEventCreateWithFlags(&a, hipEventDisableTiming));
            KF_HIP_CHECK(hipEventCreateWithFlags(&b, hipEventDisableTiming));
            pl->ev_up.push_back(a);
            pl->ev_m2l.push_back(b);
        }
    }
    // bottom-up; the trees of a forest share each level's launches (one per root-width class).
    // Pipelined: the M2L of level l needs only level l's upward expansions, so it runs on its own
    // stream while the upward pass climbs.
    if (chain) {
        switch (pl->NB) {
            case 1: launch_up_chain<1>(pl, s); break;
            case 2: launch_up_chain<2>(pl, s); break;
            case 3: launch_up_chain<3>(pl, s); break;
            default: launch_up_chain<4>(pl, s); break;
        }
    }
    for (int l = chain ? 0 : pl->L - 1; l >= 1; --l) {
        for (int c = 0; c < pl->n_classes; ++c) {
            const int op = c * pl->L + l;
            switch (pl->NB) {
                case 1: launch_up<1>(pl, op, s); break;
                case 2: launch_up<2>(pl, op, s); break;
                case 3: launch_up<3>(pl, op, s); break;
                default: launch_up<4>(pl, op, s); break;
            }
        }
        if (pipe && l >= 2) {
            KF_HIP_CHECK(hipEventRecord(pl->ev_up[l], s));
            KF_HIP_CHECK(hipStreamWaitEvent(pl->m2ls, pl->ev_up[l], 0));
            for (int c = 0; c < pl->n_classes; ++c) {
                const int op = c * pl->L + l;
                switch (pl->NB) {
                    case 1: launch_m2l_op<1>(pl, op, pl->m2ls); break;
                    case 2: launch_m2l_op<2>(pl, op, pl->m2ls); break;
                    case 3: launch_m2l_op<3>(pl, op, pl->m2ls); break;
                    default: launch_m2l_op<4>(pl, op, pl->m2ls); break;
                }
            }
            KF_HIP_CHECK(hipEventRecord(pl->ev_m2l[l], pl->m2ls));
        }
    }
    mark(pl, 3, s);
    if (overlap_near && fork == KF_FORK_AFTER_UP) fork_near();
    // M2L of all levels writes the downward check potentials D into the U slab (U is free
    // once the upward pass is done); then L2L + C2E level by level
    if (!pipe) {
        switch (pl->NB) {
            case 1: launch_m2l<1>(pl, s); break;
            case 2: launch_m2l<2>(pl, s); break;
            case 3: launch_m2l<3>(pl, s); break;
            default: launch_m2l<4>(pl, s); break;
        }
    }
    mark(pl, 8, s);  // the M2L launch's share of the downward pass (none when pipelined)
    if (n_defer > 0) {
        // the deferred share of the near field waits for the M2L (throughput-bound: the
        // near field gains nothing beside it) and overlaps the latency-bound downward chain
        if (pipe) {
            KF_HIP_CHECK(hipStreamWaitEvent(pl->side, pl->ev_m2l[2], 0));  // level 2's M2L is issued last
        } else {
            KF_HIP_CHECK(hipEventRecord(pl->ev_defer, s));
            KF_HIP_CHECK(hipStreamWaitEvent(pl->side, pl->ev_defer, 0));
        }
        run_p2p(pl, pl->side, n_first, pl->n_leaves, 9, 10);
        KF_HIP_CHECK(hipEventRecord(pl->ev_join, pl->side));
    }
    if (chain) {
        switch (pl->NB) {
            case 1: launch_down_chain<1>(pl, s); break;
            case 2: launch_down_chain<2>(pl, s); break;
            case 3: launch_down_chain<3>(pl, s); break;
            default: launch_down_chain<4>(pl, s); break;
        }
    }
    for (int l = 2; !chain && l < pl->L; ++l) {
        if (pipe) KF_HIP_CHECK(hipStreamWaitEvent(s, pl->ev_m2l[l], 0));
        for (int c = 0; c < pl->n_classes; ++c) {
            const int op = c * pl->L + l;
            switch (pl->NB) {
                case 1: launch_down<1>(pl, op, s); break;
                case 2: launch_down<2>(pl, op, s); break;
                case 3: launch_down<3>(pl, op, s); break;
                default: launch_down<4>(pl, op, s); break;
            }
        }
    }
    mark(pl, 4, s);
    KF_HIP_CHECK(hipGetLastError());
    pl->expanded = pl->mode;
    pl->near_pending = overlap_near;
}

// Far field + near field at the sources for the batch, input order, into out.p[r].
static void self_batch(kf_plan* pl, const RhsOut& out, hipStream_t s) {
    if (pl->near_pending) {
        KF_HIP_CHECK(hipStreamWaitEvent(s, pl->ev_join, 0));
    } else {
        run_p2p(pl, s, 0, pl->n_leaves, 6, 7);
        mark(pl, 11, s);  // the L2P pass starts after the near field
    }
    pl->near_pending = false;
    run_valu(pl, 2, &out, s);
    if (by_rows(pl) && !pl->cplx)  // a complex kernel's L2P writes the potentials itself
        k_scatter_pot<<<rows_grid(pl), 256, 0, s>>>(out, pl->nvalid, pl->iperm.p, pl->N, pl->lgR, pl->nearb.p);
    mark(pl, 5, s);
    KF_HIP_CHECK(hipGetLastError());
}

static int batch_width(int n) {
    int R = 1;
    while (R < n && R < 64) R <<= 1;
    return R;
}

static RhsIn strided_in(const double* base, int64_t ld, int n) {
    RhsIn in{};
    for (int r = 0; r < n; ++r) in.p[r] = base + (size_t)r * ld;
    return in;
}

static RhsOut strided_out(double* base, int64_t ld, int n) {
    RhsOut out{};
    for (int r = 0; r < n; ++r) out.p[r] = base + (size_t)r * ld;
    return out;
}

static void finish_profile(kf_plan* pl) {
    // the self pass ends the evaluation; events are read lazily (no synchronisation here)
    pl->cur = -1;
}

static void drain_profile(kf_plan* pl) {
    while (pl->n_done < pl->n_rec) accumulate(pl, (int)(pl->n_done++ % kf_plan::RING));
}

// Sets the arguments of a captured kernel node: each value is converted to the kernel's own
// parameter type, so the node receives exactly what a launch with the same expressions passes.
template <class... P, class... A>
static void set_node_args(void (*)(P...), hipGraphExec_t exec, hipGraphNode_t node, const hipKernelNodeParams& base,
                          A&&... a) {
    std::tuple<std::decay_t<P>...> vals(std::forward<A>(a)...);
    void* ptrs[sizeof...(P)];
    std::apply([&](auto&... v) {
        size_t i = 0;
        ((ptrs[i++] = (void*)&v), ...);
    }, vals);
    hipKernelNodeParams p = base;
    p.kernelParams = ptrs;
    p.extra = nullptr;
    KF_HIP_CHECK(hipGraphExecKernelNodeSetParams(exec, node, &p));
}

template <int KIND, int RC>
static void set_l2p_args(kf_plan* pl, GraphSet& G, const RhsOut& out) {
    set_node_args(pl->stored ? k_l2p<KIND, RC, true> : k_l2p<KIND, RC, false>, G.exec, G.l2p, G.lp, pl->leaves.p,
                  pl->rid.p, pl->lvl.p, pl->bop.p, pl->cx.p, pl->cy.p, pl->bot.p, pl->top.p, pl->xs, pl->ys, pl->perm,
                  pl->surf.p, pl->Psi.p, pl->nearb.p, pl->p, pl->P, pl->lgR, pl->nvalid, by_rows(pl) ? 1 : 0,
                  make_ker<KIND>(pl->ker), (const int64_t*)pl->l2pM_off.p, (const double*)pl->l2pM.p, out);
}

template <int KIND>
static void set_l2p_kind(kf_plan* pl, GraphSet& G, const RhsOut& out) {
    switch (rhs_chunk(pl->R)) {
        case 1: set_l2p_args<KIND, 1>(pl, G, out); break;
        case 2: set_l2p_args<KIND, 2>(pl, G, out); break;
        case 4: set_l2p_args<KIND, 4>(pl, G, out); break;
        case 8: set_l2p_args<KIND, 8>(pl, G, out); break;
        default: set_l2p_args<KIND, 16>(pl, G, out); break;
    }
}

// Captures the evaluation of the current width (expand + self, near field on the side stream
// as a parallel branch) on the plan's capture stream and finds the gather and L2P nodes.
static void capture_width(kf_plan* pl, GraphSet& G, const RhsIn& in, const RhsOut& out, int nvalid) {
    if (!pl->cap) KF_HIP_CHECK(hipStreamCreateWithFlags(&pl->cap, hipStreamNonBlocking));
    KF_HIP_CHECK(hipStreamBeginCapture(pl->cap, hipStreamCaptureModeThreadLocal));
    pl->capturing = true;
    try {
        expand_batch(pl, in, nvalid, pl->cap, true);
        self_batch(pl, out, pl->cap);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(pl->cap, &g);
        if (g) (void)hipGraphDestroy(g);
        pl->capturing = false;
        throw;
    }
    pl->capturing = false;
    KF_HIP_CHECK(hipStreamEndCapture(pl->cap, &G.graph));
    size_t nn = 0;
    KF_HIP_CHECK(hipGraphGetNodes(G.graph, nullptr, &nn));
    std::vector<hipGraphNode_t> nodes(nn);
    KF_HIP_CHECK(hipGraphGetNodes(G.graph, nodes.data(), &nn));
    for (hipGraphNode_t nd : nodes) {
        hipGraphNodeType t;
        KF_HIP_CHECK(hipGraphNodeGetType(nd, &t));
        if (t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams kp{};
        KF_HIP_CHECK(hipGraphKernelNodeGetParams(nd, &kp));
        if (kp.func == (void*)k_gather_tau || kp.func == (void*)k_gather_rows) {
            G.gather = nd;
            G.gp = kp;
        } else if (kp.func == (void*)k_scatter_pot) {
            G.scatter = nd;
            G.sp = kp;
        } else if (kp.func == pl->l2p_func) {
            G.l2p = nd;
            G.lp = kp;
        }
    }
    KF_REQUIRE(G.gather && (G.l2p || !pl->n_leaves || !pl->l2p_func), KF_ERR_STATE,
               "captured evaluation lacks its gather / L2P node");
    KF_HIP_CHECK(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
    KF_HIP_CHECK(hipEventCreateWithFlags(&G.done, hipEventDisableTiming));
}

// One evaluation of the current width by graph replay: new density / potential addresses in
// the gather and L2P nodes, then one launch.  A replay still in flight is waited for before its
// nodes change.
static void eval_graph(kf_plan* pl, const RhsIn& in, const RhsOut& out, int nvalid, hipStream_t s) {
    GraphSet& G = pl->gs[pl->lgR];
    if (!G.exec) capture_width(pl, G, in, out, nvalid);
    if (G.launched) KF_HIP_CHECK(hipEventSynchronize(G.done));
    pl->nvalid = nvalid;
    if (by_rows(pl))
        set_node_args(k_gather_rows, G.exec, G.gather, G.gp, in, nvalid, pl->iperm.p, pl->N, pl->lgR, pl->tau.p);
    else
        set_node_args(k_gather_tau, G.exec, G.gather, G.gp, in, nvalid, pl->perm, pl->N, pl->lgR, pl->tau.p);
    if (G.scatter)
        set_node_args(k_scatter_pot, G.exec, G.scatter, G.sp, out, nvalid, pl->iperm.p, pl->N, pl->lgR, pl->nearb.p);
    if (G.l2p) {
        if (pl->ker.kind == KF_LAPLACE) set_l2p_kind<KF_LAPLACE>(pl, G, out);
        else if (pl->ker.kind == KF_TABLE) set_l2p_kind<KF_TABLE>(pl, G, out);
        else set_l2p_kind<KF_MODHELM>(pl, G, out);
    }
    KF_HIP_CHECK(hipGraphLaunch(G.exec, s));
    KF_HIP_CHECK(hipEventRecord(G.done, s));
    G.launched = true;
    pl->expanded = KF_TGT_SOURCES;
    pl->near_pending = false;
}

// Source-target evaluation of one batch (nvalid <= 64 right-hand sides): graph replay when
// enabled (not while profiling, and not on a workspace), launches otherwise.
static void eval_batch(kf_plan* pl, const RhsIn& in, const RhsOut& out, int nvalid, hipStream_t s) {
    set_batch(pl, batch_width(nvalid), KF_TGT_SOURCES, s);
    if (pl->use_graphs && !pl->profile && !pl->ws) {
        eval_graph(pl, in, out, nvalid, s);
    } else {
        expand_batch(pl, in, nvalid, s, true);
        self_batch(pl, out, s);
    }
}

}  // namespace kf

extern "C" {

int kf_plan_create(const kf_tree* tree, const kf_kernel_desc* kernel, int32_t p, int32_t precision,
                   void* stream, kf_plan** out) {
    return kf_forest_create(&tree, 1, kernel, p, precision, stream, out);
}

int kf_forest_create(const kf_tree* const* trees, int32_t n_trees, const kf_kernel_desc* kernel, int32_t p,
                     int32_t precision, void* stream, kf_plan** out) {
    return guarded([&] {
        KF_REQUIRE(out, KF_ERR_ARG, "null output handle");
        *out = nullptr;
        KF_REQUIRE(n_trees >= 1, KF_ERR_ARG, "a forest needs at least one tree");
        kf_plan* pl = new kf_plan();
        try {
            build_plan(pl, trees, n_trees, kernel, p, precision, (hipStream_t)stream);
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

int kf_plan_verify(const kf_plan* pl, void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl, KF_ERR_ARG, "null plan");
        hipStream_t s = (hipStream_t)stream;
        std::string e, eo;
        const char* stage = "work lists";
        try {
            e = compare_with_host_lists(pl->trees, pl->p, pl->P, pl->dl, pl->lists, s);
            stage = "operators";
            if (e.empty()) eo = compare_ops_with_host(pl, s);
        } catch (const std::bad_alloc&) {
            const PlanLists& l = pl->lists;
            throw KfError{KF_ERR_ALLOC, std::string("kf_plan_verify: host allocation failed comparing the ") + stage +
                                            " (boxes " + std::to_string(l.box_count) + ", upward " +
                                            std::to_string(l.up_count) + ", downward " + std::to_string(l.dn_count) +
                                            ", leaves " + std::to_string(l.leaf_count) + ", p2m " +
                                            std::to_string(l.p2m_count) + ", ops " + std::to_string(pl->ops.n) + ")"};
        }
        KF_REQUIRE(e.empty(), KF_ERR_STATE, "device work lists differ from the host builder's: " + e);
        KF_REQUIRE(eo.empty(), KF_ERR_STATE, "device operators differ from the host's: " + eo);
        int ctl[4] = {0, 0, 0, 0};
        KF_HIP_CHECK(hipMemcpy(ctl, pl->chain_ctl.p, sizeof(ctl), hipMemcpyDeviceToHost));
        KF_REQUIRE(ctl[3] == 0, KF_ERR_STATE, "a persistent chain's bounded wait gave up in an earlier evaluation");
    });
}

int kf_plan_expand(kf_plan* pl, const double* d_tau, int64_t ld_tau, int32_t nrhs, void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl && d_tau, KF_ERR_ARG, "null argument");
        KF_REQUIRE(nrhs >= 1 && nrhs <= 64, KF_ERR_ARG, "kf_plan_expand keeps 1..64 right-hand sides");
        KF_REQUIRE(nrhs == 1 || ld_tau >= pl->N, KF_ERR_ARG, "leading dimension < n");
        hipStream_t s = (hipStream_t)stream;
        set_batch(pl, batch_width(nrhs), KF_TGT_ALL, s);
        expand_batch(pl, strided_in(d_tau, ld_tau, nrhs), nrhs, s, false);
    });
}

int kf_plan_self(kf_plan* pl, double* d_pot, int64_t ld_pot, void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl && d_pot, KF_ERR_ARG, "null argument");
        KF_REQUIRE(pl->expanded >= 0, KF_ERR_STATE, "kf_plan_expand must run first");
        KF_REQUIRE(!pl->ws || pl->ws->last == pl, KF_ERR_STATE,
                   "another plan on the workspace has evaluated since kf_plan_expand");
        KF_REQUIRE(pl->nvalid == 1 || ld_pot >= pl->N, KF_ERR_ARG, "leading dimension < n");
        if (pl->nvalid == 1) ld_pot = pl->N;
        self_batch(pl, strided_out(d_pot, pl->cplx ? 2 * ld_pot : ld_pot, pl->nvalid), (hipStream_t)stream);
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
        if (pl->cplx) ld_pot *= 2;  // complex potentials: interleaved (re, im)
        hipStream_t s = (hipStream_t)stream;
        for (int r0 = 0; r0 < nrhs; r0 += 64) {
            const int nb = std::min(64, nrhs - r0);
            eval_batch(pl, strided_in(d_tau + (size_t)r0 * ld_tau, ld_tau, nb),
                       strided_out(d_pot + (size_t)r0 * ld_pot, ld_pot, nb), nb, s);
        }
        finish_profile(pl);
    });
}

int kf_plan_eval_ptrs(kf_plan* pl, const double* const* d_taus, double* const* d_pots, int32_t nrhs,
                      void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl && d_taus && d_pots, KF_ERR_ARG, "null argument");
        KF_REQUIRE(nrhs >= 1, KF_ERR_ARG, "nrhs must be >= 1");
        for (int r = 0; r < nrhs; ++r)
            KF_REQUIRE(d_taus[r] && d_pots[r], KF_ERR_ARG, "null density or potential pointer");
        hipStream_t s = (hipStream_t)stream;
        for (int r0 = 0; r0 < nrhs; r0 += 64) {
            const int nb = std::min(64, nrhs - r0);
            RhsIn in{};
            RhsOut out{};
            for (int r = 0; r < nb; ++r) {
                in.p[r] = d_taus[r0 + r];
                out.p[r] = d_pots[r0 + r];
            }
            eval_batch(pl, in, out, nb, s);
        }
        finish_profile(pl);
    });
}

int kf_plan_get_tuning(const kf_plan* pl, kf_tuning* t) {
    return guarded([&] {
        KF_REQUIRE(pl && t, KF_ERR_ARG, "null argument");
        *t = pl->tune;
    });
}

int kf_plan_set_tuning(kf_plan* pl, const kf_tuning* t) {
    return guarded([&] {
        KF_REQUIRE(pl && t, KF_ERR_ARG, "null argument");
        KF_REQUIRE(t->near_grid >= 0 && t->near_fork >= 0 && t->near_fork <= 4 &&
                       (t->m2l_pipeline == 0 || t->m2l_pipeline == 1) &&
                       (t->level_rows == 0 || t->level_rows == 16 || t->level_rows == 32 || t->level_rows == 64 ||
                        t->level_rows == 128) &&
                       t->near_split >= -1 && t->near_split <= 1000 && t->chain >= 0 && t->chain <= 2 &&
                       t->batch_form >= 0 && t->batch_form <= 2 && t->near_form >= 0 && t->near_form <= 2,
                   KF_ERR_ARG, "invalid tuning");
        pl->drop_graphs();
        pl->tune = *t;
    });
}

int kf_plan_autotune(kf_plan* pl, int32_t nrhs, void* stream, kf_tuning* chosen, double* best_ms) {
    return guarded([&] {
        KF_REQUIRE(pl, KF_ERR_ARG, "null plan");
        KF_REQUIRE(nrhs >= 1 && nrhs <= 64, KF_ERR_ARG, "kf_plan_autotune takes 1..64 right-hand sides");
        hipStream_t s = (hipStream_t)stream;
        int dev = 0, cus = 256;
        KF_HIP_CHECK(hipGetDevice(&dev));
        KF_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        DevBuf<double> tin, tout;
        const int64_t ldo = pl->cplx ? 2 * pl->N : pl->N;  // complex potentials: (re, im) pairs
        tin.alloc((size_t)pl->N * nrhs);
        tout.alloc((size_t)ldo * nrhs);
        tin.zero(s);
        const RhsIn in = strided_in(tin.p, pl->N, nrhs);
        const RhsOut out = strided_out(tout.p, ldo, nrhs);
        hipEvent_t e0, e1;
        KF_HIP_CHECK(hipEventCreate(&e0));
        KF_HIP_CHECK(hipEventCreate(&e1));
        const bool graphs = pl->use_graphs, prof = pl->profile;
        pl->use_graphs = false;
        pl->profile = false;
        const kf_tuning before = pl->tune;
        kf_tuning best = before;
        double best_t = 1e300;
        try {
            std::vector<int> grids = {0};
            for (int k : {4, 8, 16})
                if (k * cus < pl->n_leaves) grids.push_back(k * cus);
            // the near-field split pays only where the near field starts with the upward pass
            std::vector<kf_tuning> cand;
            for (int chain : {1, 2})
                for (int fork = 1; fork <= 4; ++fork)
                    for (int pipe = 0; pipe < (chain == 2 ? 1 : 2); ++pipe)
                        for (int g : grids)
                            for (int split : {-1, 100})
                                if (split < 0 || fork != 1)
                                    cand.push_back(
                                        kf_tuning{g, fork, pipe, before.level_rows, split, chain, before.batch_form,
                                                  before.near_form});
            // stage 1: launch structure; stage 2 (after it): the tile shape of the level launches
            // (rows per workgroup: the size rule, or 16, 32 or 64 everywhere) under the best one
            const size_t n1 = cand.size();
            for (size_t ci = 0; ci < cand.size(); ++ci) {
                if (ci == n1)
                    for (int rows : {0, 16, 32, 64, 128}) {
                        kf_tuning t = best;
                        t.level_rows = rows;
                        if (rows != best.level_rows) cand.push_back(t);
                    }
                if (ci >= cand.size()) break;
                const kf_tuning c = cand[ci];
                pl->tune = c;
                // warm-up, then the better of two timings of >= ~10 ms of evaluations each
                KF_HIP_CHECK(hipEventRecord(e0, s));
                eval_batch(pl, in, out, nrhs, s);
                KF_HIP_CHECK(hipEventRecord(e1, s));
                KF_HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.f;
                KF_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                const int reps = std::max(3, std::min(200, (int)(10.0 / std::max(ms, 0.01f))));
                double t = 1e300;
                for (int round = 0; round < 2; ++round) {
                    KF_HIP_CHECK(hipEventRecord(e0, s));
                    for (int rep = 0; rep < reps; ++rep) eval_batch(pl, in, out, nrhs, s);
                    KF_HIP_CHECK(hipEventRecord(e1, s));
                    KF_HIP_CHECK(hipEventSynchronize(e1));
                    KF_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                    t = std::min(t, (double)ms / reps);
                }
                if (t < best_t) {
                    best_t = t;
                    best = pl->tune;
                }
            }
        } catch (...) {
            pl->tune = before;
            pl->use_graphs = graphs;
            pl->profile = prof;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        pl->tune = best;
        pl->use_graphs = graphs;
        pl->profile = prof;
        pl->drop_graphs();
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        KF_HIP_CHECK(hipStreamSynchronize(s));
        if (chosen) *chosen = best;
        if (best_ms) *best_ms = best_t;
    });
}

int kf_plan_set_graphs(kf_plan* pl, int32_t enable) {
    return guarded([&] {
        KF_REQUIRE(pl, KF_ERR_ARG, "null plan");
        KF_REQUIRE(!enable || !pl->cplx, KF_ERR_ARG, "graph replay takes real kernels");
        if (!enable) pl->drop_graphs();
        pl->use_graphs = enable != 0;
    });
}

int kf_plan_reserve(kf_plan* pl, int32_t nrhs, void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl, KF_ERR_ARG, "null plan");
        KF_REQUIRE(nrhs >= 1 && nrhs <= 64, KF_ERR_ARG, "kf_plan_reserve takes 1..64 right-hand sides");
        hipStream_t s = (hipStream_t)stream;
        reserve_batch(pl, batch_width(nrhs), s);
        KF_HIP_CHECK(hipStreamSynchronize(s));
    });
}

int kf_plan_store_matrices(kf_plan* pl, int32_t on, void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl, KF_ERR_ARG, "null plan");
        KF_REQUIRE(!on || !pl->cplx, KF_ERR_ARG, "stored kernel matrices take real kernels");
        hipStream_t s = (hipStream_t)stream;
        if (on != 0 && pl->stored) return;
        pl->drop_graphs();  // captured evaluations hold the launches that change
        KF_HIP_CHECK(hipStreamSynchronize(s));
        pl->stored = false;
        auto release = [pl] {
            for (DevBuf<double>* m : {&pl->nearM, &pl->p2mM, &pl->l2pM}) m->release();
            for (DevBuf<int64_t>* o : {&pl->nearM_off, &pl->p2mM_off, &pl->l2pM_off}) o->release();
        };
        release();
        if (on == 0) return;
        // a failure part way (an allocation) leaves the plan matrix-free with nothing held
        struct Guard {
            decltype(release)& r;
            bool armed = true;
            ~Guard() {
                if (armed) r();
            }
        } guard{release};
        // host copies of the lists the block offsets depend on
        PlanLists& ls = pl->lists;
        auto get = [&](std::vector<int32_t>& v, const DevBuf<int32_t>& src, int64_t n) {
            v.resize(n);
            if (n) KF_HIP_CHECK(hipMemcpyAsync(v.data(), src.p, n * 4, hipMemcpyDeviceToHost, s));
        };
        get(ls.bot, pl->bot, ls.box_count);
        get(ls.top, pl->top, ls.box_count);
        get(ls.lvl, pl->lvl, ls.box_count);
        get(ls.p2m, pl->p2m, ls.p2m_count);
        get(ls.leaves, pl->leaves, ls.leaf_count);
        get(ls.lcoll, pl->leaf_coll, 9 * ls.leaf_count);
        KF_HIP_CHECK(hipStreamSynchronize(s));
        StoredOffsets so;
        stored_offsets(ls, pl->p, so);
        const int64_t ntot = so.near.back(), ltot = so.l2p.back(), mtot = so.p2m.back();
        pl->nearM.alloc((size_t)ntot);
        pl->l2pM.alloc((size_t)ltot);
        pl->p2mM.alloc((size_t)mtot);
        pl->nearM_off.upload(so.near, s);
        pl->l2pM_off.upload(so.l2p, s);
        pl->p2mM_off.upload(so.p2m, s);
        auto fill = [&](auto kind) {
            constexpr int KIND = decltype(kind)::value;
            const KerT<KIND> ker = make_ker<KIND>(pl->ker);
            const unsigned gl = (unsigned)std::max(1, std::min(pl->n_leaves, 1 << 16));
            if (ntot)
                k_p2p_store<KIND><<<gl, 64, 0, s>>>(pl->leaves.p, pl->n_leaves, pl->leaf_coll.p, pl->bot.p, pl->top.p,
                                                    pl->xs, pl->ys, ker, pl->nearM_off.p, pl->nearM.p);
            if (ltot)
                k_l2p_store<KIND><<<gl, 64, 0, s>>>(pl->leaves.p, pl->n_leaves, pl->lvl.p, pl->bop.p, pl->cx.p, pl->cy.p,
                                                    pl->bot.p, pl->top.p, pl->xs, pl->ys, pl->surf.p, pl->p, pl->P, ker,
                                                    pl->l2pM_off.p, pl->l2pM.p);
            if (mtot)
                k_p2m_store<KIND><<<(unsigned)std::min(pl->n_p2m, 1 << 16), 64, 0, s>>>(
                    pl->p2m.p, pl->n_p2m, pl->bop.p, pl->cx.p, pl->cy.p, pl->bot.p, pl->top.p, pl->xs, pl->ys,
                    pl->surf.p, pl->p, pl->P, ker, pl->p2mM_off.p, pl->p2mM.p);
        };
        if (pl->ker.kind == KF_LAPLACE) fill(std::integral_constant<int, KF_LAPLACE>{});
        else if (pl->ker.kind == KF_TABLE) fill(std::integral_constant<int, KF_TABLE>{});
        else fill(std::integral_constant<int, KF_MODHELM>{});
        KF_HIP_CHECK(hipGetLastError());
        KF_HIP_CHECK(hipStreamSynchronize(s));  // the host offsets are released on return
        guard.armed = false;
        pl->stored = true;
    });
}

int kf_workspace_create(kf_workspace** out) {
    return guarded([&] {
        KF_REQUIRE(out, KF_ERR_ARG, "null argument");
        *out = new kf_workspace();
    });
}

int kf_workspace_destroy(kf_workspace* ws) {
    return guarded([&] {
        if (!ws) return;
        for (kf_plan* pl : ws->users) {  // back to slabs of their own, allocated on next use
            pl->drop_graphs();
            pl->U.release();
            pl->Psi.release();
            pl->X.release();
            pl->X32.release();
            pl->tau.release();
            pl->nearb.release();
            pl->ws = nullptr;
            pl->Rcap = 0;
            pl->rows_cap = 0;
            pl->R = 0;
            pl->mode = -1;
            pl->expanded = -1;
        }
        delete ws;
    });
}

int kf_workspace_fill(kf_workspace* ws, double value, void* stream) {
    return guarded([&] {
        KF_REQUIRE(ws, KF_ERR_ARG, "null workspace");
        hipStream_t s = (hipStream_t)stream;
        for (DevBuf<double>* b : {&ws->U, &ws->Psi, &ws->X, &ws->tau, &ws->nearb})
            if (b->n) k_fill_f64<<<1024, 256, 0, s>>>(b->p, b->n, value);
        if (ws->X32.n) k_fill_f32<<<1024, 256, 0, s>>>(ws->X32.p, ws->X32.n, (float)value);
        KF_HIP_CHECK(hipGetLastError());
    });
}

int kf_workspace_bytes(const kf_workspace* ws, int64_t* bytes) {
    return guarded([&] {
        KF_REQUIRE(ws && bytes, KF_ERR_ARG, "null argument");
        *bytes = (int64_t)((ws->U.n + ws->Psi.n + ws->X.n + ws->tau.n + ws->nearb.n) * 8 +
                           ws->X32.n * 4);
    });
}

int kf_plan_attach_workspace(kf_plan* pl, kf_workspace* ws) {
    return guarded([&] {
        KF_REQUIRE(pl, KF_ERR_ARG, "null plan");
        if (pl->ws == ws) return;
        pl->drop_graphs();  // captured evaluations hold the old slab addresses
        if (pl->ws) {
            auto& u = pl->ws->users;
            u.erase(std::remove(u.begin(), u.end(), pl), u.end());
            if (pl->ws->last == pl) pl->ws->last = nullptr;
        }
        pl->U.release();
        pl->Psi.release();
        pl->X.release();
        pl->X32.release();
        pl->tau.release();
        pl->nearb.release();
        pl->ws = ws;
        if (ws) ws->users.push_back(pl);
        // the next evaluation (or kf_plan_reserve) lays the slabs out again
        pl->Rcap = 0;
        pl->rows_cap = 0;
        pl->R = 0;
        pl->mode = -1;
        pl->expanded = -1;
    });
}

int kf_plan_slab_bytes(const kf_plan* pl, int32_t nrhs, int64_t* bytes) {
    return guarded([&] {
        KF_REQUIRE(pl && bytes, KF_ERR_ARG, "null argument");
        KF_REQUIRE(nrhs >= 1 && nrhs <= 64, KF_ERR_ARG, "kf_plan_slab_bytes takes 1..64 right-hand sides");
        *bytes = slab_bytes(pl, batch_width(nrhs), pl->n_src);
    });
}

int kf_plan_footprint(const kf_plan* pl, int32_t nrhs, int64_t* bytes) {
    return guarded([&] {
        KF_REQUIRE(pl && bytes, KF_ERR_ARG, "null argument");
        KF_REQUIRE(nrhs >= 1 && nrhs <= 64, KF_ERR_ARG, "kf_plan_footprint takes 1..64 right-hand sides");
        *bytes = footprint(pl, batch_width(nrhs));
    });
}

int kf_plan_targets(kf_plan* pl, const double* d_tx, const double* d_ty, int64_t m, int32_t check_self,
                    double* d_pot, void* stream) {
    return guarded([&] {
        KF_REQUIRE(pl && d_tx && d_ty && d_pot, KF_ERR_ARG, "null argument");
        // a target may lie in any leaf: the expansion must cover boxes without sources too
        KF_REQUIRE(pl->expanded == KF_TGT_ALL, KF_ERR_STATE, "kf_plan_expand must run first");
        KF_REQUIRE(!pl->ws || pl->ws->last == pl, KF_ERR_STATE,
                   "another plan on the workspace has evaluated since kf_plan_expand");
        KF_REQUIRE(m >= 0, KF_ERR_ARG, "negative target count");
        KF_REQUIRE(pl->trees.size() == 1, KF_ERR_STATE, "arbitrary targets need a single-tree plan");
        if (m == 0) return;
        hipStream_t s = (hipStream_t)stream;
        if ((int64_t)pl->loc.n < m) pl->loc.alloc((size_t)m);
        unsigned nb = (unsigned)((m + 255) / 256);
        const dim3 grid(nb, (unsigned)pl->nvalid);
        k_locate<<<nb, 256, 0, s>>>(d_tx, d_ty, m, pl->cx.p, pl->cy.p, pl->child0.p, pl->leafflag.p, pl->loc.p);
        if (pl->cplx) {  // complex potentials, pot[2 (r m + i) + c]
            launch_cplx_targets(pl->ker, cplx_args(pl), d_tx, d_ty, m, pl->loc.p, pl->coll.p, check_self ? 1 : 0,
                                d_pot, s);
            return;
        }
        auto targets = [&](auto kind) {
            constexpr int KIND = decltype(kind)::value;
            k_targets<KIND><<<grid, 256, 0, s>>>(d_tx, d_ty, m, pl->loc.p, pl->rid.p, pl->lvl.p, pl->bop.p, pl->coll.p,
                                                 pl->cx.p, pl->cy.p, pl->bot.p, pl->top.p, pl->xs, pl->ys, pl->tau.p,
                                                 pl->surf.p, pl->Psi.p, pl->p, pl->P, pl->lgR, make_ker<KIND>(pl->ker),
                                                 check_self ? 1 : 0, d_pot);
        };
        if (pl->ker.kind == KF_LAPLACE) targets(std::integral_constant<int, KF_LAPLACE>{});
        else if (pl->ker.kind == KF_TABLE) targets(std::integral_constant<int, KF_TABLE>{});
        else targets(std::integral_constant<int, KF_MODHELM>{});
        KF_HIP_CHECK(hipGetLastError());
    });
}

int kf_plan_profile(kf_plan* pl, int32_t enable, double* ms, int32_t cap) {
    int written = 0;
    int rc = guarded([&] {
        KF_REQUIRE(pl, KF_ERR_ARG, "null plan");
        if (pl->profile) drain_profile(pl);
        double out[kf_plan::NPASS + 2];
        for (int i = 0; i < kf_plan::NPASS; ++i) out[i] = pl->prof_ms[i];
        out[kf_plan::NPASS] = (double)pl->n_acc;
        out[kf_plan::NPASS + 1] = pl->prof_m2l;
        for (int i = 0; ms && i < cap && i < kf_plan::NPASS + 2; ++i) ms[written++] = out[i];
        if (enable)
            for (auto& set : pl->ev)
                for (auto& e : set)
                    if (!e) KF_HIP_CHECK(hipEventCreate(&e));
        pl->profile = enable != 0;
        pl->n_rec = pl->n_done = pl->n_acc = 0;
        pl->cur = -1;
        for (double& v : pl->prof_ms) v = 0.0;
        pl->prof_m2l = 0.0;
    });
    return rc == KF_OK ? written : rc;
}

void kf_plan_destroy(kf_plan* pl) { delete pl; }

}  // extern "C"
