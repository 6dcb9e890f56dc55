// plan_dev.hip -- the work lists of a planned evaluation, built on the GPU (see plan_dev.h).
//
// The lists restate the reference's pass structure (pykifmm2d/class_fmm.py:299-352,
// fmm.py:350-408) as plan_lists.cpp does on the host, from the device box structure of each
// tree (tree_dev.hip).  Boxes are visited in (tree, level, position) order, the reference's
// level order; each list is the subsequence of boxes it needs, grouped by a stable sort on a
// small key.
#include <algorithm>
#include <cstring>
#include <string>

#include "kf_internal.h"
#include "kf_scan.h"
#include "kf_sort.h"
#include "plan_dev.h"
#include "tree_dev.h"

namespace kf {

namespace {

constexpr int PT = 256;
inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + PT - 1) / PT); }

__device__ __forceinline__ int seg_of64(const int64_t* __restrict__ off, int n, int64_t i) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One tree of the plan: its device box structure (at its first box) and plan offsets.
struct TreeSrc {
    const int32_t *lvl, *ix, *iy, *bot, *top, *parent, *child0, *coll, *level_list;
    const uint8_t *leaf, *fake;
    int64_t pbox, ppt;
    double root_min, root_w;
    int cls;
};

// Plan-wide per-box arrays (global id g = pbox of its tree + local id) and the box sequence
// S in (tree, level, position) order.
struct Boxes {
    double *cx, *cy;
    int32_t *bot, *top, *lvl, *bop, *child0, *coll, *ix, *iy, *parent, *ns, *S;
    uint8_t *leaf, *fake;
};

__global__ void k_gather_boxes(int64_t B, int nt, const int64_t* __restrict__ pbox_off,
                               const TreeSrc* __restrict__ ts, int L, Boxes o) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= B) return;
    const int i = seg_of64(pbox_off, nt, g);
    const TreeSrc T = ts[i];
    const int64_t lid = g - T.pbox;
    const int l = T.lvl[lid];
    const double w = ldexp(T.root_w, -l);
    o.cx[g] = T.root_min + ((double)T.ix[lid] + 0.5) * w;
    o.cy[g] = T.root_min + ((double)T.iy[lid] + 0.5) * w;
    o.bot[g] = (int32_t)(T.bot[lid] + T.ppt);
    o.top[g] = (int32_t)(T.top[lid] + T.ppt);
    o.ns[g] = T.top[lid] - T.bot[lid];
    o.lvl[g] = l;
    o.bop[g] = T.cls * L + l;
    o.ix[g] = T.ix[lid];
    o.iy[g] = T.iy[lid];
    const int32_t c0 = T.child0[lid], pa = T.parent[lid];
    o.child0[g] = c0 >= 0 ? (int32_t)(c0 + T.pbox) : -1;
    o.parent[g] = pa >= 0 ? (int32_t)(pa + T.pbox) : -1;
    o.leaf[g] = T.leaf[lid];
    o.fake[g] = T.fake[lid];
    for (int j = 0; j < 9; ++j) {
        const int32_t c = T.coll[9 * lid + j];
        o.coll[9 * g + j] = c >= 0 ? (int32_t)(c + T.pbox) : -1;
    }
    o.S[g] = (int32_t)(T.pbox + T.level_list[lid]);
}

// Per position k of the sequence: membership flags of the lists and the group keys.
//   src   slab row in the first group (sources, level >= 1)
//   emp   slab row in the second group (real, no sources, level >= 2)
//   up    upward entry, key (level, class, part: 0 with children, 1 without)
//   dn    downward target, key (level, class, child position, 0 with sources / 1 without)
//   lf    near-field / L2P leaf (real leaf with points, level >= 1)
//   p2m   P2M box (sources, no children, level >= 1)
enum { F_SRC, F_EMP, F_UP, F_DN, F_LF, F_P2M, NF };
__global__ void k_flags(int64_t B, Boxes o, const int32_t* __restrict__ cls_of_tree, const int64_t* __restrict__ pbox_off,
                        int nt, int C, uint32_t* __restrict__ flags /* NF x B */, uint32_t* __restrict__ upkey,
                        uint32_t* __restrict__ dnkey) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= B) return;
    const int32_t b = o.S[k];
    const int l = o.lvl[b], ns = o.ns[b];
    const bool fk = o.fake[b], has = ns > 0, kids = o.child0[b] >= 0;
    const int c = cls_of_tree[seg_of64(pbox_off, nt, b)];
    const uint32_t src = l >= 1 && has, emp = l >= 2 && !fk && !has;
    const uint32_t up = l >= 1 && has, dn = l >= 2 && !fk;
    const uint32_t lf = l >= 1 && o.leaf[b] && has, p2m = l >= 1 && has && !kids;
    flags[F_SRC * B + k] = src;
    flags[F_EMP * B + k] = emp;
    flags[F_UP * B + k] = up;
    flags[F_DN * B + k] = dn;
    flags[F_LF * B + k] = lf;
    flags[F_P2M * B + k] = p2m;
    upkey[k] = (uint32_t)((l * C + c) * 2 + (kids ? 0 : 1));
    const int q = 2 * (o.ix[b] & 1) + (o.iy[b] & 1);
    dnkey[k] = (uint32_t)((((l * C + c) * 4 + q) * 2) + (has ? 0 : 1));
}

// Scatter of every list (positions from the flag scans) and of the slab rows.
__global__ void k_scatter_lists(int64_t B, Boxes o, const uint32_t* __restrict__ flags,
                                const uint32_t* __restrict__ idx /* NF x B scans */, const uint32_t* __restrict__ upkey,
                                const uint32_t* __restrict__ dnkey, const uint32_t* __restrict__ n_src_total,
                                int32_t* __restrict__ rid, int32_t* __restrict__ up_b, uint64_t* __restrict__ up_k,
                                int32_t* __restrict__ dn_b, uint64_t* __restrict__ dn_k, int32_t* __restrict__ leaves,
                                int32_t* __restrict__ p2m, uint32_t* __restrict__ up_hist,
                                uint32_t* __restrict__ dn_hist) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= B) return;
    const int32_t b = o.S[k];
    int32_t r = -1;
    if (flags[F_SRC * B + k]) r = (int32_t)idx[F_SRC * B + k];
    if (flags[F_EMP * B + k]) r = (int32_t)(*n_src_total + idx[F_EMP * B + k]);
    rid[b] = r;
    if (flags[F_UP * B + k]) {
        const uint32_t j = idx[F_UP * B + k];
        up_b[j] = b;
        up_k[j] = upkey[k];
        atomicAdd(&up_hist[upkey[k]], 1u);
    }
    if (flags[F_DN * B + k]) {
        const uint32_t j = idx[F_DN * B + k];
        dn_b[j] = b;
        dn_k[j] = dnkey[k];
        atomicAdd(&dn_hist[dnkey[k]], 1u);
    }
    if (flags[F_LF * B + k]) leaves[idx[F_LF * B + k]] = b;
    if (flags[F_P2M * B + k]) p2m[idx[F_P2M * B + k]] = b;
}

// Upward entries: the box's row and its children's rows (-1 for a child without sources).
__global__ void k_up_rows(int n, const int32_t* __restrict__ up_b, Boxes o, const int32_t* __restrict__ rid,
                          int32_t* __restrict__ urow, int32_t* __restrict__ crow) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t b = up_b[k];
    urow[k] = rid[b];
    const int32_t c0 = o.child0[b];
    for (int q = 0; q < 4; ++q) crow[4 * (int64_t)k + q] = c0 >= 0 && o.ns[c0 + q] > 0 ? rid[c0 + q] : -1;
}

// Downward targets: row, parent row and the 27 interaction-list source rows (the non-adjacent
// children of the parent's colleagues that hold sources; reference class_fmm.py:17-33).
struct VSl {
    int8_t ox[4][27], oy[4][27];
};
__global__ void k_down_rows(int n, const int32_t* __restrict__ dn_b, Boxes o, const int32_t* __restrict__ rid, VSl vs,
                            int32_t* __restrict__ trow, int32_t* __restrict__ prow, int32_t* __restrict__ vsrc,
                            unsigned long long* __restrict__ stats) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t t = dn_b[k], par = o.parent[t];
    const int l = o.lvl[t];
    const int X = o.ix[t], Y = o.iy[t];
    const int q = 2 * (X & 1) + (Y & 1);
    const int64_t side = (int64_t)1 << l;
    trow[k] = rid[t];
    prow[k] = rid[par];
    int used = 0;
    for (int sl = 0; sl < 27; ++sl) {
        const int gx = X + vs.ox[q][sl], gy = Y + vs.oy[q][sl];
        int32_t src = -1;
        if (gx >= 0 && gy >= 0 && gx < side && gy < side)
            for (int j = 0; j < 9; ++j) {
                const int32_t c = o.coll[9 * (int64_t)par + j];
                if (c < 0 || o.ix[c] != (gx >> 1) || o.iy[c] != (gy >> 1)) continue;
                const int32_t c0 = o.child0[c];
                if (c0 >= 0) {
                    const int32_t cand = c0 + 2 * (gx & 1) + (gy & 1);
                    if (o.ns[cand] > 0) src = rid[cand];
                }
                break;
            }
        vsrc[27 * (int64_t)k + sl] = src;
        used += src >= 0;
    }
    if (o.ns[t] > 0) {
        atomicAdd(&stats[0], (unsigned long long)used);  // m2l_pairs
        if (l >= 3) atomicAdd(&stats[1], 1ull);            // l2l_boxes
    }
}

// Leaves and their colleagues; near-field pairs and L2P evaluations.
__global__ void k_leaf_rows(int n, const int32_t* __restrict__ leaves, Boxes o, int p, int32_t* __restrict__ lcoll,
                            unsigned long long* __restrict__ stats) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t b = leaves[k];
    int64_t src = 0;
    for (int j = 0; j < 9; ++j) {
        const int32_t c = o.coll[9 * (int64_t)b + j];
        lcoll[9 * (int64_t)k + j] = c;
        if (c >= 0) src += o.ns[c];
    }
    const int64_t nb = o.ns[b];
    atomicAdd(&stats[2], (unsigned long long)(nb * src - nb));                // p2p_pairs
    if (o.lvl[b] >= 2) atomicAdd(&stats[3], (unsigned long long)(p * nb));    // l2p_evals
}

__global__ void k_p2m_evals(int n, const int32_t* __restrict__ p2m, Boxes o, int p, unsigned long long* __restrict__ stats) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    atomicAdd(&stats[4], (unsigned long long)p * (unsigned long long)o.ns[p2m[k]]);
}

__global__ void k_masks(const int4* __restrict__ tiles, int count, const int32_t* __restrict__ vsrc,
                        uint32_t* __restrict__ masks) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int4 T = tiles[k];
    uint32_t m = 0;
    for (int i = 0; i < T.y; ++i)
        for (int sl = 0; sl < 27; ++sl)
            if (vsrc[27 * (int64_t)(T.x + i) + sl] >= 0) m |= 1u << sl;
    masks[k] = m;
}

int bits_for(int64_t v) {
    int b = 1;
    while (((int64_t)1 << b) <= v) ++b;
    return b;
}

}  // namespace

void build_dev_lists(const std::vector<const kf_tree*>& trees, int p, int P, PlanLists& o, DevLists& d,
                     hipStream_t s) {
    const int nt = (int)trees.size();
    o = PlanLists{};
    o.n_trees = nt;
    // classes of equal root width share operators (first appearance order)
    std::vector<int> cls(nt);
    for (int i = 0; i < nt; ++i) {
        const double w = trees[i]->h.root_width;
        int c = 0;
        while (c < (int)o.class_width.size() && o.class_width[c] != w) ++c;
        if (c == (int)o.class_width.size()) o.class_width.push_back(w);
        cls[i] = c;
        o.L = std::max(o.L, trees[i]->levels());
    }
    o.n_classes = (int)o.class_width.size();
    const int L = o.L, C = o.n_classes;
    o.pt_off.assign(nt + 1, 0);
    o.box_off.assign(nt + 1, 0);
    std::vector<TreeSrc> ts(nt);
    for (int i = 0; i < nt; ++i) {
        const kf_tree* tr = trees[i];
        KF_REQUIRE(tr->dev, KF_ERR_STATE, "plan needs trees built on the GPU");
        const DevForest& F = *tr->dev;
        const int64_t b0 = F.box_off[tr->slot];
        o.pt_off[i + 1] = o.pt_off[i] + tr->h.n;
        o.box_off[i + 1] = o.box_off[i] + tr->n_boxes();
        TreeSrc& T = ts[i];
        T.lvl = F.lvl.p + b0;
        T.ix = F.ix.p + b0;
        T.iy = F.iy.p + b0;
        T.bot = F.bot.p + b0;
        T.top = F.top.p + b0;
        T.parent = F.parent.p + b0;
        T.child0 = F.child0.p + b0;
        T.coll = F.coll.p + 9 * b0;
        T.level_list = F.level_list.p + F.lev_off[(size_t)tr->slot * 32];
        T.leaf = F.leaf.p + b0;
        T.fake = F.fake.p + b0;
        T.pbox = o.box_off[i];
        T.ppt = o.pt_off[i];
        T.root_min = tr->h.root_min;
        T.root_w = tr->h.root_width;
        T.cls = cls[i];
    }
    const int64_t B = o.box_off[nt];
    KF_REQUIRE(B < ((int64_t)1 << 31) / 27, KF_ERR_ARG, "too many boxes for one plan");
    DevBuf<TreeSrc> d_ts;
    DevBuf<int64_t> d_pbox;
    DevBuf<int32_t> d_cls;
    {
        std::vector<TreeSrc> hv(ts);
        d_ts.alloc(nt);
        KF_HIP_CHECK(hipMemcpyAsync(d_ts.p, hv.data(), nt * sizeof(TreeSrc), hipMemcpyHostToDevice, s));
        d_pbox.upload(o.box_off, s);
        d_cls.upload(cls, s);
        KF_HIP_CHECK(hipStreamSynchronize(s));  // hv is the copy's source
    }
    // per-box arrays
    d.cx.alloc(B); d.cy.alloc(B); d.bot.alloc(B); d.top.alloc(B); d.lvl.alloc(B); d.bop.alloc(B);
    d.child0.alloc(B); d.coll.alloc((size_t)B * 9); d.rid.alloc(B); d.leaf.alloc(B);
    DevBuf<int32_t> ix, iy, parent, ns, S;
    DevBuf<uint8_t> fake;
    ix.alloc(B); iy.alloc(B); parent.alloc(B); ns.alloc(B); S.alloc(B); fake.alloc(B);
    Boxes bx{d.cx.p, d.cy.p, d.bot.p, d.top.p, d.lvl.p, d.bop.p, d.child0.p, d.coll.p, ix.p, iy.p, parent.p, ns.p,
             S.p, d.leaf.p, fake.p};
    k_gather_boxes<<<nblk(B), PT, 0, s>>>(B, nt, d_pbox.p, d_ts.p, L, bx);
    KF_HIP_CHECK(hipGetLastError());
    // memberships, scans, scatter, sorts
    DevBuf<uint32_t> flags, idx, scr, upkey, dnkey, up_hist, dn_hist;
    flags.alloc((size_t)NF * B);
    idx.alloc((size_t)NF * B);
    const int64_t sw = scan_scratch(B);
    scr.alloc((size_t)NF * sw);
    upkey.alloc(B);
    dnkey.alloc(B);
    const int n_up_keys = L * C * 2, n_dn_keys = L * C * 8;
    up_hist.alloc(n_up_keys);
    dn_hist.alloc(n_dn_keys);
    up_hist.zero(s);
    dn_hist.zero(s);
    k_flags<<<nblk(B), PT, 0, s>>>(B, bx, d_cls.p, d_pbox.p, nt, C, flags.p, upkey.p, dnkey.p);
    for (int f = 0; f < NF; ++f) scan_u32(flags.p + f * B, idx.p + f * B, B, scr.p + f * sw, s);
    KF_HIP_CHECK(hipGetLastError());
    DevBuf<int32_t> up_b, dn_b, vb;
    DevBuf<uint64_t> up_k, dn_k, kb;
    DevBuf<uint32_t> hist;
    up_b.alloc(B); dn_b.alloc(B); up_k.alloc(B); dn_k.alloc(B);
    d.leaves.alloc(B);
    d.p2m.alloc(B);
    const uint32_t* n_src_dev = scr.p + F_SRC * sw + scan_total_index(B);
    k_scatter_lists<<<nblk(B), PT, 0, s>>>(B, bx, flags.p, idx.p, upkey.p, dnkey.p, n_src_dev, d.rid.p, up_b.p, up_k.p,
                                           dn_b.p, dn_k.p, d.leaves.p, d.p2m.p, up_hist.p, dn_hist.p);
    KF_HIP_CHECK(hipGetLastError());
    std::vector<uint32_t> tot(NF);
    for (int f = 0; f < NF; ++f)
        KF_HIP_CHECK(hipMemcpyAsync(&tot[f], scr.p + f * sw + scan_total_index(B), 4, hipMemcpyDeviceToHost, s));
    std::vector<uint32_t> uh(n_up_keys), dh(n_dn_keys);
    KF_HIP_CHECK(hipMemcpyAsync(uh.data(), up_hist.p, uh.size() * 4, hipMemcpyDeviceToHost, s));
    KF_HIP_CHECK(hipMemcpyAsync(dh.data(), dn_hist.p, dh.size() * 4, hipMemcpyDeviceToHost, s));
    KF_HIP_CHECK(hipStreamSynchronize(s));
    const int64_t nU = tot[F_UP], nD = tot[F_DN], nL = tot[F_LF], nP = tot[F_P2M];
    kb.alloc(std::max<int64_t>(1, std::max(nU, nD)));
    vb.alloc(std::max<int64_t>(1, std::max(nU, nD)));
    hist.alloc((size_t)sort_hist_words(std::max<int64_t>(1, std::max(nU, nD))));
    KF_HIP_CHECK(sort_pairs(up_k.p, up_b.p, nU, bits_for(n_up_keys), kb.p, vb.p, hist.p, s));
    KF_HIP_CHECK(sort_pairs(dn_k.p, dn_b.p, nD, bits_for(n_dn_keys), kb.p, vb.p, hist.p, s));
    KF_HIP_CHECK(hipGetLastError());
    o.n_src = tot[F_SRC];
    o.n_all = o.n_src + tot[F_EMP];
    // host tables: upward ranges per op index (list order: level, class, part)
    const int n_ops = C * L;
    o.up_off.assign(n_ops, 0);
    o.up_n.assign(n_ops, 0);
    o.up_nm2m.assign(n_ops, 0);
    kf_plan_stats& st = o.stats;
    st.p = p;
    st.p_padded = P;
    {
        int64_t acc = 0;
        for (int l = 0; l < L; ++l)
            for (int c = 0; c < C; ++c) {
                const int op = c * L + l;
                const uint32_t a = uh[(l * C + c) * 2], b = uh[(l * C + c) * 2 + 1];
                o.up_off[op] = (int)acc;
                o.up_nm2m[op] = (int)a;
                o.up_n[op] = (int)(a + b);
                acc += a + b;
                st.m2m_boxes += a;
                st.c2e_up_rows += a + b;
            }
        acc = 0;
        for (int l = 0; l < L; ++l)
            for (int c = 0; c < C; ++c)
                for (int q = 0; q < 4; ++q) {
                    const int key = ((l * C + c) * 4 + q) * 2;
                    const uint32_t a = dh[key], b = dh[key + 1];
                    if (l >= 2) o.groups.push_back(DownGroup{(int)acc, (int)a, (int)(a + b), q, l, c * L + l});
                    acc += a + b;
                    if (l >= 2) st.c2e_down_rows += a;
                }
    }
    st.leaves = nL;
    // per-entry lists
    d.urow.alloc(std::max<int64_t>(1, nU));
    d.crow.alloc(std::max<int64_t>(1, 4 * nU));
    d.trow.alloc(std::max<int64_t>(1, nD));
    d.prow.alloc(std::max<int64_t>(1, nD));
    d.vsrc.alloc(std::max<int64_t>(1, 27 * nD));
    d.lcoll.alloc(std::max<int64_t>(1, 9 * nL));
    DevBuf<unsigned long long> stats;
    stats.alloc(5);
    stats.zero(s);
    const VSlots& v = vslots();
    VSl vs;
    for (int q = 0; q < 4; ++q)
        for (int sl = 0; sl < 27; ++sl) {
            vs.ox[q][sl] = (int8_t)v.ox[q][sl];
            vs.oy[q][sl] = (int8_t)v.oy[q][sl];
        }
    if (nU) k_up_rows<<<nblk(nU), PT, 0, s>>>((int)nU, up_b.p, bx, d.rid.p, d.urow.p, d.crow.p);
    if (nD) k_down_rows<<<nblk(nD), PT, 0, s>>>((int)nD, dn_b.p, bx, d.rid.p, vs, d.trow.p, d.prow.p, d.vsrc.p, stats.p);
    if (nL) k_leaf_rows<<<nblk(nL), PT, 0, s>>>((int)nL, d.leaves.p, bx, p, d.lcoll.p, stats.p);
    if (nP) k_p2m_evals<<<nblk(nP), PT, 0, s>>>((int)nP, d.p2m.p, bx, p, stats.p);
    KF_HIP_CHECK(hipGetLastError());
    unsigned long long hs[5];
    KF_HIP_CHECK(hipMemcpyAsync(hs, stats.p, sizeof(hs), hipMemcpyDeviceToHost, s));
    KF_HIP_CHECK(hipStreamSynchronize(s));  // the temporaries above are freed on return
    st.m2l_pairs = (int64_t)hs[0];
    st.l2l_boxes = (int64_t)hs[1];
    st.p2p_pairs = (int64_t)hs[2];
    st.l2p_evals = (int64_t)hs[3];
    st.p2m_evals = (int64_t)hs[4];
    // list sizes the host code reads (the arrays themselves stay on the device)
    o.p2m_count = nP;
    o.leaf_count = nL;
    o.up_count = nU;
    o.dn_count = nD;
    o.box_count = B;
}

void dev_tile_masks(const int4* tiles, int count, const int32_t* vsrc, uint32_t* masks, hipStream_t s) {
    if (count <= 0) return;
    k_masks<<<nblk(count), PT, 0, s>>>(tiles, count, vsrc, masks);
    KF_HIP_CHECK(hipGetLastError());
}

std::string compare_with_host_lists(const std::vector<const kf_tree*>& trees, int p, int P, const DevLists& d,
                                    const PlanLists& dh, hipStream_t s) {
    // host trees from the sorted keys, then the host builder's lists
    std::vector<HostTree> ht(trees.size());
    std::vector<const HostTree*> hp;
    for (size_t i = 0; i < trees.size(); ++i) {
        const kf_tree* tr = trees[i];
        const DevForest& F = *tr->dev;
        HostTree& h = ht[i];
        h.n = tr->h.n;
        h.cutoff = tr->h.cutoff;
        h.root_min = tr->h.root_min;
        h.root_width = tr->h.root_width;
        h.key_depth = tr->h.key_depth;
        h.keys.resize(h.n);
        KF_HIP_CHECK(hipMemcpyAsync(h.keys.data(), F.keys.p + F.pt_off[tr->slot], h.n * 8, hipMemcpyDeviceToHost, s));
        KF_HIP_CHECK(hipStreamSynchronize(s));
        std::string e;
        try {
            e = h.build();
        } catch (const std::bad_alloc&) {
            return "host build of tree " + std::to_string(i) + " (n " + std::to_string(h.n) + ", cutoff " +
                   std::to_string(h.cutoff) + ", depth " + std::to_string(h.key_depth) + ") ran out of memory";
        }
        if (!e.empty()) return "host build of tree " + std::to_string(i) + ": " + e;
        hp.push_back(&h);
    }
    PlanLists hl;
    build_forest_lists(hp, p, P, hl);
    auto cmp_i = [&](const char* name, const std::vector<int32_t>& want, const DevBuf<int32_t>& got,
                     int64_t n) -> std::string {
        if ((int64_t)want.size() != n) return std::string(name) + ": size " + std::to_string(n) + " vs host " + std::to_string(want.size());
        std::vector<int32_t> g(n);
        if (n) KF_HIP_CHECK(hipMemcpy(g.data(), got.p, n * 4, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < n; ++k)
            if (g[k] != want[k])
                return std::string(name) + "[" + std::to_string(k) + "] = " + std::to_string(g[k]) + ", host " +
                       std::to_string(want[k]);
        return "";
    };
    std::string e;
    const int64_t B = dh.box_count;
    if (!(e = cmp_i("bot", hl.bot, d.bot, B)).empty()) return e;
    if (!(e = cmp_i("top", hl.top, d.top, B)).empty()) return e;
    if (!(e = cmp_i("lvl", hl.lvl, d.lvl, B)).empty()) return e;
    if (!(e = cmp_i("bop", hl.bop, d.bop, B)).empty()) return e;
    if (!(e = cmp_i("child0", hl.child0, d.child0, B)).empty()) return e;
    if (!(e = cmp_i("coll", hl.coll, d.coll, 9 * B)).empty()) return e;
    if (!(e = cmp_i("rid", hl.rid, d.rid, B)).empty()) return e;
    if (!(e = cmp_i("p2m", hl.p2m, d.p2m, dh.p2m_count)).empty()) return e;
    if (!(e = cmp_i("urow", hl.urow, d.urow, dh.up_count)).empty()) return e;
    if (!(e = cmp_i("crow", hl.crow, d.crow, 4 * dh.up_count)).empty()) return e;
    if (!(e = cmp_i("trow", hl.trow, d.trow, dh.dn_count)).empty()) return e;
    if (!(e = cmp_i("prow", hl.prow, d.prow, dh.dn_count)).empty()) return e;
    if (!(e = cmp_i("vsrc", hl.vsrc, d.vsrc, 27 * dh.dn_count)).empty()) return e;
    if (!(e = cmp_i("leaves", hl.leaves, d.leaves, dh.leaf_count)).empty()) return e;
    if (!(e = cmp_i("lcoll", hl.lcoll, d.lcoll, 9 * dh.leaf_count)).empty()) return e;
    {
        std::vector<double> cx(B), cy(B);
        std::vector<uint8_t> lf(B);
        if (B) {
            KF_HIP_CHECK(hipMemcpy(cx.data(), d.cx.p, B * 8, hipMemcpyDeviceToHost));
            KF_HIP_CHECK(hipMemcpy(cy.data(), d.cy.p, B * 8, hipMemcpyDeviceToHost));
            KF_HIP_CHECK(hipMemcpy(lf.data(), d.leaf.p, B, hipMemcpyDeviceToHost));
        }
        for (int64_t b = 0; b < B; ++b)
            if (cx[b] != hl.cx[b] || cy[b] != hl.cy[b] || lf[b] != hl.leaf[b])
                return "box " + std::to_string(b) + ": centre or leaf flag differs";
    }
    if (hl.n_src != dh.n_src || hl.n_all != dh.n_all) return "slab row counts differ";
    if (hl.up_off != dh.up_off || hl.up_n != dh.up_n || hl.up_nm2m != dh.up_nm2m) return "upward ranges differ";
    if (hl.groups.size() != dh.groups.size()) return "downward group count differs";
    for (size_t g = 0; g < hl.groups.size(); ++g) {
        const DownGroup &a = hl.groups[g], &b = dh.groups[g];
        if (a.first != b.first || a.n_src != b.n_src || a.n_all != b.n_all || a.q != b.q || a.level != b.level ||
            a.op != b.op)
            return "downward group " + std::to_string(g) + " differs";
    }
    const kf_plan_stats &x = hl.stats, &y = dh.stats;
    if (x.p2p_pairs != y.p2p_pairs || x.p2m_evals != y.p2m_evals || x.l2p_evals != y.l2p_evals ||
        x.m2l_pairs != y.m2l_pairs || x.m2m_boxes != y.m2m_boxes || x.l2l_boxes != y.l2l_boxes ||
        x.c2e_up_rows != y.c2e_up_rows || x.c2e_down_rows != y.c2e_down_rows || x.leaves != y.leaves)
        return "plan statistics differ";
    if (hl.class_width != dh.class_width || hl.L != dh.L || hl.pt_off != dh.pt_off || hl.box_off != dh.box_off)
        return "plan tables differ";
    return "";
}

}  // namespace kf
