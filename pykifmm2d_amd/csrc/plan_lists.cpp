// plan_lists.cpp -- host-side work lists of a planned evaluation (see plan_lists.h).
//
// The lists restate the reference's pass structure (pykifmm2d/class_fmm.py:299-352,
// fmm.py:350-408): which boxes take part in the upward pass, the interaction lists of the
// downward pass by child position, and the leaves of the near field.
#include "plan_lists.h"

#include <algorithm>
#include <cstdlib>

namespace kf {

int m2l_offset_index(int ox, int oy) {
    int idx = 0;
    for (int x = -3; x <= 3; ++x)
        for (int y = -3; y <= 3; ++y) {
            if (std::abs(x) <= 1 && std::abs(y) <= 1) continue;
            if (x == ox && y == oy) return idx;
            ++idx;
        }
    return -1;
}

void m2l_offset_of(int idx, int& ox, int& oy) {
    int k = 0;
    for (int x = -3; x <= 3; ++x)
        for (int y = -3; y <= 3; ++y) {
            if (std::abs(x) <= 1 && std::abs(y) <= 1) continue;
            if (k == idx) {
                ox = x;
                oy = y;
                return;
            }
            ++k;
        }
    ox = oy = 0;
}

static VSlots make_vslots() {
    VSlots v{};
    for (int q = 0; q < 4; ++q) {
        int a = q >> 1, b = q & 1, s = 0;
        for (int ox = -2 - a; ox <= 3 - a; ++ox)
            for (int oy = -2 - b; oy <= 3 - b; ++oy) {
                if (std::abs(ox) <= 1 && std::abs(oy) <= 1) continue;
                v.ox[q][s] = ox;
                v.oy[q][s] = oy;
                v.op[q][s] = m2l_offset_index(ox, oy);
                ++s;
            }
    }
    return v;
}

const VSlots& vslots() {
    static const VSlots v = make_vslots();
    return v;
}

void build_plan_lists(const HostTree& h, int p, int P, PlanLists& o) {
    const int L = h.nlevels();
    const int64_t B = (int64_t)h.lvl.size();
    o = PlanLists{};
    o.L = L;
    o.class_width = {h.root_width};
    o.pt_off = {0, h.n};
    o.box_off = {0, B};
    // per-box arrays
    o.cx.resize(B);
    o.cy.resize(B);
    o.bot.resize(B);
    o.top.resize(B);
    o.lvl.resize(B);
    o.child0.resize(B);
    o.leaf.resize(B);
    o.coll.assign((size_t)B * 9, -1);
    for (int64_t b = 0; b < B; ++b) {
        o.cx[b] = h.xmid((int32_t)b);
        o.cy[b] = h.ymid((int32_t)b);
        o.bot[b] = (int32_t)h.bot[b];
        o.top[b] = (int32_t)h.top[b];
        o.lvl[b] = h.lvl[b];
        o.child0[b] = h.child0[b];
        o.leaf[b] = h.leaf[b];
        const auto& C = h.coll[h.lvl[b]];
        for (int j = 0; j < 9; ++j) o.coll[(size_t)b * 9 + j] = C[(size_t)h.pos[b] * 9 + j];
    }
    o.bop = o.lvl;
    // slab rows: boxes with sources at levels >= 1, then real boxes without sources at >= 2
    o.rid.assign(B, -1);
    int32_t nr = 0;
    for (int l = 1; l < L; ++l)
        for (int32_t b : h.levels[l])
            if (h.ns(b) > 0) o.rid[b] = nr++;
    o.n_src = nr;
    for (int l = 2; l < L; ++l)
        for (int32_t b : h.levels[l])
            if (!h.fake[b] && h.ns(b) <= 0) o.rid[b] = nr++;
    o.n_all = nr;
    kf_plan_stats& st = o.stats;
    st.p = p;
    st.p_padded = P;
    // upward lists (reference class_fmm.py:309-319): per level the boxes with sources, boxes
    // with children first
    o.up_off.assign(L, 0);
    o.up_n.assign(L, 0);
    o.up_nm2m.assign(L, 0);
    for (int l = 1; l < L; ++l) {
        o.up_off[l] = (int)o.urow.size();
        for (int32_t b : h.levels[l])
            if (h.ns(b) > 0 && h.child0[b] >= 0) {
                o.urow.push_back(o.rid[b]);
                for (int q = 0; q < 4; ++q) {
                    // a child without sources has no upward expansion (its row, if any, is a
                    // downward-only row beyond the X slab)
                    const int32_t c = h.child0[b] + q;
                    o.crow.push_back(h.ns(c) > 0 ? o.rid[c] : -1);
                }
            }
        o.up_nm2m[l] = (int)o.urow.size() - o.up_off[l];
        for (int32_t b : h.levels[l])
            if (h.ns(b) > 0 && h.child0[b] < 0) {
                o.urow.push_back(o.rid[b]);
                for (int q = 0; q < 4; ++q) o.crow.push_back(-1);
                o.p2m.push_back(b);
                st.p2m_evals += (int64_t)p * h.ns(b);
            }
        o.up_n[l] = (int)o.urow.size() - o.up_off[l];
        st.m2m_boxes += o.up_nm2m[l];
        st.c2e_up_rows += o.up_n[l];
    }
    // downward lists (reference class_fmm.py:328-350): the real boxes of levels >= 2 by child
    // position, those with sources first; interaction-list sources are the children of the
    // parent's colleagues that are not adjacent and hold sources
    const VSlots& vs = vslots();
    for (int l = 2; l < L; ++l) {
        const int64_t side = (int64_t)1 << l;
        for (int q = 0; q < 4; ++q) {
            const int start = (int)o.trow.size();
            int n_with = 0;
            for (int pass = 0; pass < 2; ++pass)
                for (int32_t t : h.levels[l]) {
                    if (h.fake[t] || 2 * (h.ix[t] & 1) + (h.iy[t] & 1) != q) continue;
                    const bool has = h.ns(t) > 0;
                    if (has != (pass == 0)) continue;
                    const int32_t par = h.parent[t];
                    o.trow.push_back(o.rid[t]);
                    o.prow.push_back(o.rid[par]);
                    int64_t used = 0;
                    for (int sl = 0; sl < 27; ++sl) {
                        const int64_t gx = h.ix[t] + vs.ox[q][sl], gy = h.iy[t] + vs.oy[q][sl];
                        int32_t src = -1;
                        if (gx >= 0 && gy >= 0 && gx < side && gy < side) {
                            const auto& C = h.coll[l - 1];
                            for (int j = 0; j < 9; ++j) {
                                const int32_t c = C[(size_t)h.pos[par] * 9 + j];
                                if (c < 0 || h.ix[c] != (gx >> 1) || h.iy[c] != (gy >> 1)) continue;
                                if (h.child0[c] < 0) {
                                    if (has) ++o.aliased_pairs;
                                    break;
                                }
                                const int32_t cand = h.child0[c] + (int32_t)(2 * (gx & 1) + (gy & 1));
                                if (h.ns(cand) > 0) src = o.rid[cand];
                                break;
                            }
                        }
                        o.vsrc.push_back(src);
                        used += src >= 0;
                    }
                    if (has) {
                        ++n_with;
                        st.m2l_pairs += used;
                        if (l >= 3) st.l2l_boxes += 1;
                        st.c2e_down_rows += 1;
                    }
                }
            o.groups.push_back(DownGroup{start, n_with, (int)o.trow.size() - start, q, l, l});
        }
    }
    // leaves (reference fmm.py:30-64, 91-115): real leaves with points, levels >= 1
    for (int l = 1; l < L; ++l)
        for (int32_t b : h.levels[l]) {
            if (!h.leaf[b] || h.ns(b) <= 0) continue;
            o.leaves.push_back(b);
            int64_t src = 0;
            for (int j = 0; j < 9; ++j) {
                const int32_t c = h.coll[l][(size_t)h.pos[b] * 9 + j];
                o.lcoll.push_back(c);
                if (c >= 0) src += h.ns(c);
            }
            st.p2p_pairs += h.ns(b) * src - h.ns(b);
            if (l >= 2) st.l2p_evals += (int64_t)p * h.ns(b);
            st.leaves += 1;
        }
}

void build_forest_lists(const std::vector<const HostTree*>& trees, int p, int P, PlanLists& o) {
    const int nt = (int)trees.size();
    std::vector<PlanLists> T(nt);
    for (int i = 0; i < nt; ++i) build_plan_lists(*trees[i], p, P, T[i]);
    if (nt == 1) {
        o = std::move(T[0]);
        return;
    }
    o = PlanLists{};
    o.n_trees = nt;
    // classes of equal root width share operators
    std::vector<int> cls(nt);
    for (int i = 0; i < nt; ++i) {
        const double w = trees[i]->root_width;
        int c = 0;
        while (c < (int)o.class_width.size() && o.class_width[c] != w) ++c;
        if (c == (int)o.class_width.size()) o.class_width.push_back(w);
        cls[i] = c;
        o.L = std::max(o.L, T[i].L);
    }
    o.n_classes = (int)o.class_width.size();
    const int L = o.L;
    // offsets of points, boxes and the two row groups
    o.pt_off.assign(nt + 1, 0);
    o.box_off.assign(nt + 1, 0);
    std::vector<int64_t> src_off(nt + 1, 0), emp_off(nt + 1, 0);
    for (int i = 0; i < nt; ++i) {
        o.pt_off[i + 1] = o.pt_off[i] + trees[i]->n;
        o.box_off[i + 1] = o.box_off[i] + (int64_t)T[i].lvl.size();
        src_off[i + 1] = src_off[i] + T[i].n_src;
        emp_off[i + 1] = emp_off[i] + (T[i].n_all - T[i].n_src);
    }
    o.n_src = src_off[nt];
    o.n_all = o.n_src + emp_off[nt];
    auto row = [&](int i, int32_t r) -> int32_t {
        if (r < 0) return -1;
        if (r < T[i].n_src) return (int32_t)(r + src_off[i]);
        return (int32_t)(o.n_src + emp_off[i] + (r - T[i].n_src));
    };
    auto box = [&](int i, int32_t b) -> int32_t { return b < 0 ? -1 : (int32_t)(b + o.box_off[i]); };
    const int64_t B = o.box_off[nt];
    o.cx.reserve(B);
    o.cy.reserve(B);
    for (int i = 0; i < nt; ++i) {
        const PlanLists& t = T[i];
        const int32_t po = (int32_t)o.pt_off[i];
        for (size_t b = 0; b < t.lvl.size(); ++b) {
            o.cx.push_back(t.cx[b]);
            o.cy.push_back(t.cy[b]);
            o.bot.push_back(t.bot[b] + po);
            o.top.push_back(t.top[b] + po);
            o.lvl.push_back(t.lvl[b]);
            o.bop.push_back(cls[i] * L + t.lvl[b]);
            o.child0.push_back(box(i, t.child0[b]));
            o.leaf.push_back(t.leaf[b]);
            o.rid.push_back(row(i, t.rid[b]));
            for (int j = 0; j < 9; ++j) o.coll.push_back(box(i, t.coll[b * 9 + j]));
        }
        for (int32_t b : t.p2m) o.p2m.push_back(box(i, b));
        for (size_t k = 0; k < t.leaves.size(); ++k) {
            o.leaves.push_back(box(i, t.leaves[k]));
            for (int j = 0; j < 9; ++j) o.lcoll.push_back(box(i, t.lcoll[k * 9 + j]));
        }
        o.aliased_pairs += t.aliased_pairs;
        const kf_plan_stats& s = t.stats;
        o.stats.p2p_pairs += s.p2p_pairs;
        o.stats.p2m_evals += s.p2m_evals;
        o.stats.l2p_evals += s.l2p_evals;
        o.stats.m2l_pairs += s.m2l_pairs;
        o.stats.m2m_boxes += s.m2m_boxes;
        o.stats.l2l_boxes += s.l2l_boxes;
        o.stats.c2e_up_rows += s.c2e_up_rows;
        o.stats.c2e_down_rows += s.c2e_down_rows;
        o.stats.leaves += s.leaves;
    }
    o.stats.p = p;
    o.stats.p_padded = P;
    // upward lists per op index: every tree of the class at that level, boxes with children first
    const int n_ops = o.n_classes * L;
    o.up_off.assign(n_ops, 0);
    o.up_n.assign(n_ops, 0);
    o.up_nm2m.assign(n_ops, 0);
    for (int l = 1; l < L; ++l)
        for (int c = 0; c < o.n_classes; ++c) {
            const int op = c * L + l;
            o.up_off[op] = (int)o.urow.size();
            for (int part = 0; part < 2; ++part) {
                for (int i = 0; i < nt; ++i) {
                    const PlanLists& t = T[i];
                    if (cls[i] != c || l >= t.L) continue;
                    const int e0 = t.up_off[l] + (part ? t.up_nm2m[l] : 0);
                    const int e1 = part ? t.up_off[l] + t.up_n[l] : t.up_off[l] + t.up_nm2m[l];
                    for (int e = e0; e < e1; ++e) {
                        o.urow.push_back(row(i, t.urow[e]));
                        for (int q = 0; q < 4; ++q) o.crow.push_back(row(i, t.crow[4 * (size_t)e + q]));
                    }
                }
                if (part == 0) o.up_nm2m[op] = (int)o.urow.size() - o.up_off[op];
            }
            o.up_n[op] = (int)o.urow.size() - o.up_off[op];
        }
    // downward groups per (level, class, child position): targets with sources of every tree,
    // then those without
    for (int l = 2; l < L; ++l)
        for (int c = 0; c < o.n_classes; ++c)
            for (int q = 0; q < 4; ++q) {
                const int start = (int)o.trow.size();
                int n_with = 0;
                for (int pass = 0; pass < 2; ++pass)
                    for (int i = 0; i < nt; ++i) {
                        const PlanLists& t = T[i];
                        if (cls[i] != c || l >= t.L) continue;
                        const DownGroup& g = t.groups[(size_t)(l - 2) * 4 + q];
                        const int t0 = g.first + (pass ? g.n_src : 0);
                        const int t1 = g.first + (pass ? g.n_all : g.n_src);
                        for (int k = t0; k < t1; ++k) {
                            o.trow.push_back(row(i, t.trow[k]));
                            o.prow.push_back(row(i, t.prow[k]));
                            for (int sl = 0; sl < 27; ++sl) o.vsrc.push_back(row(i, t.vsrc[(size_t)k * 27 + sl]));
                        }
                        if (pass == 0) n_with += g.n_src;
                    }
                o.groups.push_back(DownGroup{start, n_with, (int)o.trow.size() - start, q, l, c * L + l});
            }
}

void build_tile_table(const PlanLists& pl, int lg, int mode, std::vector<Tile>& tiles,
                      std::vector<uint32_t>* masks, std::vector<int>& off, std::vector<int>& n) {
    const int bpt = 64 >> lg;
    tiles.clear();
    if (masks) masks->clear();
    off.assign((size_t)pl.n_classes * pl.L, 0);
    n.assign((size_t)pl.n_classes * pl.L, 0);
    int cur = -1;
    for (const DownGroup& g : pl.groups) {
        if (g.op != cur) {
            cur = g.op;
            off[cur] = (int)tiles.size();
        }
        const int cnt = mode == KF_TGT_ALL ? g.n_all : g.n_src;
        for (int o = 0; o < cnt; o += bpt) {
            const int m = std::min(bpt, cnt - o);
            tiles.push_back(Tile{g.first + o, m, g.q, g.op});
            if (!masks) continue;
            uint32_t mk = 0;
            for (int k = 0; k < m; ++k)
                for (int sl = 0; sl < 27; ++sl)
                    if (pl.vsrc[(size_t)(g.first + o + k) * 27 + sl] >= 0) mk |= 1u << sl;
            masks->push_back(mk);
        }
        n[cur] = (int)tiles.size() - off[cur];
    }
}

int64_t tile_count(const PlanLists& pl, int lg, int mode) {
    const int bpt = 64 >> lg;
    int64_t cnt = 0;
    for (const DownGroup& g : pl.groups) {
        const int m = mode == KF_TGT_ALL ? g.n_all : g.n_src;
        cnt += (m + bpt - 1) / bpt;
    }
    return cnt;
}

void stored_offsets(const PlanLists& ls, int p, StoredOffsets& o) {
    const size_t nl = ls.leaves.size();
    o.near.assign(nl + 1, 0);
    o.l2p.assign(nl + 1, 0);
    o.p2m.assign(ls.p2m.size() + 1, 0);
    for (size_t li = 0; li < nl; ++li) {
        const int32_t b = ls.leaves[li];
        const int64_t nt = ls.top[b] - ls.bot[b];
        int64_t ns = nt;
        for (int j = 0; j < 9; ++j) {
            const int32_t c = ls.lcoll[li * 9 + j];
            if (c >= 0 && c != b) ns += ls.top[c] - ls.bot[c];
        }
        o.near[li + 1] = o.near[li] + nt * ns;
        o.l2p[li + 1] = o.l2p[li] + (ls.lvl[b] >= 2 ? nt * p : 0);
    }
    for (size_t i = 0; i < ls.p2m.size(); ++i) {
        const int32_t b = ls.p2m[i];
        o.p2m[i + 1] = o.p2m[i] + (int64_t)(ls.top[b] - ls.bot[b]) * p;
    }
}

}  // namespace kf
