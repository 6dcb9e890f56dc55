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
            o.groups.push_back(DownGroup{start, n_with, (int)o.trow.size() - start, q, l});
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

void build_tile_table(const PlanLists& pl, int lg, int mode, std::vector<Tile>& tiles,
                      std::vector<uint32_t>& masks, std::vector<int>& off, std::vector<int>& n) {
    const int bpt = 64 >> lg;
    tiles.clear();
    masks.clear();
    off.assign(pl.L, 0);
    n.assign(pl.L, 0);
    int cur_level = -1;
    for (const DownGroup& g : pl.groups) {
        if (g.level != cur_level) {
            cur_level = g.level;
            off[cur_level] = (int)tiles.size();
        }
        const int cnt = mode == KF_TGT_ALL ? g.n_all : g.n_src;
        for (int o = 0; o < cnt; o += bpt) {
            const int m = std::min(bpt, cnt - o);
            tiles.push_back(Tile{g.first + o, m, g.q, g.level});
            uint32_t mk = 0;
            for (int k = 0; k < m; ++k)
                for (int sl = 0; sl < 27; ++sl)
                    if (pl.vsrc[(size_t)(g.first + o + k) * 27 + sl] >= 0) mk |= 1u << sl;
            masks.push_back(mk);
        }
        n[cur_level] = (int)tiles.size() - off[cur_level];
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

}  // namespace kf
