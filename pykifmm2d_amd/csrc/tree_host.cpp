// tree_host.cpp -- box-level quadtree construction from sorted Morton keys.
//
// Each step names the reference function it reproduces (pykifmm2d/tree.py).  Boxes are
// identified by (level, ix, iy); a point belongs to a box iff its key has the box's prefix,
// which is exactly the reference's strict-">" midline classification (tree.py:33-35)
// because the keys are computed against the same exact dyadic gridlines.
#include "tree_host.h"

#include <algorithm>
#include <cmath>

namespace kf {

static inline uint64_t interleave(uint64_t x, uint64_t y) {
    // x bits go to odd positions, y bits to even positions (child class = 2*hx + hy)
    auto spread = [](uint64_t v) {
        v &= 0xffffffffull;
        v = (v | (v << 16)) & 0x0000ffff0000ffffull;
        v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
        v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
        v = (v | (v << 2)) & 0x3333333333333333ull;
        v = (v | (v << 1)) & 0x5555555555555555ull;
        return v;
    };
    return (spread(x) << 1) | spread(y);
}

double HostTree::width(int l) const { return std::ldexp(root_width, -l); }
double HostTree::xmid(int32_t b) const {
    return root_min + ((double)ix[b] + 0.5) * width(lvl[b]);
}
double HostTree::ymid(int32_t b) const {
    return root_min + ((double)iy[b] + 0.5) * width(lvl[b]);
}

int32_t HostTree::new_box(int l, int64_t x, int64_t y, int64_t b, int64_t t, int32_t par,
                          bool is_fake) {
    int32_t id = (int32_t)lvl.size();
    lvl.push_back(l);
    ix.push_back(x);
    iy.push_back(y);
    bot.push_back(b);
    top.push_back(t);
    parent.push_back(par);
    child0.push_back(-1);
    leaf.push_back(is_fake ? 0 : 1);
    fake.push_back(is_fake ? 1 : 0);
    xl.push_back(0);
    depth.push_back(0);
    if ((int)levels.size() <= l) levels.resize(l + 1);
    pos.push_back((int32_t)levels[l].size());
    levels[l].push_back(id);
    return id;
}

// tree.py:128-190 divide_and_reorder for one box: the children's point ranges are the key
// ranges of the four child prefixes, found by binary search in the sorted keys.
bool HostTree::split(int32_t b, bool is_fake, std::string& err) {
    int l = lvl[b];
    if (l + 1 > key_depth) {
        err = "tree needs more levels than the key resolution allows (coincident points or "
              "extreme clustering)";
        return false;
    }
    uint64_t m = interleave((uint64_t)ix[b], (uint64_t)iy[b]);
    int shift = 2 * (key_depth - l - 1);
    int64_t cut[5];
    cut[0] = bot[b];
    cut[4] = top[b];
    const uint64_t* k0 = keys.data();
    for (int q = 1; q < 4; ++q) {
        uint64_t lo = ((m << 2) | (uint64_t)q) << shift;
        cut[q] = std::lower_bound(k0 + cut[q - 1], k0 + top[b], lo) - k0;
    }
    int32_t first = -1;
    for (int q = 0; q < 4; ++q) {
        int32_t c = new_box(l + 1, 2 * ix[b] + (q >> 1), 2 * iy[b] + (q & 1), cut[q], cut[q + 1],
                            b, is_fake);
        if (q == 0) first = c;
    }
    child0[b] = first;
    if (!is_fake) leaf[b] = 0;
    return true;
}

// tree.py:349-375 Level.tag_colleagues: same-level boxes within the 3x3 neighbourhood.  The
// slot order follows the reference: level order for small levels (direct search, 260-272),
// parent-colleague order otherwise (274-294).
void HostTree::tag_colleagues(int l) {
    const auto& L = levels[l];
    size_t n = L.size();
    auto& C = coll[l];
    C.assign(n * 9, -1);
    auto near = [&](int32_t a, int32_t c) {
        return std::llabs(ix[a] - ix[c]) <= 1 && std::llabs(iy[a] - iy[c]) <= 1;
    };
    if (n < 100 || l == 0) {
        for (size_t i = 0; i < n; ++i) {
            int slot = 0;
            for (size_t j = 0; j < n && slot < 9; ++j)
                if (near(L[i], L[j])) C[i * 9 + slot++] = L[j];
        }
        return;
    }
    const auto& PC = coll[l - 1];
    for (size_t i = 0; i < n; ++i) {
        int32_t b = L[i];
        int32_t p = parent[b];
        int slot = 0;
        for (int j = 0; j < 9; ++j) {
            int32_t pj = PC[(size_t)pos[p] * 9 + j];
            if (pj < 0 || child0[pj] < 0) continue;
            for (int q = 0; q < 4; ++q) {
                int32_t c = child0[pj] + q;
                if (near(b, c) && slot < 9) C[i * 9 + slot++] = c;
            }
        }
    }
}

// tree.py:423-430 / 582-585: depth = 0 for a leaf, else 1 + max child depth.  After a leaf
// is split its depth becomes 1 and ancestors are raised as needed.
void HostTree::raise_depths(int32_t b) {
    int32_t p = b;
    while (p >= 0) {
        int32_t c = child0[p];
        int32_t d = 0;
        if (c >= 0 && !leaf[p])
            d = 1 + std::max(std::max(depth[c], depth[c + 1]), std::max(depth[c + 2], depth[c + 3]));
        if (d == depth[p] && p != b) break;
        depth[p] = d;
        p = parent[p];
    }
}

std::string HostTree::build() {
    std::string err;
    levels.clear();
    // tree.py:491-506: root box, then split every box with more than `cutoff` points
    new_box(0, 0, 0, 0, n, -1, false);
    if (n > cutoff) {
        int cur = 0;
        for (;;) {
            levels.resize(cur + 2);
            std::vector<int32_t> here = levels[cur];
            for (int32_t b : here)
                if (ns(b) > cutoff && !split(b, false, err)) return err;
            bool more = false;
            for (int32_t b : levels[cur + 1]) more |= ns(b) > cutoff;
            if (!more) break;
            ++cur;
        }
    }
    int L = nlevels();
    coll.assign(L, {});
    for (int l = 0; l < L; ++l) tag_colleagues(l);
    // initial depths, bottom-up
    for (int l = L - 1; l >= 0; --l)
        for (int32_t b : levels[l]) {
            int32_t c = child0[b];
            depth[b] = (c >= 0 && !leaf[b])
                ? 1 + std::max(std::max(depth[c], depth[c + 1]), std::max(depth[c + 2], depth[c + 3]))
                : 0;
        }
    // tree.py:533-545 level restriction sweeps with 432-444 numba_get_bads
    for (;;) {
        int64_t new_nodes = 0;
        for (int ind = 1; ind < L - 2; ++ind) {
            std::vector<int32_t> bads;
            const auto& lv = levels[ind];
            const auto& C = coll[ind];
            for (size_t i = 0; i < lv.size(); ++i) {
                int32_t b = lv[i];
                if (!leaf[b]) continue;
                bool bad = false;
                for (int j = 0; j < 9; ++j) {
                    int32_t c = C[i * 9 + j];
                    if (c >= 0) {
                        int d = depth[b] - depth[c];
                        if (d > 1 || d < -1) bad = true;
                    }
                }
                if (bad) bads.push_back(b);
            }
            for (int32_t b : bads)
                if (!split(b, false, err)) return err;
            if (!bads.empty()) {
                tag_colleagues(ind + 1);
                for (int32_t b : bads) raise_depths(b);
            }
            new_nodes += (int64_t)bads.size();
        }
        if (new_nodes == 0) break;
    }
    // tree.py:546-557 X-list: a real leaf with a real colleague of positive depth gets four
    // fake children (446-458; colleagues appended as fakes are ignored, tree.py:454)
    n_real.assign(L, 0);
    for (int ind = 0; ind < L; ++ind) {
        const std::vector<int32_t> lv = levels[ind];
        int32_t nr = 0;
        for (int32_t b : lv) nr += fake[b] ? 0 : 1;
        n_real[ind] = nr;
        if (ind == 0 || ind == L - 1) continue;
        std::vector<int32_t> xs;
        const auto& C = coll[ind];
        for (int32_t i = 0; i < nr; ++i) {
            int32_t b = lv[i];
            if (!leaf[b]) continue;
            bool x = false;
            for (int j = 0; j < 9; ++j) {
                int32_t c = C[(size_t)i * 9 + j];
                if (c >= 0 && pos[c] < nr && depth[b] - depth[c] < 0) x = true;
            }
            if (x) xs.push_back(b);
        }
        for (int32_t b : xs) {
            xl[b] = 1;
            if (!split(b, true, err)) return err;
        }
        if (!xs.empty()) tag_colleagues(ind + 1);
    }
    return "";
}

}  // namespace kf
