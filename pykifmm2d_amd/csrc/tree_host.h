// tree_host.h -- box-level quadtree structure built on the host from sorted Morton keys.
//
// The point-level work (Morton keys, radix sort, point gather) runs on the GPU; this
// structure holds the O(boxes) part: levels, children, colleagues, level restriction and the
// X-list fake children, with the same per-level box order as the reference Tree
// (reference pykifmm2d/tree.py:460-576).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace kf {

struct HostTree {
    // geometry of the root box (reference tree.py:475-484) and key resolution
    int64_t n = 0;
    int cutoff = 0;
    double root_min = 0.0;   // mmin
    double root_width = 0.0; // mmax - mmin
    int key_depth = 0;       // D: keys hold 2*D bits
    std::vector<uint64_t> keys;  // sorted Morton keys (x bit above y bit in each pair)

    // boxes, global ids in creation order; the four children of a box are consecutive
    std::vector<int32_t> lvl;
    std::vector<int64_t> ix, iy;          // grid coordinates at the box's level
    std::vector<int64_t> bot, top;        // point range in tree order
    std::vector<int32_t> parent, child0;  // global ids, -1 if none
    std::vector<uint8_t> leaf, fake, xl;  // reference Level.leaf, fake_leaf, Xlist
    std::vector<int32_t> depth;
    std::vector<int32_t> pos;             // index of the box within its level

    std::vector<std::vector<int32_t>> levels;  // global ids per level, reference order
    std::vector<std::vector<int32_t>> coll;    // per level: n_node*9 global ids (-1 padded)
    std::vector<int32_t> n_real;               // per level: boxes that are not fake

    int nlevels() const { return (int)levels.size(); }
    int64_t ns(int32_t b) const { return top[b] - bot[b]; }
    double width(int l) const;
    double xmid(int32_t b) const;
    double ymid(int32_t b) const;

    // Builds levels, colleagues, restriction and X-list from `keys`.  Returns "" or an error.
    std::string build();

  private:
    int32_t new_box(int l, int64_t x, int64_t y, int64_t b, int64_t t, int32_t par, bool is_fake);
    bool split(int32_t b, bool is_fake, std::string& err);
    void tag_colleagues(int l);
    void raise_depths(int32_t b);
};

}  // namespace kf
