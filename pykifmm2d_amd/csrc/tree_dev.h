// tree_dev.h -- the box structure of one tree or of a batch of trees, built on the GPU from the
// trees' sorted Morton keys (reference pykifmm2d/tree.py:464-576).
//
// Boxes of tree t have local ids in creation order (the reference's order of box creation:
// level by level while splitting by count, then the level-restriction sweeps, then the X-list
// fake children top-down); arrays are laid out tree by tree, box `lid` of tree t at box_off[t] +
// lid.  parent / child0 / colleagues hold local ids.  Within a level, boxes have the reference's
// positions and colleague slots the reference's order.
#pragma once
#include <cstdint>
#include <vector>

#include "kf_internal.h"

namespace kf {

struct DevForest {
    int nt = 0;
    int cutoff = 0;
    std::vector<int64_t> pt_off;      // points of tree t: [pt_off[t], pt_off[t + 1]) of the batch
    std::vector<int32_t> D;           // key depth per tree
    std::vector<int32_t> L;           // levels per tree
    std::vector<int64_t> box_off;     // boxes of tree t: [box_off[t], box_off[t + 1])
    std::vector<int64_t> lev_off;     // level list of (t, l) at lev_off[t * 32 + l], size nlev
    std::vector<int32_t> nlev;        // boxes at (t, l), t * 32 + l
    DevBuf<uint64_t> keys;            // sorted keys, tree by tree
    // boxes (box_off order)
    DevBuf<int32_t> lvl, ix, iy, bot, top, parent, child0, depth, pos, coll;  // coll: 9 per box
    DevBuf<uint8_t> leaf, fake, xl;
    DevBuf<int32_t> level_list;       // local ids by (t, l, position)
    int64_t n_boxes() const { return box_off.empty() ? 0 : box_off.back(); }
};

// Builds the box structure of nt trees whose sorted keys (tree t's n_t keys at pt_off[t], depth
// D[t]) are in `keys` (taken over by the result).  Throws KfError (KF_ERR_DEPTH when more than
// `cutoff` points share a finest-level box).  Synchronises the stream.
void build_dev_forest(DevForest& F, DevBuf<uint64_t>& keys, const std::vector<int64_t>& pt_off,
                      const std::vector<int32_t>& D, int cutoff, hipStream_t s);

// Host copy of tree t of a device forest in the HostTree layout (exports, host-side checks).
void dev_forest_to_host(const DevForest& F, int t, HostTree& h, hipStream_t s);

}  // namespace kf
