// plan_dev.h -- a plan's work lists built on the GPU from its trees' device box structures.
//
// The lists are those of plan_lists.h (the host builder, kept as the cross-check and for the
// CPU index checker), in the same order: every list is a subsequence of the boxes in (tree,
// level, position) order, stably sorted by a small key where the host builder groups it
// (upward lists by level, class and part; downward targets by level, class, child position and
// whether they hold sources).
#pragma once
#include <vector>

#include "kf_internal.h"
#include "plan_lists.h"

namespace kf {

struct DevLists {
    DevBuf<double> cx, cy;
    DevBuf<int32_t> bot, top, lvl, bop, child0, coll, rid;
    DevBuf<uint8_t> leaf;
    DevBuf<int32_t> p2m, urow, crow, trow, prow, vsrc, leaves, lcoll;
};

// Device lists of the plan over `trees` (each built on the GPU) and the host tables of `host`
// (every PlanLists field except the per-box and per-entry arrays, which stay empty).
// Synchronises the stream.
void build_dev_lists(const std::vector<const kf_tree*>& trees, int p, int P, PlanLists& host, DevLists& dev,
                     hipStream_t s);

// masks[k] = the interaction-list slots any target of tile k uses (bit s: slot s).
void dev_tile_masks(const int4* tiles, int count, const int32_t* vsrc, uint32_t* masks, hipStream_t s);

// Compares the device lists with the host builder's from the trees' sorted keys; "" if equal,
// else the first difference.
std::string compare_with_host_lists(const std::vector<const kf_tree*>& trees, int p, int P, const DevLists& dev,
                                    const PlanLists& host, hipStream_t s);

}  // namespace kf
