// plan_lists.h -- work lists of a planned evaluation, built on the host from box trees.
//
// Host-only C++ (no HIP): the library uploads these lists; the CPU index checker of the test
// suite replays every pass's indexing over them.
//
// A plan covers one tree or a forest of trees evaluated in the same launches.  Trees whose
// root boxes have the same width form a class and share operators; the operators of (class c,
// level l) have the op index c * L + l (L = the deepest tree's level count).  Boxes, points and
// slab rows of the trees are concatenated; downward targets of the same (level, class, child
// position) from every tree form one group, so small trees fill tiles together.
#pragma once
#include <cstdint>
#include <vector>

#include "tree_host.h"
#include "../../include/kifmm2d.h"

namespace kf {

// The 40 M2L offsets (ox, oy) in [-3,3]^2 with max(|ox|,|oy|) >= 2, x-major order.
int m2l_offset_index(int ox, int oy);
void m2l_offset_of(int idx, int& ox, int& oy);

// For a box at child position q (= 2*xbit + ybit within its parent) the 27 interaction-list
// slots: offsets of the non-adjacent children of the parent's colleagues.
struct VSlots {
    int ox[4][27], oy[4][27], op[4][27];
};
const VSlots& vslots();

// Which downward targets an expansion computes: boxes with sources (enough for potentials at
// the sources), or every real box of levels >= 2 (needed for arbitrary targets).
enum { KF_TGT_SOURCES = 0, KF_TGT_ALL = 1 };

// Downward targets of one (level, class, child position q) group: [first, first + n_src) have
// sources, [first + n_src, first + n_all) do not; op = class * L + level.
struct DownGroup {
    int first, n_src, n_all, q, level, op;
};

// A downward tile: targets [first, first + count) of one group (the layout of an int4; the
// last field is the group's op index).
struct Tile {
    int32_t first, count, q, op;
};

struct PlanLists {
    int L = 0;            // levels of the deepest tree
    int n_classes = 1;    // distinct root widths
    int n_trees = 1;
    std::vector<double> class_width;   // root width of each class
    std::vector<int64_t> pt_off, box_off;  // per tree, n_trees + 1 entries
    // boxes (global ids: tree offset + the tree's own id)
    std::vector<double> cx, cy;
    std::vector<int32_t> bot, top, lvl, bop, child0, coll;  // bop: op index; coll: 9 per box
    std::vector<uint8_t> leaf;
    // slab rows: boxes with sources at levels >= 1 take [0, n_src), the real boxes without
    // sources at levels >= 2 take [n_src, n_all); rid[box] = row or -1
    int64_t n_src = 0, n_all = 0;
    std::vector<int32_t> rid;
    // upward pass, per op index o: entries [up_off[o], up_off[o] + up_n[o]) of urow (rows),
    // boxes with children first (up_nm2m[o] of them) with their children's rows in crow (4 per
    // entry, -1: child without sources); p2m lists the boxes without children
    std::vector<int32_t> p2m, urow, crow;
    std::vector<int> up_off, up_n, up_nm2m;
    // downward pass: per target its row, its parent's row and 27 interaction-list source rows
    // (-1: unused), grouped by (level, class, child position)
    std::vector<int32_t> trow, prow, vsrc;
    std::vector<DownGroup> groups;
    // near field / L2P: real leaves with points and their 9 colleagues (box ids, -1 padded)
    std::vector<int32_t> leaves, lcoll;
    int64_t aliased_pairs = 0;
    kf_plan_stats stats{};
    // list lengths of device-built lists (plan_dev.h), whose arrays stay on the device
    int64_t box_count = 0, p2m_count = 0, leaf_count = 0, up_count = 0, dn_count = 0;
};

// Block offsets of the stored kernel matrices (kf_plan_store_matrices), one entry per block and
// the total last: near field nt x ns per leaf of the leaf list (ns: its own and its other
// colleagues' sources), L2P p x nt per leaf at level >= 2 (0 above), P2M p x ns per box of the
// p2m list.
struct StoredOffsets {
    std::vector<int64_t> near, l2p, p2m;
};
void stored_offsets(const PlanLists& ls, int p, StoredOffsets& o);

// Lists of one tree (op index = level, one class).
void build_plan_lists(const HostTree& h, int p, int P, PlanLists& out);

// Lists of a forest: trees[i] in class cls[i] (root width class_width[cls[i]]).
void build_forest_lists(const std::vector<const HostTree*>& trees, int p, int P, PlanLists& out);

// Downward tiles for batches of 2^lg right-hand sides (64 >> lg targets per tile) over the
// target set `mode`; off[o], n[o]: the tiles of op index o (contiguous); masks (unless null):
// slots any target of a tile uses (needs the host vsrc).
void build_tile_table(const PlanLists& pl, int lg, int mode, std::vector<Tile>& tiles,
                      std::vector<uint32_t>* masks, std::vector<int>& off, std::vector<int>& n);
int64_t tile_count(const PlanLists& pl, int lg, int mode);

}  // namespace kf
