// near_sym.h -- the near field of one right-hand side with each mutual-colleague pair of leaves
// evaluated once (near_sym.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kf_common.h"

namespace kf {

// Leaves of at most this many points can take the symmetric form (a wave holds a leaf's
// targets four to a lane group).
constexpr int KF_SYM_MAX_LEAF = 64;
// Partial-sum slots per point: 0 = the leaf's own row sums, 1 = the column sums of its own
// block, 2 + j = the column sums its colleague in slot j wrote as the pair's owner.
constexpr int KF_SYM_SLOTS = 11;

struct NearSymArgs {
    const int32_t *leaves, *lcoll, *pinfo, *bot, *top;
    const double *xs, *ys, *tau;
    double* near;   // N potentials, then KF_SYM_SLOTS partial slots of N each (slot s at near + (1 + s) N)
    int64_t N;
    int n_leaves;
};

// pinfo[9 li + j] for leaf li's colleague slot j: (leaf index of the colleague) * 16 + (slot of
// leaf li in the colleague's list) when the colleague is a leaf with points other than li
// itself, else -1.  max_leaf receives the most points of any leaf.  Synchronises the stream.
void build_leaf_pairs(const int32_t* leaves, int n_leaves, const int32_t* lcoll, const int32_t* bot,
                      const int32_t* top, int64_t n_boxes, int32_t* pinfo, int* max_leaf, hipStream_t s);

// Near field of the leaves [lo, hi) into the partial slots (grid > 0: that many workgroups
// looping over the leaves); with sum, then the fixed-order sum of every point's slots into
// a.near (after every leaf's launch: a pair's owner may sit in either range).
void launch_near_sym(const KfKernel& ker, const NearSymArgs& a, int lo, int hi, bool sum, int grid, hipStream_t s);

}  // namespace kf
