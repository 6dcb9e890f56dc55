// batch_mfma.h -- near field, P2M and L2P of batches of 16, 32 or 64 right-hand sides on the
// FP64 MFMA (batch_mfma.hip, a translation unit of its own: it is compiled with the MFMA
// accumulators in vector registers, which keeps them out of the accumulation registers across
// the branches of the kernel evaluation).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kf_common.h"

namespace kf {

// Device arrays of one plan that the batched passes read and write (tree order; rows of R
// doubles per point in tau / near, R rows of P per box in U / Psi).
struct BatchMfmaArgs {
    const int32_t *p2m, *rid, *bop, *bot, *top, *lvl, *leaves, *leaf_coll;
    const double *cx, *cy, *xs, *ys, *tau, *surf, *Psi;
    double *U, *near;
    int n_p2m, n_leaves, p, P, R;
};

// which: 0 = P2M (U from tau), 1 = near field of the leaves [lo, hi) into `near` (near_grid > 0:
// that many workgroups looping over the leaves), 2 = L2P added to `near` (leaves at level >= 2).
void launch_batch_mfma(const KfKernel& ker, const BatchMfmaArgs& a, int which, int lo, int hi, int near_grid,
                       hipStream_t s);

}  // namespace kf
