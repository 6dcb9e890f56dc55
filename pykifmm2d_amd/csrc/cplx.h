// cplx.h -- near field, P2M, L2P and target evaluation of a plan of a complex kernel
// (cplx.hip).  The kernel G = Gr + i Gi is a pair of tables sharing u = log2 r^2; densities are
// real; slab rows are [re | im], Ph doubles each (P = 2 Ph); the near field and the potentials
// are complex, interleaved (re, im).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kf_common.h"

namespace kf {

struct CplxArgs {
    const int32_t *p2m, *rid, *bop, *bot, *top, *lvl, *leaves, *leaf_coll, *perm;
    const double *cx, *cy, *xs, *ys, *tau, *surf, *Psi;
    double *U, *near;  // near: 2 R doubles per point, tree order
    int n_p2m, n_leaves, p, P, R, nvalid;
};

// which: 0 = P2M into U, 1 = near field of the leaves [lo, hi) into `near`, 2 = L2P plus the near
// field into the potentials out[r] (complex, input order; r < nvalid).
void launch_cplx(const KfKernel& ker, const CplxArgs& a, int which, int lo, int hi, double* const* out,
                 hipStream_t s);

// Potentials (complex) of the expansion at m arbitrary targets, nvalid right-hand sides:
// pot[2 (r m + i) + c].  loc: each target's leaf (k_locate).
void launch_cplx_targets(const KfKernel& ker, const CplxArgs& a, const double* tx, const double* ty, int64_t m,
                         const int32_t* loc, const int32_t* coll, int check, double* pot, hipStream_t s);

}  // namespace kf
