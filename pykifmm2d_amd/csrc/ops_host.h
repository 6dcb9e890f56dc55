// ops_host.h -- translation operators of one tree level, built in FP64 on the host.
#pragma once
#include <cstdint>
#include <vector>

#include "kf_common.h"
#include "plan_lists.h"

namespace kf {

// Operators of level l, each P x P row-major with zero padding beyond p (P = p rounded up
// to a multiple of 16).  Layout of the block (in doubles):
//   [0, 40 P^2)        M2LT[o][i][j] = G(small_i + o*w, small_j)       (reference fmm.py:298-312)
//   [40 P^2, 44 P^2)   L2LB[q][i][j] = M2MC_{l-1}[i][q p + j]           (transpose of fmm.py:374)
//   [44 P^2, 48 P^2)   M2MB[q][i][j] = M2MC_l[j][q p + i]               (fmm.py:280-297, 360)
//   [48 P^2, 49 P^2)   F[j][m]  real orthonormal Fourier basis of the circulant C2E matrix
//   [49 P^2, 50 P^2)   FT[m][j] = F[j][m]
//   [50 P^2, 50 P^2+P) 1/lambda_m (0 in the padding)
// so that E_l^{-1} = F diag(1/lambda) F^T (the C2E LU of fmm.py:274-279 in factored form).
struct LevelOps {
    KF_HD static size_t size(int P) { return (size_t)50 * P * P + P; }
    KF_HD static size_t m2lt(int P, int o) { return (size_t)o * P * P; }
    KF_HD static size_t l2lb(int P, int q) { return (size_t)(40 + q) * P * P; }
    KF_HD static size_t m2mb(int P, int q) { return (size_t)(44 + q) * P * P; }
    KF_HD static size_t F(int P) { return (size_t)48 * P * P; }
    KF_HD static size_t FT(int P) { return (size_t)49 * P * P; }
    KF_HD static size_t invlam(int P) { return (size_t)50 * P * P; }
};

// Equivalent / check circles of a level (reference fmm.py:12-22): small radius
// r1 = w(sqrt2+0.01)/2, large radius r2 = w(4-sqrt2-0.02)/2, angles 2 pi k/p.
void level_circles(double width, int p, std::vector<double>& sx, std::vector<double>& sy,
                   std::vector<double>& lx, std::vector<double>& ly);

// Fills `out` (LevelOps::size(P) doubles) for level l of a tree with root width W.
// `has_parent` / `has_children` select whether the L2L / M2M blocks are formed.
void build_level_ops(const KfKernel& ker, double W, int l, int p, int P, bool has_parent,
                     bool has_children, double* out);

// The operators of a complex kernel (ker.cplx) in real form for slab rows [re | im] of
// P = 2 Ph doubles (Ph = p rounded up to a multiple of 16): every complex p x p block M becomes
// [[Re M, Im M], [-Im M, Re M]] (y = x M on rows), F is blockdiag(F, F), FT carries
// diag(1/lambda) F^T (lambda complex) and 1/lambda is 1 on the used columns, so the passes
// apply them as the real ones.
void build_level_ops_complex(const KfKernel& ker, double W, int l, int p, int Ph, bool has_parent,
                             bool has_children, double* out);

// Dense FP64 E (p x p, E[j][i] = G(small_i, large_j)) for tests of the factorisation.
void c2e_matrix(const KfKernel& ker, double width, int p, std::vector<double>& E);

}  // namespace kf
