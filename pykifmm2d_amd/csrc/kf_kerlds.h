// kf_kerlds.h -- LDS staging shared by the kernels that evaluate the interaction kernel.
#pragma once
#include "kf_common.h"

namespace kf {

// LDS copy of the table-driven log's table (kf_logt.h) for the kernels that evaluate the
// kernel (the Laplace log, and the log of K0's power-series branch); every thread of the
// workgroup takes part, and the caller's next barrier publishes it.
template <int KIND>
constexpr int logt_lds() {
    return 2 << KF_LOGT_BITS;
}
template <int KIND>
__device__ __forceinline__ void stage_logt(double* lds) {
    for (int i = threadIdx.x; i < (2 << KF_LOGT_BITS); i += blockDim.x) lds[i] = KF_LOGT_TAB[i];
}
// Holds the table-driven log's leading polynomial coefficient in a vector register for the
// loops that follow: the first Horner step then reads one scalar operand (the most a VOP3
// instruction takes) instead of copying a coefficient into a vector register per evaluation.
template <int KIND>
__device__ __forceinline__ void pin_coef(KerT<KIND>& ker) {
    asm volatile("" : "+v"(ker.lc.q[KF_LOGT_QN - 1]));
}

}  // namespace kf
