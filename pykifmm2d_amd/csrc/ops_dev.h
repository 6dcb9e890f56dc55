// ops_dev.h -- translation operators of a plan built on the GPU (ops_dev.hip).
#pragma once
#include <vector>

#include "kf_internal.h"

namespace kf {

// The operators of every (class c, level l), op index c * L + l, in the LevelOps layout
// (ops_host.h) into `ops`, and the large-circle points of every op into `surf` (x then y, P
// doubles each).  Synchronises the stream.
void build_ops_dev(const KfKernel& ker, const std::vector<double>& class_width, int L, int p, int P,
                   DevBuf<double>& ops, DevBuf<double>& surf, hipStream_t s);

// Split precision: the 40 M2L blocks of every op in FP32, scaled to unit max-norm (scale[op]).
void build_ops32_dev(const DevBuf<double>& ops, int n_ops, int L, int P, DevBuf<float>& ops32, DevBuf<double>& scale,
                     hipStream_t s);

// FP16-split precision: the same scaled blocks as two FP16 planes (hi, lo) per block, each the
// transposed block [m][k] with k padded to PK = P rounded up to 32 (scale from build_ops32_dev).
void build_ops16_dev(const DevBuf<double>& ops, int n_ops, int L, int P, const DevBuf<double>& scale,
                     DevBuf<_Float16>& ops16, hipStream_t s);

}  // namespace kf
