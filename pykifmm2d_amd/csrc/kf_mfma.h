// kf_mfma.h -- register-resident row GEMMs on CDNA MFMA (16x16x4, FP64 and FP32).
//
// A wave owns a tile of 16 rows ("boxes"), each row a vector of P = 16*NB expansion
// coefficients.  Rows live in registers in the layout produced by v_mfma_*_16x16x4:
// lane l holds row n = l % 16 and, for each column block nb < NB, four columns col(nb, l/16, v)
// (v < 4) -- the output positions of the instruction:
//   FP32 (v_mfma_f32_16x16x4f32): col = 16*nb + 4*(l/16) + v   (contiguous 4-vector)
//   FP64 (v_mfma_f64_16x16x4f64): col = 16*nb + (l/16) + 4*v   (stride 4)
// (both operand layouts are A[i = l%16][k = l/16], B[k = l/16][j = l%16]; the output
// layouts were measured on gfx950).  A product Y = R * Op (Op = P x P in LDS, row-major) is
// formed as Y^T = Op^T R^T block by block: the MFMA A operand is Op read from LDS and the B
// operand is the register row data, with the reduction index of step (nb, v) taken as
// col(nb, l/16, v) -- the column the lane already holds.  The result comes back in the same
// layout, so products chain without any data movement.
#pragma once
#include <hip/hip_runtime.h>

namespace kf {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma_f64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma_f32(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// column held in register v of column block nb by lanes with l/16 = kq
__device__ __forceinline__ int rcol(double, int nb, int kq, int v) { return 16 * nb + kq + 4 * v; }
__device__ __forceinline__ int rcol(float, int nb, int kq, int v) { return 16 * nb + 4 * kq + v; }

template <int NB, class V>
__device__ __forceinline__ void rows_zero(V (&r)[NB]) {
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) r[nb] = V{0, 0, 0, 0};
}

// acc += R * Op, Op in LDS with row stride ld (elements).
template <int NB>
__device__ __forceinline__ void rows_gemm(d4 (&acc)[NB], const d4 (&R)[NB], const double* Op, int ld) {
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double* orow = Op + (size_t)rcol(0.0, nb, kq, v) * ld + r;
            double b = R[nb][v];
#pragma unroll
            for (int mb = 0; mb < NB; ++mb) acc[mb] = mfma_f64(orow[16 * mb], b, acc[mb]);
        }
}

template <int NB>
__device__ __forceinline__ void rows_gemm(f4 (&acc)[NB], const f4 (&R)[NB], const float* Op, int ld) {
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float* orow = Op + (size_t)rcol(0.0f, nb, kq, v) * ld + r;
            float b = R[nb][v];
#pragma unroll
            for (int mb = 0; mb < NB; ++mb) acc[mb] = mfma_f32(orow[16 * mb], b, acc[mb]);
        }
}

// Load / add / store a row of a [rows][P] slab in the register layout (row < 0: zeros).
template <int NB, class T, class V>
__device__ __forceinline__ void rows_load(V (&R)[NB], const T* __restrict__ slab, long row) {
    const int kq = (threadIdx.x & 63) >> 4;
    constexpr int P = 16 * NB;
    if (row >= 0) {  // one branch region for the row (not one per element)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v) R[nb][v] = slab[row * P + rcol(T(0), nb, kq, v)];
    } else {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) R[nb] = V{0, 0, 0, 0};
    }
}
template <int NB, class T, class V>
__device__ __forceinline__ void rows_add(V (&R)[NB], const T* __restrict__ slab, long row) {
    const int kq = (threadIdx.x & 63) >> 4;
    constexpr int P = 16 * NB;
    if (row < 0) return;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v) R[nb][v] += slab[row * P + rcol(T(0), nb, kq, v)];
}
template <int NB, class T, class V>
__device__ __forceinline__ void rows_store(const V (&R)[NB], T* __restrict__ slab, long row) {
    const int kq = (threadIdx.x & 63) >> 4;
    constexpr int P = 16 * NB;
    if (row < 0) return;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v) slab[row * P + rcol(T(0), nb, kq, v)] = R[nb][v];
}

// Column scale R[:, c] *= s[c].
template <int NB, class T, class V>
__device__ __forceinline__ void rows_scale(V (&R)[NB], const T* __restrict__ s) {
    const int kq = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v) R[nb][v] *= s[rcol(T(0), nb, kq, v)];
}

// Column-split products: wave w of a workgroup forms column block w (16 columns) of
// Y = R * Op for all its rows, so a short row tile still spreads over NB waves.  The wave's
// MFMA A operands -- Op[rcol(nb, kq, v)][16 w + l%16] for every step (nb, v) -- are held in
// registers (cols_load), and the steps run in the same (nb, v) order as rows_gemm, so the
// column block is bitwise the one rows_gemm produces.
template <int NB>
__device__ __forceinline__ void cols_load(d4 (&A)[NB], const double* __restrict__ Op, int w) {
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
    constexpr int P = 16 * NB;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v) A[nb][v] = Op[(size_t)rcol(0.0, nb, kq, v) * P + 16 * w + r];
}

// acc (column block of the operator slice A) += R * Op.
template <int NB>
__device__ __forceinline__ void cols_gemm(d4& acc, const d4 (&R)[NB], const d4 (&A)[NB]) {
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc = mfma_f64(A[nb][v], R[nb][v], acc);
}

// Column block w of a slab row (row < 0: zeros), its store, and its column scale.
template <int NB>
__device__ __forceinline__ void col_load(d4& a, const double* __restrict__ slab, long row, int w) {
    const int kq = (threadIdx.x & 63) >> 4;
    constexpr int P = 16 * NB;
    if (row >= 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) a[v] = slab[row * P + rcol(0.0, w, kq, v)];
    } else {
        a = d4{0, 0, 0, 0};
    }
}
template <int NB>
__device__ __forceinline__ void col_add(d4& a, const double* __restrict__ slab, long row, int w) {
    const int kq = (threadIdx.x & 63) >> 4;
    constexpr int P = 16 * NB;
    if (row < 0) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) a[v] += slab[row * P + rcol(0.0, w, kq, v)];
}
template <int NB, class T>
__device__ __forceinline__ void col_store(const d4& a, T* __restrict__ slab, long row, int w) {
    const int kq = (threadIdx.x & 63) >> 4;
    constexpr int P = 16 * NB;
    if (row < 0) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) slab[row * P + rcol(0.0, w, kq, v)] = (T)a[v];
}
__device__ __forceinline__ void col_scale(d4& a, const double* __restrict__ s, int w) {
    const int kq = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int v = 0; v < 4; ++v) a[v] *= s[rcol(0.0, w, kq, v)];
}

// Exchange through LDS: every wave writes its column block of the tile's 16 rows (block rb of
// the tile), then -- after a barrier -- reads whole rows back in the register row layout.
// The row stride is odd (cols_lds_ld): a 64-bit LDS access serves 16 lanes at a time over 32
// four-byte banks, and the 16 lanes of one access hold 16 consecutive rows, which an odd
// stride puts on 16 distinct bank pairs (a stride of P + 4 put every fourth row on one pair).
template <int NB>
constexpr int cols_lds_ld() {
    return 16 * NB + 1;
}
template <int NB>
__device__ __forceinline__ void cols_to_lds(double* Us, const d4& a, int rb, int w) {
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
    constexpr int LD = cols_lds_ld<NB>();
#pragma unroll
    for (int v = 0; v < 4; ++v) Us[(16 * rb + r) * LD + rcol(0.0, w, kq, v)] = a[v];
}
// A whole row (all NB column blocks, register layout) into row 16 rb + l%16 of an LDS tile of
// stride cols_lds_ld.
template <int NB>
__device__ __forceinline__ void rows_to_lds(double* Us, const d4 (&R)[NB], int rb) {
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
    constexpr int LD = cols_lds_ld<NB>();
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v) Us[(16 * rb + r) * LD + rcol(0.0, nb, kq, v)] = R[nb][v];
}
template <int NB>
__device__ __forceinline__ void rows_from_lds(d4 (&R)[NB], const double* Us, int rb) {
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
    constexpr int LD = cols_lds_ld<NB>();
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 4; ++v) R[nb][v] = Us[(16 * rb + r) * LD + rcol(0.0, nb, kq, v)];
}

// Register staging of one P x P operator for a workgroup of TH threads (256 by default): load() issues the
// 16-byte global reads (so they are in flight while the previous operator is in use), store()
// writes them to LDS.  Thread t takes units t, t + 256, ...: the lanes of one 128-bit LDS write
// cover consecutive 16-byte units, free of bank conflicts.  Operators are 16-byte aligned in HBM
// and LDS.
template <int NB, class T, int TH = 256>
struct OpStage {
    static constexpr int P = 16 * NB;
    static constexpr int N16 = P * P * (int)sizeof(T) / 16;
    static constexpr int PER = (N16 + TH - 1) / TH;
    typedef double d2 __attribute__((ext_vector_type(2)));
    d2 v[PER];
    __device__ __forceinline__ void load(const T* __restrict__ src) {
        const d2* s = reinterpret_cast<const d2*>(src);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            int i = threadIdx.x + TH * k;
            if (N16 % TH == 0 || i < N16) v[k] = s[i];
        }
    }
    __device__ __forceinline__ void store(T* __restrict__ dst) const {
        d2* d = reinterpret_cast<d2*>(dst);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            int i = threadIdx.x + TH * k;
            if (N16 % TH == 0 || i < N16) d[i] = v[k];
        }
    }
};

}  // namespace kf
