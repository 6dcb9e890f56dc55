// kf_sort.h -- stable LSD radix sort of (uint64 key, int32 value) pairs, 8-bit digits.
//
// Per pass: a digit histogram per tile of RS_TILE keys (k_rs_hist, digit-major so that one
// exclusive scan gives every tile's output offset per digit), the scan (kf_scan.h) and a stable
// scatter (k_rs_scatter: lanes of a wave holding the same digit are ranked by ballots).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "kf_scan.h"

namespace kf {

static constexpr int RS_THREADS = 256, RS_ITEMS = 16, RS_TILE = RS_THREADS * RS_ITEMS;
static constexpr int RS_WAVES = RS_THREADS / 64;

static __global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                        uint32_t* __restrict__ hist, int nblk) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int64_t i = base + r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

static __global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint64_t* __restrict__ kin,
                                                           const int32_t* __restrict__ vin,
                                                           uint64_t* __restrict__ kout, int32_t* __restrict__ vout,
                                                           int64_t n, int shift, const uint32_t* __restrict__ offs,
                                                           int nblk) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[RS_WAVES][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    run[threadIdx.x] = offs[(size_t)threadIdx.x * nblk + blockIdx.x];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; ++r) {
        for (int w = 0; w < RS_WAVES; ++w) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        const int64_t i = base + r * RS_THREADS + threadIdx.x;
        const bool valid = i < n;
        const uint64_t key = valid ? kin[i] : 0;
        const uint32_t d = valid ? (uint32_t)((key >> shift) & 255u) : 256u;
        // lanes of this wave holding the same digit
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t before = 0;
            for (int w = 0; w < wave; ++w) before += wcnt[w][d];
            const uint32_t dst = run[d] + before + rank;
            kout[dst] = key;
            vout[dst] = vin[i];
        }
        __syncthreads();
        uint32_t tot = 0;
        for (int w = 0; w < RS_WAVES; ++w) tot += wcnt[w][threadIdx.x];
        run[threadIdx.x] += tot;
        __syncthreads();
    }
}

// Scratch of sort_pairs for n pairs: key and value buffers of n each, hist words.
inline int sort_tiles(int64_t n) { return (int)((n + RS_TILE - 1) / RS_TILE); }
inline int64_t sort_hist_words(int64_t n) {
    const int64_t h = 256 * (int64_t)sort_tiles(n);
    return h + scan_scratch(h);
}

// Sorts n pairs (keys, vals) by the low `bits` bits of the keys, stably; kb / vb: n-element
// scratch, hist: sort_hist_words(n) words.  Results end in keys / vals.  No synchronisation;
// returns the first error of the copies it enqueues (launch errors: hipGetLastError).
inline hipError_t sort_pairs(uint64_t* keys, int32_t* vals, int64_t n, int bits, uint64_t* kb, int32_t* vb,
                       uint32_t* hist, hipStream_t s) {
    if (n <= 1 || bits <= 0) return hipSuccess;
    const int nblk = sort_tiles(n);
    const int64_t hn = 256 * (int64_t)nblk;
    uint32_t* part = hist + hn;
    uint64_t* ka = keys;
    int32_t* va = vals;
    uint64_t* kx = kb;
    int32_t* vx = vb;
    const int passes = (bits + 7) / 8;
    for (int ps = 0; ps < passes; ++ps) {
        const int shift = 8 * ps;
        k_rs_hist<<<nblk, RS_THREADS, 0, s>>>(ka, n, shift, hist, nblk);
        scan_u32(hist, hist, hn, part, s);
        k_rs_scatter<<<nblk, RS_THREADS, 0, s>>>(ka, va, kx, vx, n, shift, hist, nblk);
        std::swap(ka, kx);
        std::swap(va, vx);
    }
    if (ka != keys) {
        hipError_t e = hipMemcpyAsync(keys, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        return hipMemcpyAsync(vals, va, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s);
    }
    return hipSuccess;
}

}  // namespace kf
