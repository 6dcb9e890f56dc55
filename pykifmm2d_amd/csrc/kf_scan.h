// kf_scan.h -- device-wide exclusive scans and order-preserving compaction of 32-bit values.
//
// A scan of n values runs in three launches: per-tile sums (SCAN_TILE values per workgroup),
// one workgroup scanning the tile sums, and a per-tile scan that adds its tile's offset.  The
// tile scan is a wave scan (64 lanes, shuffles) over per-thread serial sums.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace kf {

static constexpr int SCAN_THREADS = 256, SCAN_ITEMS = 16, SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread over a workgroup of SCAN_THREADS threads; returns the
// thread's exclusive prefix and leaves the workgroup total in *total.
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan_u32(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
        const uint32_t s = wsum[w];
        before += w < wave ? s : 0u;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

// part[b] = sum of in[b * SCAN_TILE .. (b + 1) * SCAN_TILE).
static __global__ void __launch_bounds__(SCAN_THREADS) k_scan_tile_sums(const uint32_t* __restrict__ in, int64_t n,
                                                                  uint32_t* __restrict__ part) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) s += base + k < n ? in[base + k] : 0u;
    uint32_t total;
    (void)block_excl_scan_u32(s, &total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// Exclusive scan of part[0 .. m) in one workgroup (m = number of tiles); part[m] = total.
static __global__ void __launch_bounds__(SCAN_THREADS) k_scan_top(uint32_t* __restrict__ part, int m) {
    const int per = (m + SCAN_THREADS - 1) / SCAN_THREADS;
    const int beg = threadIdx.x * per, end = min(beg + per, m);
    uint32_t s = 0;
    for (int i = beg; i < end; ++i) s += part[i];
    uint32_t total;
    uint32_t run = block_excl_scan_u32(s, &total);
    for (int i = beg; i < end; ++i) {
        const uint32_t v = part[i];
        part[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) part[m] = total;
}

// out = exclusive scan of in (in == out allowed), tile offsets from part.
static __global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(const uint32_t* in, int64_t n,
                                                              const uint32_t* __restrict__ part, uint32_t* out) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_excl_scan_u32(s, &total) + part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) {
            out[base + k] = run;
            run += v[k];
        }
}

// Scratch words scan_u32 needs for n values.
inline int64_t scan_scratch(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

// Exclusive scan of n values in place or into out; part: scan_scratch(n) words.  The total is
// left in part[tiles].  Stream-ordered, no synchronisation.
inline void scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* part, hipStream_t s) {
    if (n <= 0) return;
    const int m = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
    k_scan_tile_sums<<<m, SCAN_THREADS, 0, s>>>(in, n, part);
    k_scan_top<<<1, SCAN_THREADS, 0, s>>>(part, m);
    k_scan_tiles<<<m, SCAN_THREADS, 0, s>>>(in, n, part, out);
}

// Position of the scan total in the scratch of scan_u32 over n values.
inline int64_t scan_total_index(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

}  // namespace kf
