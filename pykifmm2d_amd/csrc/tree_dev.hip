// tree_dev.hip -- the quadtree's box structure on the GPU (reference pykifmm2d/tree.py:128-576).
//
// Input: each tree's points sorted by exact Morton key (tree_gpu.hip).  A box at level l is a
// run of equal 2l-bit key prefixes, so every step below works on keys, never on coordinates.
//
// 1. Splitting by count (tree.py:499-506).  The box at level l that starts at point i holds
//    more than `cutoff` points iff point i + cutoff shares its level-l prefix; a point starts a
//    box at every level where its prefix differs from the previous point's.  Each point thus
//    yields the range of levels at which it starts a box that is split; the (tree, level, start)
//    triples, sorted, are the split boxes of every level in Morton order, and their children the
//    next level's boxes in the reference's order.  Neighbour tables (nb: the same-level box at
//    each of the 3x3 offsets, or -1) follow level by level from the parents' tables.
// 2. Level restriction (tree.py:533-545, 432-444, 296-311).  The reference sweeps levels
//    1 .. L-3, splitting every leaf with a colleague of depth >= 2 and appending its children to
//    the next level, until a sweep splits nothing.  Within a sweep, the depths a level's check
//    reads are those the previous sweep left (splits at shallower levels only add depth-0 boxes
//    at deeper levels and deepen shallower boxes), so the leaves that exist when a sweep starts
//    are checked in one parallel pass; the sweep's new boxes are checked level by level as they
//    appear (k_cascade, one workgroup), in the reference's order.  Depths then rise along the
//    parents of the split boxes (k_depth_walk, integer atomicMax).
// 3. X-list (tree.py:546-557, 446-458): leaves with a deeper real colleague get four fake
//    children; fake children change no depth and no real colleague, so every level is decided in
//    one parallel pass.
// 4. Colleague slots (tree.py:349-375, 260-294).  The reference lists a level's colleagues in
//    level order when it has fewer than 100 boxes and through the parent's slots otherwise, and
//    re-lists a level after each change of it (each restriction step re-lists the level below,
//    split or not).  A slot order is therefore the level order at the nearest level up the chain
//    of listings that was listed in level order: the colleagues of a box at level l sort by (the
//    position of their ancestor at that level m, their Morton digits below m).  m per level comes
//    from the level sizes after each phase (host, chain_level).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>

#include "kf_internal.h"
#include "kf_scan.h"
#include "kf_sort.h"
#include "tree_dev.h"

namespace kf {

namespace {

constexpr int TB_THREADS = 256;
constexpr int KMAXL = 32;  // level slots per tree (key depth <= 30)

// ------------------------------------------------------------------ helpers ---------
__device__ __forceinline__ uint64_t spread2(uint64_t v) {
    v &= 0xffffffffull;
    v = (v | (v << 16)) & 0x0000ffff0000ffffull;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}
__device__ __forceinline__ uint32_t compact2(uint64_t v) {
    v &= 0x5555555555555555ull;
    v = (v | (v >> 1)) & 0x3333333333333333ull;
    v = (v | (v >> 2)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v >> 4)) & 0x00ff00ff00ff00ffull;
    v = (v | (v >> 8)) & 0x0000ffff0000ffffull;
    v = (v | (v >> 16)) & 0x00000000ffffffffull;
    return (uint32_t)v;
}
// Morton key of grid cell (x, y) at its level: x bit above y bit in each pair.
__device__ __forceinline__ uint64_t morton2(uint32_t x, uint32_t y) { return (spread2(x) << 1) | spread2(y); }

// first index in [lo, hi) with keys[i] >= v
__device__ __forceinline__ int64_t lower_bound_u64(const uint64_t* __restrict__ k, int64_t lo, int64_t hi, uint64_t v) {
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (k[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ int64_t lower_bound_i32(const int32_t* __restrict__ k, int64_t lo, int64_t hi, int32_t v) {
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (k[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
// t with off[t] <= i < off[t + 1]
__device__ __forceinline__ int seg_of(const int64_t* __restrict__ off, int nt, int64_t i) {
    int lo = 0, hi = nt - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The build pool: boxes in build order ("slots"), structure of arrays.
struct Pool {
    int32_t *tree, *lvl, *pos, *lid, *ix, *iy, *bot, *top, *parent, *child0, *depth, *nb;
    uint8_t *leaf, *fake, *xl;
};

// ------------------------------------------------------------------ 1. split by count --
// For point i of tree t: a = the first level at which i starts a box (its key prefix differs
// from the previous point's), b = the deepest level at which i and i + cutoff share their key
// prefix (-1: the tree has no point i + cutoff).  i starts a split box at the levels a .. b.
struct Witness {
    int t, a, b;
    bool deep;
};
__device__ __forceinline__ Witness witness(const uint64_t* __restrict__ keys, const int64_t* __restrict__ poff,
                                          const int32_t* __restrict__ Dt, int nt, int64_t i, int cutoff) {
    Witness w;
    w.t = seg_of(poff, nt, i);
    const int64_t lo = poff[w.t], hi = poff[w.t + 1];
    const int D = Dt[w.t];
    if (i == lo) {
        w.a = 0;
    } else {
        const uint64_t x = keys[i] ^ keys[i - 1];
        w.a = x ? D - ((63 - __builtin_clzll(x)) >> 1) : D + 1;
    }
    w.b = -1;
    w.deep = false;
    if (i + cutoff < hi) {
        const uint64_t y = keys[i] ^ keys[i + cutoff];
        if (y == 0) {  // more than `cutoff` points in one finest-level box
            w.deep = true;
            w.b = D - 1;
        } else {
            w.b = D - ((63 - __builtin_clzll(y)) >> 1) - 1;
        }
    }
    return w;
}

__global__ void k_witness(const uint64_t* __restrict__ keys, const int64_t* __restrict__ poff,
                          const int32_t* __restrict__ Dt, int nt, int64_t n, int cutoff, uint32_t* __restrict__ cnt,
                          int32_t* __restrict__ err) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Witness w = witness(keys, poff, Dt, nt, i, cutoff);
    if (w.deep) atomicOr(err, 1);
    cnt[i] = w.a <= w.b ? (uint32_t)(w.b - w.a + 1) : 0u;
}

// One (tree, level, start) entry per split box: key t * 32 + l, value the start point.
__global__ void k_emit_splits(const uint64_t* __restrict__ keys, const int64_t* __restrict__ poff,
                              const int32_t* __restrict__ Dt, int nt, int64_t n, int cutoff,
                              const uint32_t* __restrict__ eoff, uint64_t* __restrict__ skey,
                              int32_t* __restrict__ sval, uint32_t* __restrict__ cnt_tl) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Witness w = witness(keys, poff, Dt, nt, i, cutoff);
    if (w.a > w.b) return;
    uint32_t e = eoff[i];
    for (int l = w.a; l <= w.b; ++l, ++e) {
        skey[e] = (uint64_t)w.t * KMAXL + l;
        sval[e] = (int32_t)i;
        atomicAdd(&cnt_tl[w.t * KMAXL + l], 1u);
    }
}

// Boxes of the split-by-count phase.  Slot j: tree t (slots [p1off[t], p1off[t + 1])), local id
// j - p1off[t], level l by base1[t][l] <= lid, position lid - base1[t][l]; a box at level l >= 1
// is child q = pos & 3 of the (pos >> 2)-th split box of level l - 1.
__global__ void k_boxes1(Pool P, int64_t B1, const uint64_t* __restrict__ keys, const int64_t* __restrict__ poff,
                         const int32_t* __restrict__ Dt, const int32_t* __restrict__ Lt, int nt,
                         const int64_t* __restrict__ p1off, const int32_t* __restrict__ base1,
                         const int32_t* __restrict__ offSL, const int32_t* __restrict__ sval, int cutoff,
                         int32_t* __restrict__ err) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= B1) return;
    const int t = seg_of(p1off, nt, j);
    const int32_t lid = (int32_t)(j - p1off[t]);
    const int L = Lt[t];
    int l = 0;
    while (l + 1 < L && base1[t * KMAXL + l + 1] <= lid) ++l;
    const int32_t pos = lid - base1[t * KMAXL + l];
    const int D = Dt[t];
    const int64_t lo = poff[t], hi = poff[t + 1];
    uint32_t X = 0, Y = 0;
    int64_t b0 = lo, b1 = hi;
    if (l > 0) {
        const int r = pos >> 2, q = pos & 3;
        const int64_t s = sval[offSL[t * KMAXL + l - 1] + r];
        const uint64_t pre = ((keys[s] >> (2 * (D - (l - 1)))) << 2) | (uint64_t)q;  // the box's key prefix
        X = compact2(pre >> 1);
        Y = compact2(pre);
        const int sh = 2 * (D - l);
        b0 = lower_bound_u64(keys, s, hi, pre << sh);
        b1 = lower_bound_u64(keys, b0, hi, (pre + 1) << sh);
    } else {
        P.parent[j] = -1;
    }
    P.tree[j] = t;
    P.lvl[j] = l;
    P.pos[j] = pos;
    P.lid[j] = lid;
    P.ix[j] = (int32_t)X;
    P.iy[j] = (int32_t)Y;
    P.bot[j] = (int32_t)(b0 - lo);
    P.top[j] = (int32_t)(b1 - lo);
    P.depth[j] = 0;
    P.fake[j] = 0;
    P.xl[j] = 0;
    if (b1 - b0 > cutoff) {  // split: its rank among the level's split boxes gives its children
        const int64_t s0 = offSL[t * KMAXL + l], s1 = offSL[t * KMAXL + l + 1];
        const int64_t r = lower_bound_i32(sval, s0, s1, (int32_t)b0) - s0;
        if (r >= s1 - s0 || sval[s0 + r] != (int32_t)b0 || l + 1 >= L) {  // cannot happen: every such box is listed
            atomicOr(err, 4);
            P.child0[j] = -1;
            P.leaf[j] = 1;
            return;
        }
        const int32_t c0 = (int32_t)(p1off[t] + base1[t * KMAXL + l + 1] + 4 * r);
        P.child0[j] = c0;
        P.leaf[j] = 0;
        for (int q = 0; q < 4; ++q) P.parent[c0 + q] = (int32_t)j;
    } else {
        P.child0[j] = -1;
        P.leaf[j] = 1;
    }
}

// Neighbour table of a box from its parent's: the same-level box at offset (dx, dy), slot
// (dx + 1) * 3 + dy + 1 (slot 4 the box itself), or -1.
__device__ __forceinline__ void nb_from_parent(const Pool& P, int32_t c) {
    const int l = P.lvl[c];
    const int32_t p = P.parent[c];
    const int32_t X = P.ix[c], Y = P.iy[c];
    const int32_t side = (int32_t)(1u << l);
    const int32_t px = P.ix[p], py = P.iy[p];
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            int32_t r = -1;
            const int32_t NX = X + dx, NY = Y + dy;
            if (dx == 0 && dy == 0) {
                r = c;
            } else if (NX >= 0 && NY >= 0 && NX < side && NY < side) {
                const int ox = (NX >> 1) - px, oy = (NY >> 1) - py;
                const int32_t pn = P.nb[9 * (int64_t)p + (ox + 1) * 3 + oy + 1];
                if (pn >= 0) {
                    const int32_t c0 = P.child0[pn];
                    if (c0 >= 0) r = c0 + (((NX & 1) << 1) | (NY & 1));
                }
            }
            P.nb[9 * (int64_t)c + (dx + 1) * 3 + dy + 1] = r;
        }
}

__global__ void k_nb_level(Pool P, int64_t n, int l) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n || P.lvl[j] != l) return;
    if (l == 0) {
        for (int k = 0; k < 9; ++k) P.nb[9 * j + k] = k == 4 ? (int32_t)j : -1;
        return;
    }
    nb_from_parent(P, (int32_t)j);
}

// Depth rises along the parents of a box that has just been split (reference tree.py:423-430:
// depth = 1 + max child depth); integer maxima, so the order of the walks does not matter.
__device__ __forceinline__ void depth_walk(const Pool& P, int32_t b) {
    int d = 1;
    int32_t p = b;
    while (p >= 0 && d <= KMAXL) {  // a tree has at most 31 levels
        if (atomicMax(&P.depth[p], d) >= d) break;
        p = P.parent[p];
        ++d;
    }
}

__global__ void k_depth_split1(Pool P, int64_t n) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n || P.child0[j] < 0) return;
    depth_walk(P, (int32_t)j);
}

// The boxes split in the last sweep: list[0 .. *count) (count <= cap, the list's length).
__global__ void k_depth_list(Pool P, const int32_t* __restrict__ list, const int32_t* __restrict__ count, int64_t cap) {
    const int n = (int)min((int64_t)*count, cap);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) depth_walk(P, list[i]);
}

// ------------------------------------------------------------------ 2. restriction ----
// flag[j] = leaf j (real, 1 <= level <= L - 3) has a colleague of depth >= 2 (tree.py:432-444).
__global__ void k_flag_bads(Pool P, const int32_t* __restrict__ nbox, int64_t cap, const int32_t* __restrict__ Lt,
                            uint32_t* __restrict__ flag) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= cap) return;
    uint32_t f = 0;
    if (j < *nbox && P.leaf[j] && !P.fake[j]) {
        const int l = P.lvl[j];
        if (l >= 1 && l <= Lt[P.tree[j]] - 3)
            for (int k = 0; k < 9; ++k) {
                const int32_t c = P.nb[9 * j + k];
                if (k != 4 && c >= 0 && P.depth[c] >= 2) f = 1;
            }
    }
    flag[j] = f;
}

// flag[j] = real leaf j (1 <= level <= L - 2) has a real colleague of positive depth
// (tree.py:446-458; no fake box exists yet).
__global__ void k_flag_xlist(Pool P, const int32_t* __restrict__ nbox, int64_t cap, const int32_t* __restrict__ Lt,
                             uint32_t* __restrict__ flag) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= cap) return;
    uint32_t f = 0;
    if (j < *nbox && P.leaf[j] && !P.fake[j]) {
        const int l = P.lvl[j];
        if (l >= 1 && l <= Lt[P.tree[j]] - 2)
            for (int k = 0; k < 9; ++k) {
                const int32_t c = P.nb[9 * j + k];
                if (k != 4 && c >= 0 && P.depth[c] > 0) f = 1;
            }
    }
    flag[j] = f;
}

// list[idx[j]] = j for flagged j; keys by (level, tree) or (tree, level).
__global__ void k_compact_boxes(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ idx, int64_t n,
                                Pool P, int by_level, int tbits, int32_t* __restrict__ list,
                                uint64_t* __restrict__ key) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n || !flag[j]) return;
    const uint32_t o = idx[j];
    list[o] = (int32_t)j;
    const uint64_t t = (uint64_t)P.tree[j], l = (uint64_t)P.lvl[j];
    key[o] = by_level ? (l << tbits) | t : t * KMAXL + l;
}

// Block-wide exclusive scan for a workgroup of CT threads (CT a multiple of 64).
template <int CT>
__device__ __forceinline__ uint32_t cascade_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan_u32(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < CT / 64; ++w) {
        const uint32_t s = wsum[w];
        before += w < wave ? s : 0u;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

// One level-restriction sweep's splits, level by level (one workgroup).  old: the leaves found
// bad when the sweep started, sorted by (level, tree, position); the leaves the sweep creates
// at a level are checked when they appear and split at the next step.  Children go to slots
// from *nbox on, local ids from cnt_box[t] and positions from nlev[t][l + 1] on, in the order of
// their parents (per tree: old bads by position, then new ones by position).
constexpr int CT = 1024;
__global__ void __launch_bounds__(CT) k_cascade(Pool P, const int32_t* __restrict__ Lt, int Lmax, int tbits,
                                                const int32_t* __restrict__ old, const uint64_t* __restrict__ okey,
                                                int m, int32_t* __restrict__ merged, int32_t* __restrict__ newA,
                                                int32_t* __restrict__ newB, int32_t* __restrict__ inc,
                                                int32_t* __restrict__ split, int32_t* __restrict__ nsplit,
                                                int32_t* __restrict__ nbox, int32_t* __restrict__ cnt_box,
                                                int32_t* __restrict__ nlev, int64_t cap, int32_t* __restrict__ err) {
    __shared__ uint32_t wsum[CT / 64];
    __shared__ int s_nbox, s_nsplit, s_nnew, s_o0, s_o1, s_bad;
    if (threadIdx.x == 0) {
        s_nbox = *nbox;
        s_nsplit = 0;
        s_nnew = 0;
        s_bad = 0;
    }
    __syncthreads();
    int32_t* cur = newA;
    int32_t* nxt = newB;
    for (int ind = 1; ind <= Lmax - 3; ++ind) {
        if (threadIdx.x == 0) {
            int64_t lo = 0, hi = m;  // old bads of this level: keys in [ind << tbits, (ind + 1) << tbits)
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if ((okey[mid] >> tbits) < (uint64_t)ind) lo = mid + 1;
                else hi = mid;
            }
            s_o0 = (int)lo;
            hi = m;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if ((okey[mid] >> tbits) < (uint64_t)(ind + 1)) lo = mid + 1;
                else hi = mid;
            }
            s_o1 = (int)lo;
        }
        __syncthreads();
        const int o0 = s_o0, nold = s_o1 - s_o0, nnew = s_nnew, cnt = nold + nnew;
        const int nb0 = s_nbox;
        if (cnt == 0) {
            __syncthreads();
            if (threadIdx.x == 0) s_nnew = 0;
            __syncthreads();
            continue;
        }
        if ((int64_t)nb0 + 4 * (int64_t)cnt > cap) {
            // the pool is full: stop here (uniform: every thread reads the same values), leaving
            // the counts of the steps done, and the host retries with a larger pool
            if (threadIdx.x == 0) atomicOr(err, 2);
            break;
        }
        // merge old (by tree, position) and new (by tree, position; after the old of their tree)
        for (int k = threadIdx.x; k < cnt; k += CT) {
            int32_t b;
            int64_t r;
            if (k < nold) {
                b = old[o0 + k];
                const int t = P.tree[b];
                int64_t lo = 0, hi = nnew;  // new bads of trees < t
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (P.tree[cur[mid]] < t) lo = mid + 1;
                    else hi = mid;
                }
                r = k + lo;
            } else {
                b = cur[k - nold];
                const int t = P.tree[b];
                int64_t lo = 0, hi = nold;  // old bads of trees <= t
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if ((int)(okey[o0 + mid] & ((1ull << tbits) - 1)) <= t) lo = mid + 1;
                    else hi = mid;
                }
                r = (k - nold) + lo;
            }
            merged[r] = b;
        }
        __threadfence();
        __syncthreads();
        // children of the merged bads
        for (int k = threadIdx.x; k < cnt; k += CT) {
            const int32_t b = merged[k];
            const int t = P.tree[b];
            int64_t lo = 0, hi = k;  // first merged entry of tree t
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (P.tree[merged[mid]] < t) lo = mid + 1;
                else hi = mid;
            }
            const int first = (int)lo;
            lo = k + 1;
            hi = cnt;  // one past the last entry of tree t
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (P.tree[merged[mid]] <= t) lo = mid + 1;
                else hi = mid;
            }
            const int rank = k - first;
            inc[k] = k == (int)lo - 1 ? (int)lo - first : 0;
            const int32_t lid0 = cnt_box[t] + 4 * rank;
            const int32_t pos0 = nlev[t * KMAXL + ind + 1] + 4 * rank;
            const int32_t c0 = nb0 + 4 * k;
            for (int q = 0; q < 4; ++q) {
                const int32_t c = c0 + q;
                P.tree[c] = t;
                P.lvl[c] = ind + 1;
                P.pos[c] = pos0 + q;
                P.lid[c] = lid0 + q;
                P.ix[c] = 2 * P.ix[b] + (q >> 1);
                P.iy[c] = 2 * P.iy[b] + (q & 1);
                P.bot[c] = 0;
                P.top[c] = 0;
                P.parent[c] = b;
                P.child0[c] = -1;
                P.depth[c] = 0;
                P.leaf[c] = 1;
                P.fake[c] = 0;
                P.xl[c] = 0;
            }
            P.child0[b] = c0;
            P.leaf[b] = 0;
            split[s_nsplit + k] = b;
        }
        __threadfence();
        __syncthreads();
        for (int k = threadIdx.x; k < cnt; k += CT)
            if (inc[k]) {
                const int t = P.tree[merged[k]];
                cnt_box[t] += 4 * inc[k];
                nlev[t * KMAXL + ind + 1] += 4 * inc[k];
            }
        for (int c = nb0 + threadIdx.x; c < nb0 + 4 * cnt; c += CT) nb_from_parent(P, c);
        __threadfence();
        __syncthreads();
        // the existing boxes next to a new one list it
        for (int c = nb0 + threadIdx.x; c < nb0 + 4 * cnt; c += CT)
            for (int k = 0; k < 9; ++k) {
                const int32_t e = P.nb[9 * (int64_t)c + k];
                if (k != 4 && e >= 0 && e < nb0) P.nb[9 * (int64_t)e + 8 - k] = c;
            }
        __threadfence();
        __syncthreads();
        // new leaves that are bad at level ind + 1, in order
        uint32_t run = 0;
        for (int base = 0; base < 4 * cnt; base += CT) {
            const int i = base + threadIdx.x;
            uint32_t f = 0;
            int32_t c = -1;
            if (i < 4 * cnt) {
                c = nb0 + i;
                if (ind + 1 <= Lt[P.tree[c]] - 3)
                    for (int k = 0; k < 9; ++k) {
                        const int32_t e = P.nb[9 * (int64_t)c + k];
                        if (k != 4 && e >= 0 && P.depth[e] >= 2) f = 1;
                    }
            }
            uint32_t total;
            const uint32_t o = cascade_scan<CT>(f, wsum, &total);
            if (f) nxt[run + o] = c;
            run += total;
        }
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            s_nbox = nb0 + 4 * cnt;
            s_nsplit += cnt;
            s_nnew = (int)run;
        }
        __syncthreads();
        int32_t* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (threadIdx.x == 0) {
        *nbox = s_nbox;
        *nsplit = s_nsplit;
    }
}

// ------------------------------------------------------------------ 3. X-list ---------
// Fake children of the X-list leaves xlist[0 .. m), sorted by (tree, level, position): local
// ids from cnt_box[t] on in that order, positions from nlev[t][l + 1] on per level.
__global__ void k_make_fakes(Pool P, const int32_t* __restrict__ xlist, const uint64_t* __restrict__ xkey, int m,
                             int32_t nb0, const int32_t* __restrict__ cnt_box, const int32_t* __restrict__ nlev,
                             int32_t* __restrict__ inc_t, int32_t* __restrict__ inc_tl) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int32_t b = xlist[k];
    const uint64_t key = xkey[k];
    const int t = (int)(key / KMAXL), l = (int)(key % KMAXL);
    int lo = 0, hi = k;  // first entry of tree t
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (xkey[mid] / KMAXL < (uint64_t)t) lo = mid + 1;
        else hi = mid;
    }
    const int ft = lo;
    lo = ft;
    hi = k;  // first entry of (t, l)
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (xkey[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    const int ftl = lo;
    const bool last_tl = k + 1 == m || xkey[k + 1] != key;
    const bool last_t = k + 1 == m || xkey[k + 1] / KMAXL != (uint64_t)t;
    inc_tl[k] = last_tl ? k - ftl + 1 : 0;
    inc_t[k] = last_t ? k - ft + 1 : 0;
    const int32_t lid0 = cnt_box[t] + 4 * (k - ft);
    const int32_t pos0 = nlev[t * KMAXL + l + 1] + 4 * (k - ftl);
    const int32_t c0 = nb0 + 4 * k;
    for (int q = 0; q < 4; ++q) {
        const int32_t c = c0 + q;
        P.tree[c] = t;
        P.lvl[c] = l + 1;
        P.pos[c] = pos0 + q;
        P.lid[c] = lid0 + q;
        P.ix[c] = 2 * P.ix[b] + (q >> 1);
        P.iy[c] = 2 * P.iy[b] + (q & 1);
        P.parent[c] = b;
        P.child0[c] = -1;
        P.depth[c] = 0;
        P.leaf[c] = 0;
        P.fake[c] = 1;
        P.xl[c] = 0;
    }
    P.child0[b] = c0;
    P.xl[b] = 1;
}

__global__ void k_fake_counts(Pool P, const int32_t* __restrict__ xlist, const uint64_t* __restrict__ xkey, int m,
                              const int32_t* __restrict__ inc_t, const int32_t* __restrict__ inc_tl,
                              int32_t* __restrict__ cnt_box, int32_t* __restrict__ nlev, int32_t* __restrict__ nbox) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const uint64_t key = xkey[k];
    const int t = (int)(key / KMAXL), l = (int)(key % KMAXL);
    if (inc_t[k]) cnt_box[t] += 4 * inc_t[k];
    if (inc_tl[k]) nlev[t * KMAXL + l + 1] += 4 * inc_tl[k];
    if (k == 0) *nbox += 4 * m;
}

__global__ void k_nb_range(Pool P, int32_t c0, int32_t c1) {
    const int32_t c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c1) return;
    nb_from_parent(P, c);
}

__global__ void k_nb_announce(Pool P, int32_t c0, int32_t c1) {
    const int32_t c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c1) return;
    for (int k = 0; k < 9; ++k) {
        const int32_t e = P.nb[9 * (int64_t)c + k];
        if (k != 4 && e >= 0 && e < c0) P.nb[9 * (int64_t)e + 8 - k] = c;
    }
}

// Point ranges of the boxes made after the split-by-count phase.
__global__ void k_ranges(Pool P, int32_t c0, int32_t c1, const uint64_t* __restrict__ keys,
                         const int64_t* __restrict__ poff, const int32_t* __restrict__ Dt) {
    const int32_t c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c1) return;
    const int t = P.tree[c], l = P.lvl[c], D = Dt[t];
    const int64_t lo = poff[t], hi = poff[t + 1];
    const uint64_t M = morton2((uint32_t)P.ix[c], (uint32_t)P.iy[c]);
    const int sh = 2 * (D - l);
    const int64_t b0 = lower_bound_u64(keys, lo, hi, M << sh);
    const int64_t b1 = lower_bound_u64(keys, b0, hi, (M + 1) << sh);
    P.bot[c] = (int32_t)(b0 - lo);
    P.top[c] = (int32_t)(b1 - lo);
}

// ------------------------------------------------------------------ 4. final layout ----
// anchor[c]: position of c's ancestor at the level m[t][l] its level's colleague order refers to.
__global__ void k_anchor(Pool P, int32_t n, const int32_t* __restrict__ mlev, int32_t* __restrict__ anchor) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int l = P.lvl[c];
    const int m = mlev[P.tree[c] * KMAXL + l];
    int32_t a = c;
    for (int k = l; k > m && a >= 0; --k) a = P.parent[a];
    anchor[c] = a >= 0 ? P.pos[a] : 0;
}

struct Final {
    int32_t *lvl, *ix, *iy, *bot, *top, *parent, *child0, *depth, *pos, *coll, *level_list;
    uint8_t *leaf, *fake, *xl;
};

__global__ void k_final(Pool P, int32_t n, const int64_t* __restrict__ box_off, const int64_t* __restrict__ lev_off,
                        const int32_t* __restrict__ mlev, const int32_t* __restrict__ anchor, Final F,
                        int32_t* __restrict__ err) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int t = P.tree[c], l = P.lvl[c];
    const int64_t g = box_off[t] + P.lid[c];
    const int64_t li = lev_off[t * KMAXL + l] + P.pos[c];
    if (P.lid[c] < 0 || g >= box_off[t + 1] || P.pos[c] < 0 || li >= lev_off[t * KMAXL + l + 1]) {
        atomicOr(err, 8);  // cannot happen: local ids and positions are dense per tree and level
        return;
    }
    F.lvl[g] = l;
    F.ix[g] = P.ix[c];
    F.iy[g] = P.iy[c];
    F.bot[g] = P.bot[c];
    F.top[g] = P.top[c];
    F.parent[g] = P.parent[c] >= 0 ? P.lid[P.parent[c]] : -1;
    F.child0[g] = P.child0[c] >= 0 ? P.lid[P.child0[c]] : -1;
    F.depth[g] = P.depth[c];
    F.pos[g] = P.pos[c];
    F.leaf[g] = P.leaf[c];
    F.fake[g] = P.fake[c];
    F.xl[g] = P.xl[c];
    F.level_list[li] = P.lid[c];
    // colleagues: the 3x3 same-level boxes by (anchor position, Morton digits below the anchor level)
    const int m = mlev[t * KMAXL + l];
    const uint64_t low = l > m ? (1ull << (2 * (l - m))) - 1 : 0ull;
    uint64_t ka[9], kb[9];
    int32_t id[9];
    int nc = 0;
    for (int k = 0; k < 9; ++k) {
        const int32_t e = P.nb[9 * (int64_t)c + k];
        if (e < 0) continue;
        uint64_t a, b;
        if (l == m) {
            a = (uint64_t)P.pos[e];
            b = 0;
        } else {
            a = (uint64_t)anchor[e];
            b = morton2((uint32_t)P.ix[e], (uint32_t)P.iy[e]) & low;
        }
        int i = nc++;
        while (i > 0 && (ka[i - 1] > a || (ka[i - 1] == a && kb[i - 1] > b))) {
            ka[i] = ka[i - 1];
            kb[i] = kb[i - 1];
            id[i] = id[i - 1];
            --i;
        }
        ka[i] = a;
        kb[i] = b;
        id[i] = P.lid[e];
    }
    for (int k = 0; k < 9; ++k) F.coll[9 * g + k] = k < nc ? id[k] : -1;
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned nblocks(int64_t n, int thr = TB_THREADS) { return (unsigned)std::max<int64_t>(1, (n + thr - 1) / thr); }

// The level m whose level order a listing of level l of tree t refers to, following the
// reference's listings (tree.py:523-529 initially, 533-545 and 296-311 in the restriction
// sweeps, 546-557 for the X-list): a listing of level l is in level order if l == 0 or the level
// then had fewer than 100 boxes, else it orders through the listing of level l - 1 current at
// that time.  n0, nR, nF: level sizes after splitting by count, after the restriction, final.
enum { V_INIT = 0, V_RESTRICT = 1, V_XLIST = 2 };
int final_kind(int l, int L, const int32_t* n_r, const int32_t* n_f) {
    const bool has_x = l >= 2 && l <= L - 1 && n_f[l] > n_r[l];
    if (has_x) return V_XLIST;
    if (l >= 2 && l <= L - 2) return V_RESTRICT;
    return V_INIT;
}
int chain_level(int l, int L, const int32_t* n0, const int32_t* nr, const int32_t* nf) {
    static constexpr int CROSSOVER = 100;  // tree.py:12 tree_search_crossover
    int kind = final_kind(l, L, nr, nf);
    for (;;) {
        const int size = kind == V_INIT ? n0[l] : kind == V_RESTRICT ? nr[l] : nf[l];
        if (l == 0 || size < CROSSOVER) return l;
        if (kind == V_RESTRICT) kind = l - 1 >= 2 ? V_RESTRICT : V_INIT;
        else if (kind == V_XLIST) kind = final_kind(l - 1, L, nr, nf);
        --l;
    }
}

template <class T>
T read_scalar(const T* d, hipStream_t s) {
    T v{};
    KF_HIP_CHECK(hipMemcpyAsync(&v, d, sizeof(T), hipMemcpyDeviceToHost, s));
    KF_HIP_CHECK(hipStreamSynchronize(s));
    return v;
}

}  // namespace

void build_dev_forest(DevForest& F, DevBuf<uint64_t>& keys, const std::vector<int64_t>& pt_off,
                      const std::vector<int32_t>& D, int cutoff, hipStream_t s) {
    const int nt = (int)D.size();
    const int64_t N = pt_off[nt];
    F.nt = nt;
    F.cutoff = cutoff;
    F.pt_off = pt_off;
    F.D = D;
    F.keys.p = keys.p;
    F.keys.n = keys.n;
    keys.p = nullptr;
    keys.n = 0;
    int tbits = 0;
    while ((1 << tbits) < nt) ++tbits;
    DevBuf<int64_t> d_poff;
    DevBuf<int32_t> d_D;
    d_poff.upload(pt_off, s);
    d_D.upload(D, s);
    // ---- 1. split by count
    DevBuf<uint32_t> cnt, scratch;
    cnt.alloc((size_t)N);
    scratch.alloc((size_t)scan_scratch(N));
    DevBuf<int32_t> err;
    err.alloc(1);
    err.zero(s);
    k_witness<<<nblocks(N), TB_THREADS, 0, s>>>(F.keys.p, d_poff.p, d_D.p, nt, N, cutoff, cnt.p, err.p);
    KF_HIP_CHECK(hipGetLastError());
    scan_u32(cnt.p, cnt.p, N, scratch.p, s);
    const uint32_t K = read_scalar(scratch.p + scan_total_index(N), s);
    KF_REQUIRE(read_scalar(err.p, s) == 0, KF_ERR_DEPTH,
               "tree needs more levels than the key resolution allows (coincident points or extreme clustering)");
    DevBuf<uint64_t> skey, kb;
    DevBuf<int32_t> sval, vb;
    DevBuf<uint32_t> cnt_tl, hist;
    cnt_tl.alloc((size_t)nt * KMAXL);
    cnt_tl.zero(s);
    skey.alloc(std::max<uint32_t>(K, 1));
    sval.alloc(std::max<uint32_t>(K, 1));
    if (K > 0) {
        k_emit_splits<<<nblocks(N), TB_THREADS, 0, s>>>(F.keys.p, d_poff.p, d_D.p, nt, N, cutoff, cnt.p, skey.p,
                                                        sval.p, cnt_tl.p);
        KF_HIP_CHECK(hipGetLastError());
        kb.alloc(K);
        vb.alloc(K);
        hist.alloc((size_t)sort_hist_words(K));
        KF_HIP_CHECK(sort_pairs(skey.p, sval.p, K, tbits + 5, kb.p, vb.p, hist.p, s));
        KF_HIP_CHECK(hipGetLastError());
    }
    std::vector<uint32_t> h_cnt((size_t)nt * KMAXL);
    KF_HIP_CHECK(hipMemcpyAsync(h_cnt.data(), cnt_tl.p, h_cnt.size() * 4, hipMemcpyDeviceToHost, s));
    KF_HIP_CHECK(hipStreamSynchronize(s));
    // phase-1 tables
    std::vector<int32_t> L(nt), base1((size_t)nt * KMAXL, 0), offSL((size_t)nt * KMAXL + 1, 0);
    std::vector<int32_t> n0((size_t)nt * KMAXL, 0);
    std::vector<int64_t> p1off(nt + 1, 0);
    int Lmax = 1;
    {
        int64_t acc = 0;
        for (int t = 0; t < nt; ++t)
            for (int l = 0; l < KMAXL; ++l) {
                offSL[(size_t)t * KMAXL + l] = (int32_t)acc;
                acc += h_cnt[(size_t)t * KMAXL + l];
            }
        offSL[(size_t)nt * KMAXL] = (int32_t)acc;
    }
    for (int t = 0; t < nt; ++t) {
        int top = -1;
        for (int l = 0; l < KMAXL; ++l)
            if (h_cnt[(size_t)t * KMAXL + l]) top = l;
        L[t] = top + 2;
        int32_t b = 0;
        for (int l = 0; l < L[t]; ++l) {
            base1[(size_t)t * KMAXL + l] = b;
            n0[(size_t)t * KMAXL + l] = l == 0 ? 1 : 4 * (int32_t)h_cnt[(size_t)t * KMAXL + l - 1];
            b += n0[(size_t)t * KMAXL + l];
        }
        for (int l = L[t]; l < KMAXL; ++l) base1[(size_t)t * KMAXL + l] = b;
        p1off[t + 1] = p1off[t] + b;
        Lmax = std::max(Lmax, L[t]);
    }
    const int64_t B1 = p1off[nt];
    // pool: room for the restriction's and the X-list's boxes (points on curves need up to about
    // 5.5 times the boxes of the split-by-count phase); a build that needs more retries with a
    // pool twice as large.  KF_TREE_POOL sets the first pool's factor (tests force retries with 1).
    const int pool_factor = [] {
        const char* e = std::getenv("KF_TREE_POOL");
        const int f = e ? std::atoi(e) : 0;
        return f >= 1 && f <= 64 ? f : 6;
    }();
    int64_t cap = pool_factor * B1 + 4096;
    DevBuf<int32_t> d_L, d_base1, d_offSL, d_nlev, d_cnt_box, d_nbox;
    DevBuf<int64_t> d_p1off;
    d_L.upload(L, s);
    d_base1.upload(base1, s);
    d_offSL.upload(offSL, s);
    d_p1off.upload(p1off, s);
    for (;;) {
        DevBuf<int32_t> tree, lvl, pos, lid, ix, iy, bot, top, parent, child0, depth, nb;
        DevBuf<uint8_t> leaf, fake, xl;
        tree.alloc(cap); lvl.alloc(cap); pos.alloc(cap); lid.alloc(cap); ix.alloc(cap); iy.alloc(cap);
        bot.alloc(cap); top.alloc(cap); parent.alloc(cap); child0.alloc(cap); depth.alloc(cap);
        nb.alloc((size_t)cap * 9);
        leaf.alloc(cap); fake.alloc(cap); xl.alloc(cap);
        Pool P{tree.p, lvl.p, pos.p, lid.p, ix.p, iy.p, bot.p, top.p, parent.p, child0.p, depth.p, nb.p,
               leaf.p, fake.p, xl.p};
        err.zero(s);
        k_boxes1<<<nblocks(B1), TB_THREADS, 0, s>>>(P, B1, F.keys.p, d_poff.p, d_D.p, d_L.p, nt, d_p1off.p, d_base1.p,
                                                     d_offSL.p, sval.p, cutoff, err.p);
        KF_HIP_CHECK(hipGetLastError());
        KF_HIP_CHECK(hipGetLastError());
        KF_REQUIRE(read_scalar(err.p, s) == 0, KF_ERR_HIP, "device tree build: inconsistent split lists");
        k_depth_split1<<<nblocks(B1), TB_THREADS, 0, s>>>(P, B1);
        for (int l = 0; l < Lmax; ++l) k_nb_level<<<nblocks(B1), TB_THREADS, 0, s>>>(P, B1, l);
        KF_HIP_CHECK(hipGetLastError());
        // counters: local ids and level sizes per tree, slots in use
        std::vector<int32_t> cnt_box(nt), nlev((size_t)nt * KMAXL, 0);
        for (int t = 0; t < nt; ++t) cnt_box[t] = (int32_t)(p1off[t + 1] - p1off[t]);
        nlev = n0;
        d_cnt_box.upload(cnt_box, s);
        d_nlev.upload(nlev, s);
        d_nbox.upload(std::vector<int32_t>{(int32_t)B1}, s);
        // ---- 2. level restriction sweeps
        DevBuf<uint32_t> flag, fidx, fscr;
        flag.alloc(cap);
        fidx.alloc(cap);
        fscr.alloc((size_t)scan_scratch(cap));
        DevBuf<int32_t> list, lvb, merged, newA, newB, inc, split, nsplit, inc2;
        DevBuf<uint64_t> lkey, lkb;
        DevBuf<uint32_t> lhist;
        list.alloc(cap); lvb.alloc(cap); merged.alloc(cap); newA.alloc(cap); newB.alloc(cap); inc.alloc(cap);
        split.alloc(cap); nsplit.alloc(1); inc2.alloc(cap);
        nsplit.zero(s);
        lkey.alloc(cap); lkb.alloc(cap);
        lhist.alloc((size_t)sort_hist_words(cap));
        bool overflow = false;
        std::vector<int32_t> nR;
        if (Lmax >= 4) {
            for (int sweep = 0;; ++sweep) {
                KF_REQUIRE(sweep < 4 * KMAXL * KMAXL, KF_ERR_HIP, "device tree build: level restriction did not settle");
                k_flag_bads<<<nblocks(cap), TB_THREADS, 0, s>>>(P, d_nbox.p, cap, d_L.p, flag.p);
                scan_u32(flag.p, fidx.p, cap, fscr.p, s);
                k_compact_boxes<<<nblocks(cap), TB_THREADS, 0, s>>>(flag.p, fidx.p, cap, P, 1, tbits, list.p, lkey.p);
                KF_HIP_CHECK(hipGetLastError());
                const uint32_t m = read_scalar(fscr.p + scan_total_index(cap), s);
                if (m == 0) break;
                KF_HIP_CHECK(sort_pairs(lkey.p, list.p, m, tbits + 5, lkb.p, lvb.p, lhist.p, s));
                k_cascade<<<1, CT, 0, s>>>(P, d_L.p, Lmax, tbits, list.p, lkey.p, (int)m, merged.p, newA.p, newB.p,
                                           inc.p, split.p, nsplit.p, d_nbox.p, d_cnt_box.p, d_nlev.p, cap, err.p);
                KF_HIP_CHECK(hipGetLastError());
                k_depth_list<<<256, TB_THREADS, 0, s>>>(P, split.p, nsplit.p, cap);
                KF_HIP_CHECK(hipGetLastError());
                const int32_t e = read_scalar(err.p, s);
                KF_REQUIRE((e & ~2) == 0, KF_ERR_HIP, "device tree build: inconsistent box structure");
                if (e & 2) {
                    overflow = true;
                    break;
                }
            }
        }
        if (!overflow) {
            nR.resize((size_t)nt * KMAXL);
            KF_HIP_CHECK(hipMemcpyAsync(nR.data(), d_nlev.p, nR.size() * 4, hipMemcpyDeviceToHost, s));
            const int32_t nbR = read_scalar(d_nbox.p, s);
            // ---- 3. X-list fake children
            k_flag_xlist<<<nblocks(cap), TB_THREADS, 0, s>>>(P, d_nbox.p, cap, d_L.p, flag.p);
            scan_u32(flag.p, fidx.p, cap, fscr.p, s);
            k_compact_boxes<<<nblocks(cap), TB_THREADS, 0, s>>>(flag.p, fidx.p, cap, P, 0, tbits, list.p, lkey.p);
            KF_HIP_CHECK(hipGetLastError());
            const uint32_t mx = read_scalar(fscr.p + scan_total_index(cap), s);
            if ((int64_t)nbR + 4 * (int64_t)mx > cap) {
                overflow = true;
            } else if (mx > 0) {
                KF_HIP_CHECK(sort_pairs(lkey.p, list.p, mx, tbits + 5, lkb.p, lvb.p, lhist.p, s));
                k_make_fakes<<<nblocks(mx), TB_THREADS, 0, s>>>(P, list.p, lkey.p, (int)mx, nbR, d_cnt_box.p, d_nlev.p,
                                                                 inc.p, inc2.p);
                k_fake_counts<<<nblocks(mx), TB_THREADS, 0, s>>>(P, list.p, lkey.p, (int)mx, inc.p, inc2.p,
                                                                  d_cnt_box.p, d_nlev.p, d_nbox.p);
                const int32_t f1 = nbR + 4 * (int32_t)mx;
                k_nb_range<<<nblocks(f1 - nbR), TB_THREADS, 0, s>>>(P, nbR, f1);
                k_nb_announce<<<nblocks(f1 - nbR), TB_THREADS, 0, s>>>(P, nbR, f1);
                KF_HIP_CHECK(hipGetLastError());
            }
        }
        if (overflow) {
            KF_REQUIRE(cap < 64 * B1 + (1 << 24), KF_ERR_HIP, "device tree build: box pool overflow");
            cap *= 2;
            continue;
        }
        const int32_t nbox = read_scalar(d_nbox.p, s);
        if (nbox > B1) k_ranges<<<nblocks(nbox - B1), TB_THREADS, 0, s>>>(P, (int32_t)B1, nbox, F.keys.p, d_poff.p, d_D.p);
        KF_HIP_CHECK(hipGetLastError());
        std::vector<int32_t> nF((size_t)nt * KMAXL), cb(nt);
        KF_HIP_CHECK(hipMemcpyAsync(nF.data(), d_nlev.p, nF.size() * 4, hipMemcpyDeviceToHost, s));
        KF_HIP_CHECK(hipMemcpyAsync(cb.data(), d_cnt_box.p, cb.size() * 4, hipMemcpyDeviceToHost, s));
        KF_HIP_CHECK(hipStreamSynchronize(s));
        // ---- 4. final layout
        F.L = L;
        F.nlev = nF;
        F.box_off.assign(nt + 1, 0);
        F.lev_off.assign((size_t)nt * KMAXL + 1, 0);
        std::vector<int32_t> mlev((size_t)nt * KMAXL, 0);
        int64_t lacc = 0;
        for (int t = 0; t < nt; ++t) {
            F.box_off[t + 1] = F.box_off[t] + cb[t];
            const int32_t* r0 = &n0[(size_t)t * KMAXL];
            const int32_t* rr = &nR[(size_t)t * KMAXL];
            const int32_t* rf = &nF[(size_t)t * KMAXL];
            for (int l = 0; l < KMAXL; ++l) {
                F.lev_off[(size_t)t * KMAXL + l] = lacc;
                if (l < L[t]) {
                    lacc += rf[l];
                    mlev[(size_t)t * KMAXL + l] = chain_level(l, L[t], r0, rr, rf);
                }
            }
        }
        F.lev_off[(size_t)nt * KMAXL] = lacc;
        const int64_t B = F.box_off[nt];
        KF_REQUIRE(B == nbox && lacc == nbox, KF_ERR_HIP, "device tree build: inconsistent box counts");
        DevBuf<int32_t> d_mlev, anchor;
        DevBuf<int64_t> d_box_off, d_lev_off;
        d_mlev.upload(mlev, s);
        d_box_off.upload(F.box_off, s);
        d_lev_off.upload(F.lev_off, s);
        anchor.alloc(nbox);
        F.lvl.alloc(B); F.ix.alloc(B); F.iy.alloc(B); F.bot.alloc(B); F.top.alloc(B); F.parent.alloc(B);
        F.child0.alloc(B); F.depth.alloc(B); F.pos.alloc(B); F.coll.alloc((size_t)B * 9);
        F.leaf.alloc(B); F.fake.alloc(B); F.xl.alloc(B); F.level_list.alloc(B);
        k_anchor<<<nblocks(nbox), TB_THREADS, 0, s>>>(P, nbox, d_mlev.p, anchor.p);
        Final Fv{F.lvl.p, F.ix.p, F.iy.p, F.bot.p, F.top.p, F.parent.p, F.child0.p, F.depth.p, F.pos.p, F.coll.p,
                 F.level_list.p, F.leaf.p, F.fake.p, F.xl.p};
        k_final<<<nblocks(nbox), TB_THREADS, 0, s>>>(P, nbox, d_box_off.p, d_lev_off.p, d_mlev.p, anchor.p, Fv,
                                                      err.p);
        KF_HIP_CHECK(hipGetLastError());
        KF_REQUIRE(read_scalar(err.p, s) == 0, KF_ERR_HIP, "device tree build: inconsistent box structure");
        return;  // read_scalar synchronised: the pool and the tables above are freed on return
    }
}

void dev_forest_to_host(const DevForest& F, int t, HostTree& h, hipStream_t s) {
    const int64_t b0 = F.box_off[t], B = F.box_off[t + 1] - b0;
    std::vector<int32_t> lvl(B), ix(B), iy(B), bot(B), top(B), parent(B), child0(B), depth(B), pos(B), coll(B * 9);
    std::vector<uint8_t> leaf(B), fake(B), xl(B);
    auto get = [&](auto& v, const auto& d, int64_t per) {
        KF_HIP_CHECK(hipMemcpyAsync(v.data(), d.p + b0 * per, v.size() * sizeof(v[0]), hipMemcpyDeviceToHost, s));
    };
    get(lvl, F.lvl, 1); get(ix, F.ix, 1); get(iy, F.iy, 1); get(bot, F.bot, 1); get(top, F.top, 1);
    get(parent, F.parent, 1); get(child0, F.child0, 1); get(depth, F.depth, 1); get(pos, F.pos, 1);
    get(coll, F.coll, 9); get(leaf, F.leaf, 1); get(fake, F.fake, 1); get(xl, F.xl, 1);
    KF_HIP_CHECK(hipStreamSynchronize(s));
    h.lvl = lvl;
    h.ix.assign(ix.begin(), ix.end());
    h.iy.assign(iy.begin(), iy.end());
    h.bot.assign(bot.begin(), bot.end());
    h.top.assign(top.begin(), top.end());
    h.parent = parent;
    h.child0 = child0;
    h.depth = depth;
    h.pos = pos;
    h.leaf = leaf;
    h.fake = fake;
    h.xl = xl;
    const int L = F.L[t];
    h.levels.assign(L, {});
    h.coll.assign(L, {});
    h.n_real.assign(L, 0);
    for (int l = 0; l < L; ++l) h.levels[l].assign(F.nlev[(size_t)t * KMAXL + l], -1);
    for (int64_t b = 0; b < B; ++b) h.levels[lvl[b]][pos[b]] = (int32_t)b;
    for (int l = 0; l < L; ++l) {
        const auto& lv = h.levels[l];
        auto& C = h.coll[l];
        C.resize(lv.size() * 9);
        int32_t nr = 0;
        for (size_t i = 0; i < lv.size(); ++i) {
            const int32_t b = lv[i];
            for (int k = 0; k < 9; ++k) C[i * 9 + k] = coll[(size_t)b * 9 + k];
            nr += fake[b] ? 0 : 1;
        }
        h.n_real[l] = nr;
    }
}

}  // namespace kf
