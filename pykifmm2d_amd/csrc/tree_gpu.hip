// tree_gpu.hip -- point-level quadtree work on the GPU: bounding box, exact Morton keys,
// a stable LSD radix sort of (key, index) pairs (kf_sort.h) and the gather of points into tree
// order.  The box-level structure is then built on the GPU from the sorted keys (tree_dev.hip).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>

#include "kf_internal.h"
#include "kf_sort.h"
#include "tree_dev.h"

namespace kf {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

// ------------------------------------------------------------------ bounding box ----
static constexpr int BB_BLOCKS = 512, BB_THREADS = 256;

__global__ void __launch_bounds__(BB_THREADS) k_bbox(const double* __restrict__ x,
                                                     const double* __restrict__ y, int64_t n,
                                                     double* __restrict__ part) {
    double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY, bad = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double a = x[i], b = y[i];
        if (!isfinite(a) || !isfinite(b)) bad = 1.0;
        mnx = fmin(mnx, a);
        mxx = fmax(mxx, a);
        mny = fmin(mny, b);
        mxy = fmax(mxy, b);
    }
    __shared__ double s[5][BB_THREADS];
    s[0][threadIdx.x] = mnx;
    s[1][threadIdx.x] = mxx;
    s[2][threadIdx.x] = mny;
    s[3][threadIdx.x] = mxy;
    s[4][threadIdx.x] = bad;
    __syncthreads();
    for (int w = BB_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            int j = threadIdx.x + w;
            s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][j]);
            s[1][threadIdx.x] = fmax(s[1][threadIdx.x], s[1][j]);
            s[2][threadIdx.x] = fmin(s[2][threadIdx.x], s[2][j]);
            s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][j]);
            s[4][threadIdx.x] = fmax(s[4][threadIdx.x], s[4][j]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) part[blockIdx.x * 5 + threadIdx.x] = s[threadIdx.x][0];
}

// ------------------------------------------------------------------ Morton keys -----
// ix = number of level-D gridlines root_min + k*h (k = 1 .. 2^D-1) strictly below x.
// Every gridline is an exact FP64 value (the key depth is chosen so), so the strict ">"
// comparisons reproduce the reference's midline classification (tree.py:33-35) at every
// level: the level-l index is ix >> (D - l).
__device__ __forceinline__ uint64_t grid_index(double v, double lo, double h, int D) {
    const int64_t top = ((int64_t)1 << D) - 1;
    double t = (v - lo) / h;
    int64_t m = (int64_t)ceil(t) - 1;
    m = m < 0 ? 0 : (m > top ? top : m);
    while (m > 0 && !(v > fma((double)m, h, lo))) --m;
    while (m < top && v > fma((double)(m + 1), h, lo)) ++m;
    return (uint64_t)m;
}

__device__ __forceinline__ uint64_t spread_bits(uint64_t v) {
    v &= 0xffffffffull;
    v = (v | (v << 16)) & 0x0000ffff0000ffffull;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}

__global__ void k_keys(const double* __restrict__ x, const double* __restrict__ y, int64_t n,
                       double lo, double h, int D, uint64_t* __restrict__ keys,
                       int32_t* __restrict__ idx) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t a = grid_index(x[i], lo, h, D), b = grid_index(y[i], lo, h, D);
    keys[i] = (spread_bits(a) << 1) | spread_bits(b);
    idx[i] = (int32_t)i;
}

__global__ void k_gather_points(const double* __restrict__ x, const double* __restrict__ y,
                                const int32_t* __restrict__ perm, int64_t n,
                                double* __restrict__ xs, double* __restrict__ ys) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t j = perm[i];
    xs[i] = x[j];
    ys[i] = y[j];
}

// Sorts (keys, vals) in place by the low `bits` key bits (stable); stream-ordered.
static void radix_sort(DevBuf<uint64_t>& keys, DevBuf<int32_t>& vals, int64_t n, int bits, hipStream_t s) {
    if (n <= 1) return;
    DevBuf<uint64_t> k2;
    DevBuf<int32_t> v2;
    DevBuf<uint32_t> hist;
    k2.alloc(n);
    v2.alloc(n);
    hist.alloc((size_t)sort_hist_words(n));
    KF_HIP_CHECK(sort_pairs(keys.p, vals.p, n, bits, k2.p, v2.p, hist.p, s));
    KF_HIP_CHECK(hipGetLastError());
    KF_HIP_CHECK(hipStreamSynchronize(s));  // the scratch buffers are freed on return
}

// Root box and key depth of one tree from its bounding box (reference tree.py:475-484, with the
// y.min() upper bound of tree.py:478; points above it are rejected).
static void root_box(double xmin, double xmax, double ymin, double ymax, double bad, double& mmin, double& W,
                     int& D, const std::string& who) {
    KF_REQUIRE(bad == 0.0, KF_ERR_ARG, who + "point coordinates must be finite");
    mmin = std::floor(std::min(xmin, ymin));
    const double mmax = std::ceil(std::max(xmax, ymin));
    KF_REQUIRE(ymax <= mmax, KF_ERR_DOMAIN,
               who + "y range exceeds the root box [floor(min(x.min,y.min)), ceil(max(x.max,y.min))] "
                     "that the reference tree rule produces");
    KF_REQUIRE(mmax > mmin, KF_ERR_DOMAIN, who + "degenerate root box (all points on one integer)");
    W = mmax - mmin;
    // key depth: every gridline mmin + k W 2^-D must be an exact double
    const double B = std::fabs(mmin) + W + 1.0;
    D = std::min(30, 52 - (int)std::ceil(std::log2(B)));
    KF_REQUIRE(D >= 4, KF_ERR_DOMAIN, who + "coordinates too large for exact midline keys");
}

static void build_tree(kf_tree* t, const double* d_x, const double* d_y, int64_t n, int cutoff,
                       hipStream_t s) {
    KF_REQUIRE(n >= 1, KF_ERR_ARG, "need at least one point");
    KF_REQUIRE(n < ((int64_t)1 << 31), KF_ERR_ARG, "at most 2^31-1 points per tree");
    KF_REQUIRE(cutoff >= 1, KF_ERR_ARG, "cutoff must be >= 1");
    KF_REQUIRE(d_x && d_y, KF_ERR_ARG, "null point pointer");
    // bounding box (reference tree.py:475-480)
    DevBuf<double> part;
    part.alloc(BB_BLOCKS * 5);
    k_bbox<<<BB_BLOCKS, BB_THREADS, 0, s>>>(d_x, d_y, n, part.p);
    KF_HIP_CHECK(hipGetLastError());
    std::vector<double> hp(BB_BLOCKS * 5);
    KF_HIP_CHECK(hipMemcpyAsync(hp.data(), part.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    KF_HIP_CHECK(hipStreamSynchronize(s));
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY, bad = 0;
    for (int b = 0; b < BB_BLOCKS; ++b) {
        xmin = std::min(xmin, hp[b * 5 + 0]);
        xmax = std::max(xmax, hp[b * 5 + 1]);
        ymin = std::min(ymin, hp[b * 5 + 2]);
        ymax = std::max(ymax, hp[b * 5 + 3]);
        bad = std::max(bad, hp[b * 5 + 4]);
    }
    double mmin, W;
    int D;
    root_box(xmin, xmax, ymin, ymax, bad, mmin, W, D, "");
    HostTree& h = t->h;
    h.n = n;
    h.cutoff = cutoff;
    h.root_min = mmin;
    h.root_width = W;
    h.key_depth = D;
    t->ymax = ymax;
    DevBuf<uint64_t> keys;
    DevBuf<int32_t> idx;
    keys.alloc(n);
    idx.alloc(n);
    int thr = 256;
    int64_t nb = (n + thr - 1) / thr;
    k_keys<<<(unsigned)nb, thr, 0, s>>>(d_x, d_y, n, mmin, std::ldexp(W, -D), D, keys.p, idx.p);
    KF_HIP_CHECK(hipGetLastError());
    radix_sort(keys, idx, n, 2 * D, s);
    t->pts = std::make_shared<TreePoints>();
    TreePoints& P = *t->pts;
    P.xs.alloc(n);
    P.ys.alloc(n);
    k_gather_points<<<(unsigned)nb, thr, 0, s>>>(d_x, d_y, idx.p, n, P.xs.p, P.ys.p);
    KF_HIP_CHECK(hipGetLastError());
    P.perm.p = idx.p;
    P.perm.n = idx.n;
    idx.p = nullptr;
    idx.n = 0;
    t->xs = P.xs.p;
    t->ys = P.ys.p;
    t->perm = P.perm.p;
    t->dev = std::make_shared<DevForest>();
    build_dev_forest(*t->dev, keys, {0, n}, {D}, cutoff, s);
}

// ------------------------------------------------------------------ batch builds ----
// Many trees at once: the points of n_trees trees are concatenated (tree t owns input
// [off[t], off[t+1])).  Bounding boxes, keys, sort and gather are shared launches; the sort
// orders by Morton key, then (stably) by tree, which leaves each tree's points contiguous and
// in its own key order.

// Tree of input point i: off[t] <= i < off[t+1].
__device__ __forceinline__ int tree_of(const int64_t* __restrict__ off, int nt, int64_t i) {
    int lo = 0, hi = nt - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One workgroup per tree; part[t*5 ..] = (xmin, xmax, ymin, ymax, non-finite flag).
__global__ void __launch_bounds__(BB_THREADS) k_bbox_seg(const double* __restrict__ x, const double* __restrict__ y,
                                                         const int64_t* __restrict__ off,
                                                         double* __restrict__ part) {
    const int t = blockIdx.x;
    double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY, bad = 0.0;
    for (int64_t i = off[t] + threadIdx.x; i < off[t + 1]; i += blockDim.x) {
        double a = x[i], b = y[i];
        if (!isfinite(a) || !isfinite(b)) bad = 1.0;
        mnx = fmin(mnx, a);
        mxx = fmax(mxx, a);
        mny = fmin(mny, b);
        mxy = fmax(mxy, b);
    }
    __shared__ double sh[5][BB_THREADS];
    sh[0][threadIdx.x] = mnx;
    sh[1][threadIdx.x] = mxx;
    sh[2][threadIdx.x] = mny;
    sh[3][threadIdx.x] = mxy;
    sh[4][threadIdx.x] = bad;
    __syncthreads();
    for (int w = BB_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            int j = threadIdx.x + w;
            sh[0][threadIdx.x] = fmin(sh[0][threadIdx.x], sh[0][j]);
            sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][j]);
            sh[2][threadIdx.x] = fmin(sh[2][threadIdx.x], sh[2][j]);
            sh[3][threadIdx.x] = fmax(sh[3][threadIdx.x], sh[3][j]);
            sh[4][threadIdx.x] = fmax(sh[4][threadIdx.x], sh[4][j]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) part[(size_t)t * 5 + threadIdx.x] = sh[threadIdx.x][0];
}

__device__ __forceinline__ uint64_t morton_of(double x, double y, double lo, double h, int D) {
    return (spread_bits(grid_index(x, lo, h, D)) << 1) | spread_bits(grid_index(y, lo, h, D));
}

__global__ void k_keys_seg(const double* __restrict__ x, const double* __restrict__ y, int64_t n,
                           const int64_t* __restrict__ off, int nt, const double* __restrict__ glo,
                           const double* __restrict__ gh, const int32_t* __restrict__ gD,
                           uint64_t* __restrict__ keys, int32_t* __restrict__ idx) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = tree_of(off, nt, i);
    keys[i] = morton_of(x[i], y[i], glo[t], gh[t], gD[t]);
    idx[i] = (int32_t)i;
}

// Second sort key: the tree of each element in the current order.
__global__ void k_tree_keys(const int32_t* __restrict__ idx, int64_t n, const int64_t* __restrict__ off, int nt,
                            uint64_t* __restrict__ keys) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = (uint64_t)tree_of(off, nt, idx[i]);
}

// Points, permutations (into the concatenated input and into the own tree's input) and Morton
// keys in (tree, key) order.
__global__ void k_gather_seg(const double* __restrict__ x, const double* __restrict__ y,
                             const int32_t* __restrict__ idx, int64_t n, const int64_t* __restrict__ off, int nt,
                             const double* __restrict__ glo, const double* __restrict__ gh,
                             const int32_t* __restrict__ gD, double* __restrict__ xs, double* __restrict__ ys,
                             int32_t* __restrict__ gperm, int32_t* __restrict__ lperm,
                             uint64_t* __restrict__ keys) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t j = idx[i];
    const int t = tree_of(off, nt, j);
    const double a = x[j], b = y[j];
    xs[i] = a;
    ys[i] = b;
    gperm[i] = j;
    lperm[i] = (int32_t)(j - off[t]);
    keys[i] = morton_of(a, b, glo[t], gh[t], gD[t]);
}

static void build_tree_batch(const double* d_x, const double* d_y, const int64_t* h_off, int nt, int cutoff,
                             hipStream_t s, std::vector<kf_tree*>& out) {
    KF_REQUIRE(d_x && d_y && h_off, KF_ERR_ARG, "null argument");
    KF_REQUIRE(nt >= 1, KF_ERR_ARG, "need at least one tree");
    KF_REQUIRE(cutoff >= 1, KF_ERR_ARG, "cutoff must be >= 1");
    KF_REQUIRE(h_off[0] == 0, KF_ERR_ARG, "offsets must start at 0");
    for (int t = 0; t < nt; ++t) KF_REQUIRE(h_off[t + 1] > h_off[t], KF_ERR_ARG, "every tree needs at least one point");
    const int64_t n = h_off[nt];
    KF_REQUIRE(n < ((int64_t)1 << 31), KF_ERR_ARG, "at most 2^31-1 points per batch");
    std::vector<int64_t> off(h_off, h_off + nt + 1);
    DevBuf<int64_t> doff;
    doff.upload(off, s);
    DevBuf<double> part;
    part.alloc((size_t)nt * 5);
    k_bbox_seg<<<nt, BB_THREADS, 0, s>>>(d_x, d_y, doff.p, part.p);
    KF_HIP_CHECK(hipGetLastError());
    std::vector<double> hp((size_t)nt * 5);
    KF_HIP_CHECK(hipMemcpyAsync(hp.data(), part.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    KF_HIP_CHECK(hipStreamSynchronize(s));
    std::vector<double> glo(nt), gh(nt), W(nt);
    std::vector<int32_t> gD(nt);
    int maxD = 1;
    for (int t = 0; t < nt; ++t) {
        const double* q = &hp[(size_t)t * 5];
        double mmin, w;
        int D;
        root_box(q[0], q[1], q[2], q[3], q[4], mmin, w, D, "tree " + std::to_string(t) + ": ");
        glo[t] = mmin;
        W[t] = w;
        gh[t] = std::ldexp(w, -D);
        gD[t] = D;
        maxD = std::max(maxD, D);
    }
    DevBuf<double> dlo, dh;
    DevBuf<int32_t> dD;
    dlo.upload(glo, s);
    dh.upload(gh, s);
    dD.upload(gD, s);
    DevBuf<uint64_t> keys;
    DevBuf<int32_t> idx;
    keys.alloc(n);
    idx.alloc(n);
    const int thr = 256;
    const unsigned nb = (unsigned)((n + thr - 1) / thr);
    k_keys_seg<<<nb, thr, 0, s>>>(d_x, d_y, n, doff.p, nt, dlo.p, dh.p, dD.p, keys.p, idx.p);
    KF_HIP_CHECK(hipGetLastError());
    radix_sort(keys, idx, n, 2 * maxD, s);
    if (nt > 1) {
        int bits = 1;
        while ((1 << bits) < nt) ++bits;
        k_tree_keys<<<nb, thr, 0, s>>>(idx.p, n, doff.p, nt, keys.p);
        KF_HIP_CHECK(hipGetLastError());
        radix_sort(keys, idx, n, bits, s);
    }
    auto pts = std::make_shared<TreePoints>();
    pts->n_trees = nt;
    pts->xs.alloc(n);
    pts->ys.alloc(n);
    pts->perm.alloc(n);
    pts->gperm.alloc(n);
    k_gather_seg<<<nb, thr, 0, s>>>(d_x, d_y, idx.p, n, doff.p, nt, dlo.p, dh.p, dD.p, pts->xs.p, pts->ys.p,
                                    pts->gperm.p, pts->perm.p, keys.p);
    KF_HIP_CHECK(hipGetLastError());
    // box structure of every tree on the GPU, in shared launches
    auto dev = std::make_shared<DevForest>();
    build_dev_forest(*dev, keys, off, gD, cutoff, s);
    std::vector<std::unique_ptr<kf_tree>> made(nt);
    for (int t = 0; t < nt; ++t) {
        made[t].reset(new kf_tree());
        kf_tree* tr = made[t].get();
        HostTree& h = tr->h;
        h.n = off[t + 1] - off[t];
        h.cutoff = cutoff;
        h.root_min = glo[t];
        h.root_width = W[t];
        h.key_depth = gD[t];
        tr->ymax = hp[(size_t)t * 5 + 3];
        tr->dev = dev;
        tr->pts = pts;
        tr->slot = t;
        tr->off = off[t];
        tr->xs = pts->xs.p + off[t];
        tr->ys = pts->ys.p + off[t];
        tr->perm = pts->perm.p + off[t];
    }
    out.clear();
    for (int t = 0; t < nt; ++t) out.push_back(made[t].release());
}

// Leaf of each point (reference tree.py:686-727 numba_locate_points): from the root, while the
// box is not a leaf, go to child 2 (x > xmid) + (y > ymid).  lvl / ix / iy / child0 / leaf / pos:
// the tree's device boxes (local ids).
__global__ void k_tree_locate(const double* __restrict__ x, const double* __restrict__ y, int64_t m,
                              const int32_t* __restrict__ lvl, const int32_t* __restrict__ ix,
                              const int32_t* __restrict__ iy, const int32_t* __restrict__ child0,
                              const uint8_t* __restrict__ leaf, const int32_t* __restrict__ pos, double root_min,
                              double root_w, int64_t* __restrict__ out_level, int64_t* __restrict__ out_pos) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double px = x[i], py = y[i];
    int32_t b = 0;
    for (int guard = 0; guard < 32 && !leaf[b]; ++guard) {
        const double w = ldexp(root_w, -lvl[b]);
        const double xmid = (root_min + (double)ix[b] * w) + 0.5 * w;  // Level.xmin + half_width
        const double ymid = (root_min + (double)iy[b] * w) + 0.5 * w;
        b = child0[b] + (px > xmid ? 2 : 0) + (py > ymid ? 1 : 0);
    }
    out_level[i] = lvl[b];
    out_pos[i] = pos[b];
}

}  // namespace kf

const kf::HostTree& kf_tree::host() const {
    if (dev) std::call_once(mirrored, [this] { kf::dev_forest_to_host(*dev, slot, h, nullptr); });
    return h;
}
int kf_tree::levels() const { return dev ? dev->L[slot] : h.nlevels(); }
int64_t kf_tree::n_boxes() const { return dev ? dev->box_off[slot + 1] - dev->box_off[slot] : (int64_t)h.lvl.size(); }
int64_t kf_tree::level_size(int l) const {
    return dev ? dev->nlev[(size_t)slot * 32 + l] : (int64_t)h.levels[l].size();
}

using namespace kf;

extern "C" {

int kf_version(void) { return KF_VERSION; }

const char* kf_last_error(void) { return kf::g_err.c_str(); }

double kf_k0(double x) { return ::k0_f64(x); }

int kf_tree_build(const double* d_x, const double* d_y, int64_t n, int32_t cutoff, void* stream,
                  kf_tree** out) {
    return guarded([&] {
        KF_REQUIRE(out, KF_ERR_ARG, "null output handle");
        *out = nullptr;
        kf_tree* t = new kf_tree();
        try {
            build_tree(t, d_x, d_y, n, cutoff, (hipStream_t)stream);
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
    });
}

int kf_tree_build_batch(const double* d_x, const double* d_y, const int64_t* h_offsets, int32_t n_trees,
                        int32_t cutoff, void* stream, kf_tree** out) {
    return guarded([&] {
        KF_REQUIRE(out, KF_ERR_ARG, "null output array");
        for (int t = 0; t < n_trees; ++t) out[t] = nullptr;
        std::vector<kf_tree*> made;
        build_tree_batch(d_x, d_y, h_offsets, n_trees, cutoff, (hipStream_t)stream, made);
        for (int t = 0; t < n_trees; ++t) out[t] = made[t];
    });
}

int kf_tree_get_info(const kf_tree* t, kf_tree_info* info) {
    return guarded([&] {
        KF_REQUIRE(t && info, KF_ERR_ARG, "null argument");
        info->levels = t->levels();
        info->key_depth = t->h.key_depth;
        info->n_points = t->h.n;
        info->n_boxes = t->n_boxes();
        info->root_min = t->h.root_min;
        info->root_width = t->h.root_width;
    });
}

int kf_tree_level_size(const kf_tree* t, int32_t level, int64_t* n_node) {
    return guarded([&] {
        KF_REQUIRE(t && n_node, KF_ERR_ARG, "null argument");
        KF_REQUIRE(level >= 0 && level < t->levels(), KF_ERR_ARG, "level out of range");
        *n_node = t->level_size(level);
    });
}

int kf_tree_export_level(const kf_tree* t, int32_t level, double* xmin, double* ymin,
                         int64_t* bot_ind, int64_t* top_ind, int64_t* parent_ind,
                         int64_t* children_ind, uint8_t* leaf, uint8_t* fake_leaf,
                         uint8_t* xlist, int64_t* colleagues) {
    return guarded([&] {
        KF_REQUIRE(t, KF_ERR_ARG, "null tree");
        const HostTree& h = t->host();
        KF_REQUIRE(level >= 0 && level < h.nlevels(), KF_ERR_ARG, "level out of range");
        const auto& L = h.levels[level];
        double w = h.width(level);
        for (size_t i = 0; i < L.size(); ++i) {
            int32_t b = L[i];
            if (xmin) xmin[i] = h.root_min + (double)h.ix[b] * w;
            if (ymin) ymin[i] = h.root_min + (double)h.iy[b] * w;
            if (bot_ind) bot_ind[i] = h.bot[b];
            if (top_ind) top_ind[i] = h.top[b];
            if (parent_ind) parent_ind[i] = h.parent[b] >= 0 ? h.pos[h.parent[b]] : -1;
            if (children_ind) children_ind[i] = h.child0[b] >= 0 ? h.pos[h.child0[b]] : -1;
            if (leaf) leaf[i] = h.leaf[b];
            if (fake_leaf) fake_leaf[i] = h.fake[b];
            if (xlist) xlist[i] = h.xl[b];
            if (colleagues)
                for (int j = 0; j < 9; ++j) {
                    int32_t c = h.coll[level][i * 9 + j];
                    colleagues[i * 9 + j] = c >= 0 ? h.pos[c] : -1;
                }
        }
    });
}

int kf_tree_export_points(const kf_tree* t, int64_t* ordv, double* x, double* y) {
    return guarded([&] {
        KF_REQUIRE(t, KF_ERR_ARG, "null tree");
        KF_REQUIRE(t->perm, KF_ERR_STATE, "tree has no device points");
        int64_t n = t->h.n;
        if (ordv) {
            std::vector<int32_t> p(n);
            KF_HIP_CHECK(hipMemcpy(p.data(), t->perm, n * sizeof(int32_t), hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) ordv[i] = p[i];
        }
        if (x) KF_HIP_CHECK(hipMemcpy(x, t->xs, n * sizeof(double), hipMemcpyDeviceToHost));
        if (y) KF_HIP_CHECK(hipMemcpy(y, t->ys, n * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int kf_tree_device_points(const kf_tree* t, const double** d_x, const double** d_y,
                          const int32_t** d_ordv) {
    return guarded([&] {
        KF_REQUIRE(t, KF_ERR_ARG, "null tree");
        if (d_x) *d_x = t->xs;
        if (d_y) *d_y = t->ys;
        if (d_ordv) *d_ordv = t->perm;
    });
}

int kf_tree_from_sorted_keys(const uint64_t* keys, int64_t n, int32_t cutoff, double root_min,
                             double root_width, int32_t key_depth, kf_tree** out) {
    return guarded([&] {
        KF_REQUIRE(out && (keys || n == 0), KF_ERR_ARG, "null argument");
        KF_REQUIRE(n >= 1 && cutoff >= 1 && key_depth >= 1 && key_depth <= 30 && root_width > 0,
                   KF_ERR_ARG, "invalid tree parameters");
        *out = nullptr;
        kf_tree* t = new kf_tree();
        HostTree& h = t->h;
        h.n = n;
        h.cutoff = cutoff;
        h.root_min = root_min;
        h.root_width = root_width;
        h.key_depth = key_depth;
        h.keys.assign(keys, keys + n);
        for (int64_t i = 1; i < n; ++i)
            if (h.keys[i] < h.keys[i - 1]) {
                delete t;
                throw KfError{KF_ERR_ARG, "keys are not sorted"};
            }
        std::string err = h.build();
        if (!err.empty()) {
            delete t;
            throw KfError{KF_ERR_DEPTH, err};
        }
        *out = t;
    });
}

int kf_tree_locate(const kf_tree* t, const double* d_x, const double* d_y, int64_t m, int64_t* d_level,
                   int64_t* d_pos, void* stream) {
    return guarded([&] {
        KF_REQUIRE(t && t->dev, KF_ERR_STATE, "locate needs a tree built on the GPU");
        KF_REQUIRE(m >= 0, KF_ERR_ARG, "negative count");
        if (m == 0) return;
        KF_REQUIRE(d_x && d_y && d_level && d_pos, KF_ERR_ARG, "null pointer");
        const DevForest& F = *t->dev;
        const int64_t b0 = F.box_off[t->slot];
        k_tree_locate<<<(unsigned)((m + 255) / 256), 256, 0, (hipStream_t)stream>>>(
            d_x, d_y, m, F.lvl.p + b0, F.ix.p + b0, F.iy.p + b0, F.child0.p + b0, F.leaf.p + b0, F.pos.p + b0,
            t->h.root_min, t->h.root_width, d_level, d_pos);
        KF_HIP_CHECK(hipGetLastError());
    });
}

int kf_tree_export_depths(const kf_tree* t, int32_t level, int64_t* depths) {
    return guarded([&] {
        KF_REQUIRE(t && depths, KF_ERR_ARG, "null argument");
        const HostTree& h = t->host();
        KF_REQUIRE(level >= 0 && level < h.nlevels(), KF_ERR_ARG, "level out of range");
        const auto& L = h.levels[level];
        for (int32_t i = 0; i < h.n_real[level]; ++i) depths[i] = h.depth[L[i]];
    });
}

int kf_tree_level_real(const kf_tree* t, int32_t level, int64_t* n_real) {
    return guarded([&] {
        KF_REQUIRE(t && n_real, KF_ERR_ARG, "null argument");
        const HostTree& h = t->host();
        KF_REQUIRE(level >= 0 && level < h.nlevels(), KF_ERR_ARG, "level out of range");
        *n_real = h.n_real[level];
    });
}

void kf_tree_destroy(kf_tree* t) { delete t; }

}  // extern "C"
