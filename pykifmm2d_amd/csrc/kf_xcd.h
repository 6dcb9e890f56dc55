// kf_xcd.h -- workgroup-to-work mapping for the 8 XCDs of an MI355X.
//
// The dispatcher hands workgroups to the XCDs round-robin (workgroup b runs on XCD b % 8), and
// each XCD has its own L2.  Work lists here are in Morton order, so neighbouring items share
// their sources; xcd_contiguous(b, n) maps the n workgroups of a launch so that XCD x works on
// the x-th contiguous range of items, which keeps the shared sources in one L2 instead of
// fetching them into all eight.  It is a bijection of [0, n) (tests/plan_index_check.cpp).
// Used by the M2L, whose tiles of neighbouring targets share source rows and operators: it
// halves the M2L's HBM fetch at the same time.  The VALU passes (P2M, P2P, L2P) keep the
// round-robin order: contiguous ranges of leaves carry unequal work on clustered points, and
// the slowest XCD then sets their time (measured 681 -> 875 us for P2P at 1M clustered points).
// Host-only translation units (the index checker) include it too, hence the qualifier macro.
#pragma once

#if defined(__HIPCC__)
#define KF_XCD_FN __host__ __device__ inline
#else
#define KF_XCD_FN inline
#endif

namespace kf {

KF_XCD_FN int xcd_contiguous(int b, int n) {
    const int x = b & 7, k = b >> 3;
    int start = 0;
    for (int y = 0; y < x; ++y) start += (n - y + 7) >> 3;  // items of XCDs before x
    return start + k;
}

// Runs of K neighbouring items per XCD, dealt round-robin: the items form groups of 8K, and
// XCD x takes the x-th run of K items of each group, so neighbours that share sources meet in
// one L2 while the XCDs still draw their work from the whole list (the balance that
// xcd_contiguous gives up).  Items past the last whole group keep the identity map.  A bijection
// of [0, n) (tests/plan_index_check.cpp).
KF_XCD_FN int xcd_chunked(int b, int n, int K) {
    const int G = 8 * K, full = n - n % G;
    if (b >= full) return b;
    const int g = b / G, r = b - g * G;
    return g * G + (r & 7) * K + (r >> 3);
}

}  // namespace kf
