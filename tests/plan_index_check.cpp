// plan_index_check.cpp -- CPU replay of the planned evaluation's indexing (test tool).
//
// Builds the box tree (tree_host.cpp) and the work lists (plan_lists.cpp) exactly as the
// library does, then walks every pass of one evaluation the way its kernel indexes memory,
// for a batch width R and a target set, and checks that
//   * every list index and slab row is inside what the library allocates for that layout
//     (U and Psi: rows_cap x R rows, X: n_src x R rows);
//   * every slab row a pass reads was written earlier in the same evaluation;
//   * the results the caller reads (L2P at the source leaves; for the all-target set the
//     expansion of every real leaf) are complete.
// Input file: int64 n, int32 cutoff, double root_min, double root_width, int32 key_depth,
// then n sorted uint64 Morton keys.  Usage: plan_index_check <keys file> <lg R> <mode>.
// Prints "OK <counts>" or the first violations and exits non-zero.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "plan_lists.h"
#include "tree_host.h"

using namespace kf;

static int n_err = 0;
#define CHECK(cond, ...)                         \
    do {                                         \
        if (!(cond)) {                           \
            if (n_err < 20) {                    \
                std::printf("FAIL %s:%d: ", __func__, __LINE__); \
                std::printf(__VA_ARGS__);        \
                std::printf("\n");               \
            }                                    \
            ++n_err;                             \
        }                                        \
    } while (0)

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s keys.bin lgR mode\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    HostTree h;
    int64_t n;
    int32_t cutoff, depth;
    double rmin, rw;
    if (std::fread(&n, 8, 1, f) != 1 || std::fread(&cutoff, 4, 1, f) != 1 || std::fread(&rmin, 8, 1, f) != 1 ||
        std::fread(&rw, 8, 1, f) != 1 || std::fread(&depth, 4, 1, f) != 1)
        return 2;
    h.n = n;
    h.cutoff = cutoff;
    h.root_min = rmin;
    h.root_width = rw;
    h.key_depth = depth;
    h.keys.resize(n);
    if (std::fread(h.keys.data(), 8, n, f) != (size_t)n) return 2;
    std::fclose(f);
    const std::string err = h.build();
    if (!err.empty()) {
        std::printf("tree build failed: %s\n", err.c_str());
        return 2;
    }
    const int lg = std::atoi(argv[2]), mode = std::atoi(argv[3]);
    const int R = 1 << lg, bpt = 64 >> lg;
    PlanLists ls;
    build_plan_lists(h, 32, 32, ls);
    const int L = ls.L;
    const int64_t B = (int64_t)h.lvl.size();
    const int64_t rows_cap = mode == KF_TGT_ALL ? ls.n_all : ls.n_src;

    // rid: boxes with sources at levels >= 1 -> [0, n_src); real empty boxes at >= 2 -> [n_src, n_all)
    for (int64_t b = 0; b < B; ++b) {
        const int32_t r = ls.rid[b];
        if (h.lvl[b] >= 1 && h.ns((int32_t)b) > 0) CHECK(r >= 0 && r < ls.n_src, "box %lld row %d", (long long)b, r);
        else if (h.lvl[b] >= 2 && !h.fake[b]) CHECK(r >= ls.n_src && r < ls.n_all, "empty box %lld row %d", (long long)b, r);
        else CHECK(r == -1, "box %lld should have no row (%d)", (long long)b, r);
    }
    std::vector<char> Uw(rows_cap, 0), Xw(ls.n_src, 0), Dw(rows_cap, 0), Psiw(rows_cap, 0);
    auto slab_ok = [&](int64_t row, int64_t rows) { return row >= 0 && row < rows && (row << lg) + (R - 1) < rows * R; };

    // P2M (k_p2m): U[rid[b]] for boxes of the p2m list
    for (int32_t b : ls.p2m) {
        CHECK(b >= 0 && b < B, "p2m box %d", b);
        const int32_t row = ls.rid[b];
        CHECK(slab_ok(row, rows_cap), "p2m row %d of box %d (rows_cap %lld)", row, b, (long long)rows_cap);
        if (slab_ok(row, rows_cap)) Uw[row] = 1;
    }
    // upward (k_up), bottom-up
    for (int l = L - 1; l >= 1; --l) {
        const int off = ls.up_off[l], nn = ls.up_n[l], nm = ls.up_nm2m[l];
        const int grid = (nn + bpt - 1) / bpt;
        std::vector<int32_t> wrote;
        for (int blk = 0; blk < grid; ++blk)
            for (int j = 0; j < 64; ++j) {
                const int i = blk * bpt + (j >> lg);
                if (i >= nn) continue;
                CHECK((size_t)(off + i) < ls.urow.size(), "urow index %d", off + i);
                const int32_t b = ls.urow[off + i];
                CHECK(slab_ok(b, ls.n_src), "level %d up row %d (n_src %lld)", l, b, (long long)ls.n_src);
                if (i < nm) {
                    for (int q = 0; q < 4; ++q) {
                        const size_t ci = 4 * (size_t)(off + i) + q;
                        CHECK(ci < ls.crow.size(), "crow index %zu", ci);
                        const int32_t c = ls.crow[ci];
                        if (c >= 0) {
                            int32_t cb = -1;
                            for (int64_t bb = 0; bb < B && cb < 0; ++bb)
                                if (ls.rid[bb] == c) cb = (int32_t)bb;
                            CHECK(slab_ok(c, ls.n_src) && Xw[c], "level %d child row %d (box %d, level %d, fake %d, ns %lld, "
                                  "child0 %d) unwritten/out", l, c, cb, cb >= 0 ? h.lvl[cb] : -1, cb >= 0 ? (int)h.fake[cb] : -1,
                                  cb >= 0 ? (long long)h.ns(cb) : -1LL, cb >= 0 ? h.child0[cb] : -2);
                        }
                    }
                } else {
                    CHECK(slab_ok(b, rows_cap) && Uw[b], "level %d leaf row %d has no P2M", l, b);
                }
                wrote.push_back(b);
            }
        for (int32_t b : wrote)
            if (slab_ok(b, ls.n_src)) Xw[b] = 1;
    }
    // every box with sources has its upward expansion
    for (int64_t b = 0; b < B; ++b)
        if (h.lvl[b] >= 1 && h.ns((int32_t)b) > 0) CHECK(ls.rid[b] >= 0 && Xw[ls.rid[b]], "box %lld has no X", (long long)b);

    std::vector<Tile> tiles;
    std::vector<uint32_t> masks;
    std::vector<int> toff, tn;
    build_tile_table(ls, lg, mode, tiles, masks, toff, tn);
    CHECK((int64_t)tiles.size() == tile_count(ls, lg, mode), "tile count");
    // M2L (k_m2l): D[trow[t]] from X[vsrc[t][s]]
    for (size_t ti = 0; ti < tiles.size(); ++ti) {
        const Tile& T = tiles[ti];
        CHECK(T.count >= 1 && T.count <= bpt, "tile %zu count %d", ti, T.count);
        CHECK(T.first >= 0 && (size_t)(T.first + T.count) <= ls.trow.size(), "tile %zu range", ti);
        for (int k = 0; k < T.count; ++k) {
            const int t = T.first + k;
            const int32_t b = ls.trow[t];
            CHECK(slab_ok(b, rows_cap), "m2l target row %d (rows_cap %lld)", b, (long long)rows_cap);
            for (int s = 0; s < 27; ++s) {
                const int32_t src = ls.vsrc[(size_t)t * 27 + s];
                if (src < 0) continue;
                CHECK((masks[ti] >> s) & 1, "tile %zu mask misses slot %d", ti, s);
                CHECK(slab_ok(src, ls.n_src) && Xw[src], "m2l source row %d unwritten/out", src);
            }
            if (slab_ok(b, rows_cap)) Dw[b] = 1;
        }
    }
    // L2L + C2E (k_l2l_c2e), top-down
    for (int l = 2; l < L; ++l) {
        std::vector<int32_t> wrote;
        for (int ti = toff[l]; ti < toff[l] + tn[l]; ++ti) {
            const Tile& T = tiles[ti];
            CHECK(T.level == l, "tile %d of level %d in level %d's range", ti, T.level, l);
            for (int k = 0; k < T.count; ++k) {
                const int t = T.first + k;
                const int32_t b = ls.trow[t];
                CHECK(slab_ok(b, rows_cap) && Dw[b], "down row %d without D", b);
                if (l >= 3) {
                    const int32_t pr = ls.prow[t];
                    CHECK(slab_ok(pr, rows_cap) && Psiw[pr], "level %d target row %d: parent row %d without Psi", l, b, pr);
                }
                wrote.push_back(b);
            }
        }
        for (int32_t b : wrote)
            if (slab_ok(b, rows_cap)) Psiw[b] = 1;
    }
    // L2P (k_l2p) at the source leaves
    for (int32_t b : ls.leaves) {
        CHECK(h.leaf[b] && h.ns(b) > 0, "leaf list box %d", b);
        if (h.lvl[b] >= 2) CHECK(slab_ok(ls.rid[b], rows_cap) && Psiw[ls.rid[b]], "leaf %d without Psi", b);
    }
    for (int64_t b = 0; b < B; ++b)
        if (h.leaf[b] && !h.fake[b] && h.ns((int32_t)b) > 0) {
            bool found = false;
            for (int32_t c : ls.leaves) found |= c == b;
            CHECK(found, "leaf %lld with points not in the leaf list", (long long)b);
        }
    // arbitrary targets (k_targets) need every real leaf of levels >= 2
    if (mode == KF_TGT_ALL)
        for (int64_t b = 0; b < B; ++b)
            if (h.leaf[b] && !h.fake[b] && h.lvl[b] >= 2)
                CHECK(slab_ok(ls.rid[b], rows_cap) && Psiw[ls.rid[b]], "target leaf %lld without Psi", (long long)b);
    if (n_err) {
        std::printf("%d violations\n", n_err);
        return 1;
    }
    std::printf("OK levels %d boxes %lld n_src %lld n_all %lld tiles %zu\n", L, (long long)B, (long long)ls.n_src,
                (long long)ls.n_all, tiles.size());
    return 0;
}
