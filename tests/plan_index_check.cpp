// plan_index_check.cpp -- CPU replay of the planned evaluation's indexing (test tool).
//
// Builds the box trees (tree_host.cpp) and the work lists of one tree or a forest
// (plan_lists.cpp) exactly as the library does, then walks every pass of one evaluation the way
// its kernel indexes memory and in the library's launch order, for a batch width R and a target
// set, and checks that
//   * every list index and slab row is inside what the library allocates for that layout
//     (U and Psi: rows_cap x R rows, X: n_src x R rows; points and boxes of the forest);
//   * every slab row a pass reads was written earlier in the same evaluation;
//   * the results the caller reads (L2P at the source leaves; for the all-target set the
//     expansion of every real leaf) are complete.
// Each input file: int64 n, int32 cutoff, double root_min, double root_width, int32 key_depth,
// then n sorted uint64 Morton keys.
// Usage: plan_index_check <lg R> <mode> <keys file> [<keys file> ...]  (several files: a forest),
//        plan_index_check xcd  (the workgroup-to-item map of kf_xcd.h for every n < 5000).
// Prints "OK <counts>" or the first violations and exits non-zero.
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "kf_xcd.h"
#include "plan_lists.h"
#include "tree_host.h"

using namespace kf;

static int n_err = 0;
#define CHECK(cond, ...)                                         \
    do {                                                         \
        if (!(cond)) {                                           \
            if (n_err < 20) {                                    \
                std::printf("FAIL line %d: ", __LINE__);         \
                std::printf(__VA_ARGS__);                        \
                std::printf("\n");                               \
            }                                                    \
            ++n_err;                                             \
        }                                                        \
    } while (0)

static bool load_tree(const char* path, HostTree& h) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    int64_t n;
    int32_t cutoff, depth;
    double rmin, rw;
    bool ok = std::fread(&n, 8, 1, f) == 1 && std::fread(&cutoff, 4, 1, f) == 1 && std::fread(&rmin, 8, 1, f) == 1 &&
              std::fread(&rw, 8, 1, f) == 1 && std::fread(&depth, 4, 1, f) == 1;
    if (ok) {
        h.n = n;
        h.cutoff = cutoff;
        h.root_min = rmin;
        h.root_width = rw;
        h.key_depth = depth;
        h.keys.resize(n);
        ok = std::fread(h.keys.data(), 8, n, f) == (size_t)n;
    }
    std::fclose(f);
    return ok && h.build().empty();
}

// xcd_contiguous(., n) must be a bijection of [0, n) giving each XCD (b % 8) one contiguous range.
static int check_xcd() {
    for (int n = 1; n < 5000; ++n) {
        std::vector<int> hit(n, 0), lo(8, n), hi(8, -1), cnt(8, 0);
        for (int b = 0; b < n; ++b) {
            const int i = xcd_contiguous(b, n);
            CHECK(i >= 0 && i < n, "n %d: workgroup %d -> %d", n, b, i);
            if (i < 0 || i >= n) continue;
            hit[i] += 1;
            lo[b & 7] = std::min(lo[b & 7], i);
            hi[b & 7] = std::max(hi[b & 7], i);
            cnt[b & 7] += 1;
        }
        for (int i = 0; i < n; ++i) CHECK(hit[i] == 1, "n %d: item %d covered %d times", n, i, hit[i]);
        for (int x = 0; x < 8; ++x) CHECK(cnt[x] == 0 || hi[x] - lo[x] + 1 == cnt[x], "n %d: XCD %d not contiguous", n, x);
        // xcd_chunked(., n, K): a bijection whose whole groups give XCD b % 8 runs of K items
        for (int K : {1, 4, 16, 32}) {
            std::vector<int> seen(n, 0);
            for (int b = 0; b < n; ++b) {
                const int i = xcd_chunked(b, n, K);
                CHECK(i >= 0 && i < n, "n %d K %d: workgroup %d -> %d", n, K, b, i);
                if (i < 0 || i >= n) continue;
                seen[i] += 1;
                if (b < n - n % (8 * K)) CHECK((i / K) % 8 == (b & 7), "n %d K %d: item %d off its XCD run", n, K, i);
            }
            for (int i = 0; i < n; ++i) CHECK(seen[i] == 1, "n %d K %d: item %d covered %d times", n, K, i, seen[i]);
        }
        if (n_err) break;
    }
    if (n_err) return 1;
    std::printf("OK xcd\n");
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 2 && std::string(argv[1]) == "xcd") return check_xcd();
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s lgR mode keys.bin [keys.bin ...]\n", argv[0]);
        return 2;
    }
    const int lg = std::atoi(argv[1]), mode = std::atoi(argv[2]);
    std::vector<std::unique_ptr<HostTree>> owned;
    std::vector<const HostTree*> trees;
    for (int a = 3; a < argc; ++a) {
        owned.emplace_back(new HostTree());
        if (!load_tree(argv[a], *owned.back())) {
            std::printf("cannot build tree from %s\n", argv[a]);
            return 2;
        }
        trees.push_back(owned.back().get());
    }
    const int nt = (int)trees.size();
    const int R = 1 << lg, bpt = 64 >> lg;
    PlanLists ls;
    build_forest_lists(trees, 32, 32, ls);
    const int L = ls.L, NC = ls.n_classes;
    const int64_t B = (int64_t)ls.lvl.size();
    int64_t N = 0;
    for (const HostTree* t : trees) N += t->n;
    const int64_t rows_cap = mode == KF_TGT_ALL ? ls.n_all : ls.n_src;
    CHECK((int)ls.pt_off.size() == nt + 1 && ls.pt_off[nt] == N, "point offsets");
    CHECK((int)ls.box_off.size() == nt + 1 && ls.box_off[nt] == B, "box offsets");
    auto tree_of = [&](int64_t b) {
        int i = 0;
        while (i + 1 < nt && ls.box_off[i + 1] <= b) ++i;
        return i;
    };
    // per-box arrays and rows: boxes with sources at levels >= 1 -> [0, n_src);
    // real empty boxes at >= 2 -> [n_src, n_all)
    for (int64_t b = 0; b < B; ++b) {
        const int i = tree_of(b);
        const HostTree& h = *trees[i];
        const int32_t lb = (int32_t)(b - ls.box_off[i]);
        const int64_t po = ls.pt_off[i];
        CHECK(ls.lvl[b] == h.lvl[lb] && ls.bot[b] == h.bot[lb] + po && ls.top[b] == h.top[lb] + po,
              "box %lld arrays", (long long)b);
        CHECK(ls.bop[b] >= 0 && ls.bop[b] < NC * L && ls.bop[b] % L == ls.lvl[b], "box %lld op index %d",
              (long long)b, ls.bop[b]);
        CHECK(ls.child0[b] < 0 || (ls.child0[b] >= ls.box_off[i] && ls.child0[b] + 3 < ls.box_off[i + 1]),
              "box %lld children", (long long)b);
        for (int j = 0; j < 9; ++j) {
            const int32_t c = ls.coll[b * 9 + j];
            CHECK(c < 0 || (c >= ls.box_off[i] && c < ls.box_off[i + 1]), "box %lld colleague %d", (long long)b, c);
        }
        const int32_t r = ls.rid[b];
        if (h.lvl[lb] >= 1 && h.ns(lb) > 0) CHECK(r >= 0 && r < ls.n_src, "box %lld row %d", (long long)b, r);
        else if (h.lvl[lb] >= 2 && !h.fake[lb]) CHECK(r >= ls.n_src && r < ls.n_all, "empty box %lld row %d", (long long)b, r);
        else CHECK(r == -1, "box %lld should have no row (%d)", (long long)b, r);
    }
    std::vector<char> Uw(rows_cap, 0), Xw(ls.n_src, 0), Dw(rows_cap, 0), Psiw(rows_cap, 0);
    auto slab_ok = [&](int64_t row, int64_t rows) { return row >= 0 && row < rows && (row << lg) + (R - 1) < rows * R; };

    // P2M (k_p2m): U[rid[b]] for the boxes of the p2m list; sources bot..top
    for (int32_t b : ls.p2m) {
        CHECK(b >= 0 && b < B, "p2m box %d", b);
        if (b < 0 || b >= B) continue;
        CHECK(ls.bot[b] >= 0 && ls.top[b] <= N && ls.bot[b] < ls.top[b], "p2m box %d points", b);
        const int32_t row = ls.rid[b];
        CHECK(slab_ok(row, rows_cap), "p2m row %d of box %d (rows_cap %lld)", row, b, (long long)rows_cap);
        if (slab_ok(row, rows_cap)) Uw[row] = 1;
    }
    // upward (k_up), bottom-up, one launch per op index of the level
    for (int l = L - 1; l >= 1; --l) {
        std::vector<int32_t> wrote;
        for (int c = 0; c < NC; ++c) {
            const int op = c * L + l;
            const int off = ls.up_off[op], nn = ls.up_n[op], nm = ls.up_nm2m[op];
            const int grid = (nn + bpt - 1) / bpt;
            for (int blk = 0; blk < grid; ++blk)
                for (int j = 0; j < 64; ++j) {
                    const int i = blk * bpt + (j >> lg);
                    if (i >= nn) continue;
                    CHECK((size_t)(off + i) < ls.urow.size(), "urow index %d", off + i);
                    const int32_t b = ls.urow[off + i];
                    CHECK(slab_ok(b, ls.n_src), "op %d up row %d (n_src %lld)", op, b, (long long)ls.n_src);
                    if (i < nm) {
                        for (int q = 0; q < 4; ++q) {
                            const size_t ci = 4 * (size_t)(off + i) + q;
                            CHECK(ci < ls.crow.size(), "crow index %zu", ci);
                            const int32_t ch = ls.crow[ci];
                            if (ch >= 0) CHECK(slab_ok(ch, ls.n_src) && Xw[ch], "op %d child row %d unwritten/out", op, ch);
                        }
                    } else {
                        CHECK(slab_ok(b, rows_cap) && Uw[b], "op %d leaf row %d has no P2M", op, b);
                    }
                    wrote.push_back(b);
                }
        }
        for (int32_t b : wrote)
            if (slab_ok(b, ls.n_src)) Xw[b] = 1;
    }
    // k_up_cs: 16 * RB rows per workgroup, global row g -> list entry g >> lgR, rhs g & (R-1);
    // every (entry, rhs) of every op index exactly once
    for (int RB = 1; RB <= 2; RB *= 2)
        for (int op = 0; op < NC * L; ++op) {
            const int nn = ls.up_n[op];
            if (!nn) continue;
            const long rows = (long)nn << lg;
            const long grid = (rows + 16 * RB - 1) / (16 * RB);
            std::vector<int> hit((size_t)rows, 0);
            for (long blk = 0; blk < grid; ++blk)
                for (int j = 0; j < 16 * RB; ++j) {
                    const long g = 16L * RB * blk + j;
                    const long i = g >> lg;
                    if (i >= nn) continue;  // the kernel reads no list entry past n
                    hit[(size_t)g] += 1;
                }
            for (long g = 0; g < rows; ++g) CHECK(hit[(size_t)g] == 1, "k_up_cs RB %d op %d row %ld covered %d times", RB, op, g, hit[(size_t)g]);
        }
    // every box with sources has its upward expansion
    for (int64_t b = 0; b < B; ++b) {
        const int i = tree_of(b);
        const int32_t lb = (int32_t)(b - ls.box_off[i]);
        if (ls.lvl[b] >= 1 && trees[i]->ns(lb) > 0)
            CHECK(ls.rid[b] >= 0 && Xw[ls.rid[b]], "box %lld has no X", (long long)b);
    }

    std::vector<Tile> tiles;
    std::vector<uint32_t> masks;
    std::vector<int> toff, tn;
    build_tile_table(ls, lg, mode, tiles, &masks, toff, tn);
    CHECK((int64_t)tiles.size() == tile_count(ls, lg, mode), "tile count");
    // M2L (k_m2l, one launch): D[trow[t]] from X[vsrc[t][s]], operators of the tile's op index
    for (size_t ti = 0; ti < tiles.size(); ++ti) {
        const Tile& T = tiles[ti];
        CHECK(T.count >= 1 && T.count <= bpt, "tile %zu count %d", ti, T.count);
        CHECK(T.first >= 0 && (size_t)(T.first + T.count) <= ls.trow.size(), "tile %zu range", ti);
        CHECK(T.op >= 0 && T.op < NC * L && T.op % L >= 2 && T.q >= 0 && T.q < 4, "tile %zu op %d q %d", ti, T.op, T.q);
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
    // L2L + C2E (k_l2l_c2e), top-down, one launch per op index of the level
    for (int l = 2; l < L; ++l) {
        std::vector<int32_t> wrote;
        for (int c = 0; c < NC; ++c) {
            const int op = c * L + l;
            for (int ti = toff[op]; ti < toff[op] + tn[op]; ++ti) {
                const Tile& T = tiles[ti];
                CHECK(T.op == op, "tile %d of op %d in op %d's range", ti, T.op, op);
                for (int k = 0; k < T.count; ++k) {
                    const int t = T.first + k;
                    const int32_t b = ls.trow[t];
                    CHECK(slab_ok(b, rows_cap) && Dw[b], "down row %d without D", b);
                    if (l >= 3) {
                        const int32_t pr = ls.prow[t];
                        CHECK(slab_ok(pr, rows_cap) && Psiw[pr], "level %d target row %d: parent row %d without Psi",
                              l, b, pr);
                    }
                    wrote.push_back(b);
                }
            }
        }
        for (int32_t b : wrote)
            if (slab_ok(b, rows_cap)) Psiw[b] = 1;
    }
    // k_l2l_c2e_cs: tile blockIdx / (4 / RB), rows j0 + [0, 16 RB) of it, j0 = 16 RB (blockIdx %
    // (4 / RB)); every (target, rhs) of every tile exactly once
    for (int RB = 1; RB <= 2; RB *= 2) {
        const int SUB = 4 / RB;
        for (size_t ti = 0; ti < tiles.size(); ++ti) {
            const Tile& T = tiles[ti];
            std::vector<int> hit((size_t)T.count << lg, 0);
            for (int sub = 0; sub < SUB; ++sub) {
                const int j0 = 16 * RB * sub;
                if ((j0 >> lg) >= T.count) continue;  // the workgroup returns at once
                for (int j = j0; j < j0 + 16 * RB; ++j) {
                    CHECK(j < 64, "sub-tile row %d", j);
                    const int k = j >> lg;
                    if (k < T.count) hit[(size_t)j] += 1;
                }
            }
            for (size_t j = 0; j < hit.size(); ++j)
                CHECK(hit[j] == 1, "k_l2l_c2e_cs RB %d tile %zu row %zu covered %d times", RB, ti, j, hit[j]);
        }
    }
    // P2P (k_p2p) + L2P (k_l2p) at the source leaves
    size_t n_leaf = 0;
    for (size_t k = 0; k < ls.leaves.size(); ++k) {
        const int32_t b = ls.leaves[k];
        CHECK(b >= 0 && b < B && ls.leaf[b] && ls.top[b] > ls.bot[b], "leaf list box %d", b);
        if (b < 0 || b >= B) continue;
        for (int j = 0; j < 9; ++j) {
            const int32_t c = ls.lcoll[k * 9 + j];
            CHECK(c < 0 || (c < B && ls.bot[c] >= 0 && ls.top[c] <= N), "leaf %d colleague %d", b, c);
        }
        if (ls.lvl[b] >= 2) CHECK(slab_ok(ls.rid[b], rows_cap) && Psiw[ls.rid[b]], "leaf %d without Psi", b);
        n_leaf += 1;
    }
    size_t want = 0;
    for (int64_t b = 0; b < B; ++b) {
        const int i = tree_of(b);
        const int32_t lb = (int32_t)(b - ls.box_off[i]);
        if (trees[i]->leaf[lb] && !trees[i]->fake[lb] && trees[i]->ns(lb) > 0) ++want;
        // arbitrary targets (k_targets) need every real leaf of levels >= 2
        if (mode == KF_TGT_ALL && trees[i]->leaf[lb] && !trees[i]->fake[lb] && ls.lvl[b] >= 2)
            CHECK(slab_ok(ls.rid[b], rows_cap) && Psiw[ls.rid[b]], "target leaf %lld without Psi", (long long)b);
    }
    CHECK(n_leaf == want, "leaf list has %zu of %zu leaves with points", n_leaf, want);
    // symmetric near field of one right-hand side (near_sym.hip; leaves of at most 64 points):
    // the pair table k_leaf_pairs builds, every partial slot k_p2p_sym writes (inside the 1 + 11
    // slots of N behind the potentials, each written once), every slot k_near_sum reads written
    // before, and every point's near field made of exactly its colleagues' points but itself
    // (count, sum and sum of squares of the source indices through the slots).
    int64_t max_leaf = 0;
    for (int32_t b : ls.leaves) max_leaf = std::max<int64_t>(max_leaf, ls.top[b] - ls.bot[b]);
    const bool sym_checked = lg == 0 && mode == KF_TGT_SOURCES && max_leaf <= 64;
    if (sym_checked) {
        const size_t nl = ls.leaves.size();
        std::vector<int32_t> lidx(B, -1), pinfo(9 * nl, -1);
        for (size_t k = 0; k < nl; ++k) lidx[ls.leaves[k]] = (int32_t)k;
        for (size_t e = 0; e < 9 * nl; ++e) {
            const int32_t b = ls.leaves[e / 9], c = ls.lcoll[e];
            if (c < 0 || c == b || lidx[c] < 0) continue;
            for (int jj = 0; jj < 9; ++jj)
                if (ls.lcoll[9 * (size_t)lidx[c] + jj] == b) pinfo[e] = lidx[c] * 16 + jj;
        }
        for (size_t e = 0; e < 9 * nl; ++e)
            if (pinfo[e] >= 0) {
                const size_t back = 9 * (size_t)(pinfo[e] >> 4) + (pinfo[e] & 15);
                CHECK(pinfo[back] == (int32_t)((e / 9) * 16 + e % 9), "pair table not symmetric at %zu", e);
            }
        constexpr int NS = 11;  // KF_SYM_SLOTS
        struct Tally {
            int64_t n = 0;
            double s1 = 0, s2 = 0;
            void add(int64_t src) {
                n += 1;
                s1 += (double)src;
                s2 += (double)src * (double)src;
            }
            void add(const Tally& o) {
                n += o.n;
                s1 += o.s1;
                s2 += o.s2;
            }
        };
        std::vector<Tally> slot((size_t)NS * N);
        std::vector<int> wrote((size_t)NS * N, 0);
        for (size_t li = 0; li < nl; ++li) {  // k_p2p_sym, one wave per leaf
            const int32_t b = ls.leaves[li], ab = ls.bot[b];
            const int na = ls.top[b] - ab;
            // the own block lane by lane (sym_block<SYM_SELF>): a lane that evaluates (on) pairs
            // a real target with another point, and every pair a < b is evaluated once
            std::vector<int> seen((size_t)na * na, 0);
            for (int jb = 0; 16 * jb < na; ++jb)
                for (int ia = 0; ia < 16 && 4 * ia < na && 4 * ia < 16 * jb + 15; ++ia)
                    for (int lane = 0; lane < 64; ++lane) {
                        const int a = 4 * ia + (lane >> 4), bs = 16 * jb + (lane & 15);
                        const int at_pt = std::min(a, na - 1), b_pt = std::min(bs, na - 1);  // padded copies
                        const bool on = a < bs && bs < na;
                        if (!on) continue;
                        CHECK(a < na && at_pt != b_pt, "own block of leaf %d: lane pairs point %d with %d", b, at_pt, b_pt);
                        if (a < na && bs < na) seen[(size_t)a * na + bs] += 1;
                    }
            for (int a = 0; a < na; ++a)
                for (int s = a + 1; s < na; ++s)
                    CHECK(seen[(size_t)a * na + s] == 1, "own block of leaf %d: pair %d,%d evaluated %d times", b, a, s,
                          seen[(size_t)a * na + s]);
            std::vector<Tally> rows(na);
            for (int a = 0; a < na; ++a)
                for (int s = a + 1; s < na; ++s) {  // own block: pairs a < s
                    rows[a].add(ab + s);
                    slot[(size_t)1 * N + ab + s].add(ab + a);
                }
            for (int s = 0; s < na; ++s) wrote[(size_t)1 * N + ab + s] += 1;
            for (int j = 0; j < 9; ++j) {
                const int32_t c = ls.lcoll[9 * li + j];
                if (c < 0 || c == b) continue;
                const int32_t c0 = ls.bot[c];
                const int nb = ls.top[c] - c0;
                if (nb <= 0) continue;
                const int32_t pi = pinfo[9 * li + j];
                if (pi >= 0 && (pi >> 4) < (int32_t)li) continue;  // the partner owns the pair
                for (int a = 0; a < na; ++a)
                    for (int s = 0; s < nb; ++s) rows[a].add(c0 + s);
                if (pi >= 0) {
                    const int sl = 2 + (pi & 15);
                    CHECK(sl < NS, "slot %d", sl);
                    for (int s = 0; s < nb; ++s) {
                        CHECK(c0 + s >= 0 && c0 + s < N, "column point %d", c0 + s);
                        wrote[(size_t)sl * N + c0 + s] += 1;
                        for (int a = 0; a < na; ++a) slot[(size_t)sl * N + c0 + s].add(ab + a);
                    }
                }
            }
            for (int a = 0; a < na; ++a) {
                slot[ab + a].add(rows[a]);
                wrote[ab + a] += 1;
            }
        }
        for (size_t e = 0; e < wrote.size(); ++e) CHECK(wrote[e] <= 1, "slot %zu point %zu written %d times",
                                                       e / N, e % N, wrote[e]);
        for (size_t li = 0; li < nl; ++li) {  // k_near_sum
            const int32_t b = ls.leaves[li], tb = ls.bot[b];
            const int nt = ls.top[b] - tb;
            Tally want;
            for (int j = 0; j < 9; ++j) {
                const int32_t c = ls.lcoll[9 * li + j];
                if (c >= 0)
                    for (int32_t s = ls.bot[c]; s < ls.top[c]; ++s) want.add(s);
            }
            for (int t = 0; t < nt; ++t) {
                Tally got;
                for (int sl = 0; sl < NS; ++sl) {
                    bool read = sl < 2;
                    if (sl >= 2) {
                        const int32_t pi = pinfo[9 * li + (sl - 2)];
                        read = pi >= 0 && (pi >> 4) < (int32_t)li;
                    }
                    if (!read) continue;
                    CHECK(wrote[(size_t)sl * N + tb + t] == 1, "slot %d of point %d read unwritten", sl, tb + t);
                    got.add(slot[(size_t)sl * N + tb + t]);
                }
                const double self = tb + t;
                CHECK(got.n == want.n - 1 && got.s1 == want.s1 - self && got.s2 == want.s2 - self * self,
                      "near field of point %d: %lld sources of %lld", tb + t, (long long)got.n, (long long)want.n - 1);
            }
        }
    }
    // near-field symmetry: pair work between leaves that are each other's colleagues
    if (std::getenv("KF_LANE_STATS")) {
        // work-weighted lane utilisation of the near field, one or two targets per lane
        double w1 = 0, w2 = 0, wb = 0, tot = 0;
        for (size_t k = 0; k < ls.leaves.size(); ++k) {
            const int32_t b = ls.leaves[k];
            int64_t src = 0;
            for (int j = 0; j < 9; ++j) {
                const int32_t c = ls.lcoll[k * 9 + j];
                if (c >= 0) src += ls.top[c] - ls.bot[c];
            }
            for (int64_t t0 = 0; t0 < ls.top[b] - ls.bot[b]; t0 += 64) {
                const int cnt = (int)std::min<int64_t>(64, ls.top[b] - ls.bot[b] - t0);
                const double work = (double)cnt * src;
                const double u1 = (double)((64 / cnt) * cnt) / 64.0;
                const int lg = (cnt + 1) / 2;
                const double u2 = (double)((64 / lg) * cnt) / 128.0;
                tot += work;
                w1 += work * u1;
                w2 += work * u2;
                wb += work * std::max(u1, u2);
            }
        }
        std::printf("lane utilisation: one target per lane %.4f, two %.4f, best of both %.4f\n", w1 / tot, w2 / tot,
                    wb / tot);
    }
    if (std::getenv("KF_SYM_STATS")) {
        std::vector<int> is_leaf_listed(B, 0);
        std::vector<std::vector<int32_t>> cl(B);
        for (size_t k = 0; k < ls.leaves.size(); ++k) {
            is_leaf_listed[ls.leaves[k]] = 1;
            for (int j = 0; j < 9; ++j) cl[ls.leaves[k]].push_back(ls.lcoll[k * 9 + j]);
        }
        double mutual = 0, total = 0;
        for (size_t k = 0; k < ls.leaves.size(); ++k) {
            const int32_t b = ls.leaves[k];
            const double nb = ls.top[b] - ls.bot[b];
            for (int j = 0; j < 9; ++j) {
                const int32_t c = ls.lcoll[k * 9 + j];
                if (c < 0) continue;
                const double w = nb * (ls.top[c] - ls.bot[c]);
                total += w;
                bool mut = c == b;
                if (!mut && is_leaf_listed[c])
                    for (int32_t d : cl[c]) mut = mut || d == b;
                if (mut) mutual += w;
            }
        }
        std::printf("near-field pair work between mutual leaf colleagues: %.4f of %.4g\n", mutual / total, total);
    }
    // stored kernel matrices (kf_plan_store_matrices): each block holds exactly what its fill
    // kernel writes and its stored-form kernel reads -- near field (k_p2p_store / k_p2p_stored)
    // column s * nt + t for the nt targets and ns = own + other colleagues' sources, walked in
    // near_source order; L2P (k_l2p_store / k_l2p) jj * nt + i for leaves at level >= 2; P2M
    // (k_p2m_store / k_p2m) k * p + j -- and the totals are the plan's evaluation counts
    {
        StoredOffsets so;
        stored_offsets(ls, 32, so);
        const size_t nl = ls.leaves.size();
        CHECK(so.near.size() == nl + 1 && so.l2p.size() == nl + 1 && so.p2m.size() == ls.p2m.size() + 1,
              "stored offset sizes");
        for (size_t li = 0; li + 1 < so.near.size(); ++li) {
            const int32_t b = ls.leaves[li];
            const int64_t nt = ls.top[b] - ls.bot[b];
            int64_t ns = nt, n_self = 0;
            for (int j = 0; j < 9; ++j) {
                const int32_t c = ls.lcoll[li * 9 + j];
                if (c == b) ++n_self;
                if (c < 0 || c == b) continue;
                CHECK(ls.bot[c] >= 0 && ls.top[c] <= N && ls.bot[c] <= ls.top[c], "leaf %d colleague %d points", b, c);
                ns += ls.top[c] - ls.bot[c];
            }
            CHECK(n_self == 1, "leaf %d is %lld times its own colleague", b, (long long)n_self);
            CHECK(so.near[li + 1] - so.near[li] == nt * ns, "near block of leaf %d", b);
            CHECK(so.l2p[li + 1] - so.l2p[li] == (ls.lvl[b] >= 2 ? nt * 32 : 0), "L2P block of leaf %d", b);
            CHECK(ls.bot[b] >= 0 && ls.top[b] <= N && nt > 0, "leaf %d points", b);
        }
        for (size_t i = 0; i + 1 < so.p2m.size(); ++i) {
            const int32_t b = ls.p2m[i];
            CHECK(so.p2m[i + 1] - so.p2m[i] == (int64_t)(ls.top[b] - ls.bot[b]) * 32, "P2M block of box %d", b);
        }
        CHECK(so.near.back() == ls.stats.p2p_pairs + N, "near total %lld, pairs %lld + points %lld",
              (long long)so.near.back(), (long long)ls.stats.p2p_pairs, (long long)N);
        CHECK(so.p2m.back() == ls.stats.p2m_evals, "P2M total");
        CHECK(so.l2p.back() == ls.stats.l2p_evals, "L2P total");
    }
    if (n_err) {
        std::printf("%d violations\n", n_err);
        return 1;
    }
    std::printf("OK trees %d classes %d levels %d boxes %lld n_src %lld n_all %lld tiles %zu%s\n", nt, NC, L,
                (long long)B, (long long)ls.n_src, (long long)ls.n_all, tiles.size(),
                sym_checked ? " symmetric near field" : "");
    return 0;
}
