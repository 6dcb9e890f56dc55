/*
 * oracle/ref_loops.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C (OpenMP) restatement of the reference's numba loops for the 2-D KIFMM.
 * It is the parity checker for the HIP product path and the timed CPU baseline
 * ("port").  Nothing in the product package links or loads this file.
 *
 * Each function names the reference loop it restates (paths are relative to the
 * reference repository root).  Index types are int64 like the reference's i8 arrays.
 *
 * Kernels (kind): 0 = Laplace  G = -(1/4pi) log(r^2)           kernels/laplace.py:5-8
 *                 1 = modified Helmholtz  G = scale*K0(k r)     kernels/modified_helmholtz.py:7-9
 * K0 comes from a function pointer installed by the Python side (scipy's Cephes k0,
 * the same Cephes method as misc/numba_special_functions.py:167-181).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double (*k0_fn_t)(double, int);
static k0_fn_t g_k0 = 0;

void orc_set_k0(void *fn) { g_k0 = (k0_fn_t)fn; }

static inline double kern(int kind, double kk, double scale,
                          double sx, double sy, double tx, double ty) {
    double dx = tx - sx, dy = ty - sy;
    if (kind == 0) return (-0.25 / M_PI) * log(dx * dx + dy * dy);
    return scale * g_k0(kk * sqrt(dx * dx + dy * dy), 0);
}

/* ---------------------------------------------------------------- tree ---- */

/* tree.py:14-126 (classify, get_target, swap, get_inds, reorder_inplace):
 * in-place cycle partition of one box's slice into the 4 classes 2*(x>midx)+(y>midy).
 * Returns the class start offsets in nns[4]. */
static int reorder_slice(double *x, double *y, int64_t *ordv, int64_t n,
                         double midx, double midy, int64_t nns[4]) {
    int8_t *cl = (int8_t *)malloc(n > 0 ? (size_t)n : 1);
    if (!cl) return -1;
    int64_t cnt[4] = {0, 0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        int c = 2 * (x[i] > midx) + (y[i] > midy);
        cl[i] = (int8_t)c;
        cnt[c] += 1;
    }
    int64_t next[4];
    next[0] = 0;
    for (int c = 0; c < 3; ++c) next[c + 1] = next[c] + cnt[c];
    for (int c = 0; c < 4; ++c) nns[c] = next[c];
    for (int64_t i = 0; i < n; ++i) {
        int want = (i < nns[1]) ? 0 : (i < nns[2]) ? 1 : (i < nns[3]) ? 2 : 3;
        for (;;) {
            int c = cl[i];
            if (c == want) { next[want] += 1; break; }
            int64_t j = next[c];
            if (j >= n) { free(cl); return -2; }   /* cannot happen for a valid partition */
            int8_t tc = cl[i]; cl[i] = cl[j]; cl[j] = tc;
            double t = x[i]; x[i] = x[j]; x[j] = t;
            t = y[i]; y[i] = y[j]; y[j] = t;
            int64_t to = ordv[i]; ordv[i] = ordv[j]; ordv[j] = to;
            next[c] += 1;
        }
    }
    free(cl);
    return 0;
}

/* tree.py:128-190 divide_and_reorder: split every flagged box of a level, reorder the
 * points of each split box in place and emit the children's geometry and links. */
int orc_divide_and_reorder(double *x, double *y, int64_t *ordv, const uint8_t *tosplit,
                           double half_width, const double *xmid, const double *ymid,
                           const int64_t *bot, const int64_t *top, uint8_t *leaf,
                           double *new_xmin, double *new_ymin, int64_t *new_bot,
                           int64_t *new_top, int64_t *parent_ind, int64_t *children_ind,
                           int64_t children_start, int xlist, int64_t num_nodes) {
    int64_t *sid = (int64_t *)malloc((num_nodes > 0 ? num_nodes : 1) * sizeof(int64_t));
    if (!sid) return -1;
    int64_t k = 0;
    for (int64_t i = 0; i < num_nodes; ++i) { sid[i] = k; if (tosplit[i]) ++k; }
    int err = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : err)
    for (int64_t i = 0; i < num_nodes; ++i) {
        if (!tosplit[i]) continue;
        int64_t s = sid[i], bi = bot[i], ti = top[i], nns[4];
        err |= reorder_slice(x + bi, y + bi, ordv + bi, ti - bi, xmid[i], ymid[i], nns) != 0;
        for (int c = 0; c < 4; ++c) {
            new_xmin[4 * s + c] = (c < 2) ? xmid[i] - half_width : xmid[i];
            new_ymin[4 * s + c] = (c & 1) ? ymid[i] : ymid[i] - half_width;
            new_bot[4 * s + c] = bi + nns[c];
            new_top[4 * s + c] = (c < 3) ? bi + nns[c + 1] : ti;
            parent_ind[4 * s + c] = i;
        }
        if (!xlist) leaf[i] = 0;
        children_ind[i] = children_start + 4 * s;
    }
    free(sid);
    return err ? -2 : 0;
}

/* tree.py:260-272 numba_tag_colleagues: direct O(n^2) colleague search. */
void orc_tag_colleagues_direct(const double *xmid, const double *ymid, int64_t n,
                               double dist, int64_t *coll) {
    double d2max = dist * dist;
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
        int slot = 0;
        for (int64_t j = 0; j < n; ++j) {
            double dx = xmid[i] - xmid[j], dy = ymid[i] - ymid[j];
            if (dx * dx + dy * dy < d2max && slot < 9) coll[9 * i + slot++] = j;
        }
    }
}

/* tree.py:274-294 numba_loop_colleagues: colleagues through the parent's colleagues'
 * children; boxes with children_ind >= 0 count as having children (tree.py:371). */
void orc_loop_colleagues(const double *xmid, const double *ymid, int64_t n, double dist,
                         const int64_t *parent_ind, const int64_t *anc_coll,
                         const int64_t *anc_children, int64_t *coll) {
    double d2max = dist * dist;
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
        int slot = 0;
        int64_t pi = parent_ind[i];
        for (int j = 0; j < 9; ++j) {
            int64_t pj = anc_coll[9 * pi + j];
            if (pj < 0 || anc_children[pj] < 0) continue;
            int64_t c0 = anc_children[pj];
            for (int kq = 0; kq < 4; ++kq) {
                int64_t c = c0 + kq;
                double dx = xmid[i] - xmid[c], dy = ymid[i] - ymid[c];
                if (dx * dx + dy * dy < d2max && slot < 9) coll[9 * i + slot++] = c;
            }
        }
    }
}

/* tree.py:423-430 numba_get_depths */
void orc_get_depths(int64_t *depths, const uint8_t *leaf, const int64_t *children_ind,
                    const int64_t *desc_depths, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        if (leaf[i]) continue;
        int64_t c = children_ind[i], m = desc_depths[c];
        for (int q = 1; q < 4; ++q) if (desc_depths[c + q] > m) m = desc_depths[c + q];
        depths[i] = m + 1;
    }
}

/* tree.py:432-444 numba_get_bads */
void orc_get_bads(const int64_t *depths, const int64_t *coll, const uint8_t *leaf,
                  uint8_t *bads, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        if (!leaf[i]) continue;
        int bad = 0;
        for (int j = 0; j < 9; ++j) {
            int64_t c = coll[9 * i + j];
            if (c >= 0) { int64_t d = depths[i] - depths[c]; if (d > 1 || d < -1) bad = 1; }
        }
        bads[i] = (uint8_t)bad;
    }
}

/* tree.py:446-458 numba_get_Xlist: only colleagues with index < n (the boxes that existed
 * when depths were gathered) are considered. */
void orc_get_xlist(const int64_t *depths, int64_t n, const int64_t *coll,
                   const uint8_t *leaf, uint8_t *xl) {
    for (int64_t i = 0; i < n; ++i) {
        if (!leaf[i]) continue;
        int v = 0;
        for (int j = 0; j < 9; ++j) {
            int64_t c = coll[9 * i + j];
            if (c >= 0 && c < n && depths[i] - depths[c] < 0) v = 1;
        }
        xl[i] = (uint8_t)v;
    }
}

/* tree.py:712-727 numba_locate_points, flattened: level arrays are concatenated and
 * lvl_off[l] is the first box of level l. */
void orc_locate_points(const double *tx, const double *ty, int64_t n, const uint8_t *leaf,
                       const double *xmid, const double *ymid, const int64_t *children,
                       const int64_t *lvl_off, int64_t *inds, int64_t *locs) {
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
        int64_t ind = 0, loc = 0;
        while (!leaf[lvl_off[ind] + loc]) {
            int64_t g = lvl_off[ind] + loc;
            int64_t c = children[g];
            if (tx[i] > xmid[g]) c += 2;
            if (ty[i] > ymid[g]) c += 1;
            loc = c;
            ind += 1;
        }
        inds[i] = ind;
        locs[i] = loc;
    }
}

/* ------------------------------------------------------------- passes ---- */

/* fmm.py:30-64 evaluate_neighbor_interactions: leaf-to-colleague direct sums with the
 * index self-exclusion (i != ci or ks != kt). */
void orc_neighbor_interactions(int kind, double kk, double scale, const double *x,
                               const double *y, const uint8_t *leaf, const int64_t *bot,
                               const int64_t *top, const double *tau, const int64_t *coll,
                               double *sol, int64_t n) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < n; ++i) {
        if (!leaf[i]) continue;
        int64_t b1 = bot[i], nt = top[i] - b1;
        if (nt <= 0) continue;
        double *acc = (double *)calloc((size_t)nt, sizeof(double));
        for (int j = 0; j < 9; ++j) {
            int64_t ci = coll[9 * i + j];
            if (ci < 0) continue;
            int64_t b2 = bot[ci], ns = top[ci] - b2;
            for (int64_t ks = 0; ks < ns; ++ks)
                for (int64_t kt = 0; kt < nt; ++kt)
                    if (i != ci || ks != kt)
                        acc[kt] += tau[b2 + ks] *
                                   kern(kind, kk, scale, x[b2 + ks], y[b2 + ks], x[b1 + kt], y[b1 + kt]);
        }
        for (int64_t kt = 0; kt < nt; ++kt) sol[b1 + kt] += acc[kt];
        free(acc);
    }
}

/* class_fmm.py:57-71 numba_upwards_pass (P2M into density rows li[i]). */
void orc_upwards_pass(int kind, double kk, double scale, const double *x, const double *y,
                      const int64_t *bot, const int64_t *top, const uint8_t *cu,
                      const double *xt, const double *yt, int64_t ne, const double *xmid,
                      const double *ymid, const double *tau, double *ucheck, const int64_t *li,
                      int64_t n) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = li[i];
        if (!cu[i] || r < 0) continue;
        for (int64_t ks = bot[i]; ks < top[i]; ++ks)
            for (int64_t kt = 0; kt < ne; ++kt)
                ucheck[r * ne + kt] += kern(kind, kk, scale, x[ks], y[ks], xt[kt] + xmid[i],
                                            yt[kt] + ymid[i]) * tau[ks];
    }
}

/* fmm.py:91-115 numba_downwards_pass2 (L2P for leaves, box-relative targets). */
void orc_downwards_pass2(int kind, double kk, double scale, const double *x, const double *y,
                         const int64_t *bot, const int64_t *top, const uint8_t *leaf,
                         const double *xs, const double *ys, int64_t ne, const double *xmid,
                         const double *ymid, const double *loc, double *sol, int64_t n) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < n; ++i) {
        if (!leaf[i] || top[i] - bot[i] <= 0) continue;
        for (int64_t t = bot[i]; t < top[i]; ++t) {
            double tx = x[t] - xmid[i], ty = y[t] - ymid[i], acc = 0.0;
            for (int64_t ks = 0; ks < ne; ++ks)
                acc += kern(kind, kk, scale, xs[ks], ys[ks], tx, ty) * loc[i * ne + ks];
            sol[t] += acc;
        }
    }
}

/* class_fmm.py:11-15 numba_distribute: ucs[li2[pi[g]]] = temp[li[g]] (assignment). */
void orc_distribute(double *ucs, const double *temp, const int64_t *pi, const int64_t *li,
                    const int64_t *li2, int64_t ngroups, int64_t width) {
    for (int64_t g = 0; g < ngroups; ++g)
        if (li[g] >= 0)
            memcpy(ucs + li2[pi[g]] * width, temp + li[g] * width, (size_t)width * sizeof(double));
}

/* class_fmm.py:17-33 numba_add_interactions.  m2ms is [3][3][4p][nsrc] (C order);
 * ci4 is trunc(children_ind/4), so a box without children aliases group 0. */
void orc_add_interactions(const uint8_t *doit, const int64_t *ci4, const int64_t *coll,
                          const double *xmid, const double *ymid, double *local_sol,
                          const double *m2ms, int64_t p, const int64_t *li, int64_t nsrc,
                          int64_t n) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < n; ++i) {
        if (!doit[i]) continue;
        int64_t dii = ci4[i];
        for (int j = 0; j < 9; ++j) {
            int64_t ci = coll[9 * i + j];
            if (ci < 0 || ci == i) continue;
            double ddx = xmid[ci] - xmid[i], ddy = ymid[ci] - ymid[i];
            int xd = (ddx > 0) - (ddx < 0), yd = (ddy > 0) - (ddy < 0);
            int64_t di = li[ci4[ci]];
            if (di < 0) continue;
            const double *blk = m2ms + (size_t)((xd + 1) * 3 + (yd + 1)) * (4 * p) * nsrc;
            for (int k = 0; k < 4; ++k)
                for (int64_t ll = 0; ll < p; ++ll)
                    local_sol[(4 * dii + k) * p + ll] += blk[(k * p + ll) * nsrc + di];
        }
    }
}

/* class_fmm.py:73-88 numba_target_local_expansion_evaluation (flattened level arrays;
 * expansions are stored per box of every level >= 1, row = lvl_off[ind]+loc). */
void orc_target_l2p(int kind, double kk, double scale, const double *tx, const double *ty,
                    int64_t n, const int64_t *inds, const int64_t *locs, const double *lx,
                    const double *ly, int64_t ne, const double *xmid, const double *ymid,
                    const int64_t *lvl_off, const double *expans, double *pot) {
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
        int64_t g = lvl_off[inds[i]] + locs[i];
        double rx = tx[i] - xmid[g], ry = ty[i] - ymid[g], acc = 0.0;
        const double *ex = expans + g * ne;
        const double *lxx = lx + inds[i] * ne, *lyy = ly + inds[i] * ne;
        for (int64_t k = 0; k < ne; ++k) acc += kern(kind, kk, scale, lxx[k], lyy[k], rx, ry) * ex[k];
        pot[i] = acc;
    }
}

/* class_fmm.py:90-111 numba_target_neighbor_evaluation; check_self excludes exactly
 * coincident source points. */
void orc_target_p2p(int kind, double kk, double scale, const double *tx, const double *ty,
                    int64_t n, const double *sx, const double *sy, const int64_t *inds,
                    const int64_t *locs, const int64_t *bot, const int64_t *top,
                    const int64_t *coll, const int64_t *lvl_off, const double *tau,
                    double *pot, int check) {
#pragma omp parallel for
    for (int64_t i = 0; i < n; ++i) {
        int64_t off = lvl_off[inds[i]], g = off + locs[i];
        double acc = 0.0;
        for (int j = 0; j < 9; ++j) {
            int64_t ci = coll[9 * g + j];
            if (ci < 0) continue;
            int64_t c = off + ci;
            for (int64_t ks = bot[c]; ks < top[c]; ++ks) {
                if (check && tx[i] - sx[ks] == 0.0 && ty[i] - sy[ks] == 0.0) continue;
                acc += kern(kind, kk, scale, sx[ks], sy[ks], tx[i], ty[i]) * tau[ks];
            }
        }
        pot[i] += acc;
    }
}

/* fmm.py:119-127 KF: out[j,i] = G(s_i, t_j), out is (nt, ns) C order. */
void orc_kernel_form(int kind, double kk, double scale, const double *sx, const double *sy,
                     int64_t ns, const double *tx, const double *ty, int64_t nt, double *out) {
#pragma omp parallel for
    for (int64_t j = 0; j < nt; ++j)
        for (int64_t i = 0; i < ns; ++i)
            out[j * ns + i] = kern(kind, kk, scale, sx[i], sy[i], tx[j], ty[j]);
}

/* fmm.py:128-147 KA / KAS direct sums.  self_mode excludes i == j (KAS).  The sum over
 * sources is compensated (Neumaier) so the result is an FP64 direct-sum oracle. */
void orc_direct(int kind, double kk, double scale, const double *sx, const double *sy,
                const double *tau, int64_t ns, const double *tx, const double *ty, int64_t nt,
                const int64_t *tidx, int self_mode, double *out) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t j = 0; j < nt; ++j) {
        double s = 0.0, c = 0.0;
        int64_t jj = tidx ? tidx[j] : j;
        for (int64_t i = 0; i < ns; ++i) {
            if (self_mode && i == jj) continue;
            double v = kern(kind, kk, scale, sx[i], sy[i], tx[j], ty[j]) * tau[i];
            double t = s + v;
            c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
            s = t;
        }
        out[j] = s + c;
    }
}
