/*
 * kifmm2d.h -- C-ABI of libkifmm2d, the MI355X (gfx950) 2-D kernel-independent FMM.
 *
 * Conventions
 *   - Pointers named d_* are device (HBM) pointers owned by the caller; everything else is
 *     host memory.  Handles (kf_tree, kf_plan) are owned by the library.
 *   - Every function returns KF_OK (0) or a negative KF_ERR_* code; kf_last_error() gives a
 *     thread-local message for the last failure on the calling thread.
 *   - `stream` is a hipStream_t (NULL = default stream).  Calls are stream-ordered; a function
 *     that returns host data synchronises the stream first.
 *   - One plan may be used by one stream at a time; different plans may run concurrently.
 *     No floating-point atomics are used: results are bitwise deterministic.
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * reference repository dbstein/pykifmm2d).
 */
#ifndef KIFMM2D_H
#define KIFMM2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KF_VERSION 100

#define KF_OK 0
#define KF_ERR_ARG (-1)      /* invalid argument */
#define KF_ERR_HIP (-2)      /* HIP runtime error */
#define KF_ERR_DEPTH (-3)    /* tree deeper than the Morton key resolution */
#define KF_ERR_STATE (-4)    /* call order (e.g. targets before expansions) */
#define KF_ERR_ALLOC (-5)    /* allocation failure */
#define KF_ERR_DOMAIN (-6)   /* points outside the reference root box rule */

#define KF_LAPLACE 0         /* G = -(1/4pi) log r^2          kernels/laplace.py:5-8 */
#define KF_MODHELM 1         /* G = scale * K0(k r)           kernels/modified_helmholtz.py:7-9 */
#define KF_TABLE 2           /* G = a user's radial kernel, tabulated (kf_ktable)   fmm.py:119 */

#define KF_PREC_FP64 0       /* every pass in FP64 */
#define KF_PREC_SPLIT 1      /* FP32 M2L (per-level scaled operators), every other pass FP64 */
#define KF_PREC_SPLIT16 2    /* M2L products from FP16 halves on the matrix cores (hi*hi + lo*hi
                                + hi*lo of scaled hi/lo splits, FP32 sums per slot, FP64 across
                                slots), every other pass FP64 */

typedef struct kf_tree kf_tree;
typedef struct kf_plan kf_plan;
typedef struct kf_workspace kf_workspace;
typedef struct kf_ktable kf_ktable;

typedef struct {
    int32_t kind;             /* KF_LAPLACE, KF_MODHELM or KF_TABLE */
    double k;                 /* modified Helmholtz wavenumber */
    double scale;             /* modified Helmholtz prefactor (1 in the reference module, 1/(2pi) in its examples) */
    const kf_ktable* table;   /* KF_TABLE: the tabulated kernel (NULL otherwise); plans keep a copy */
    const kf_ktable* table_im; /* KF_TABLE, complex kernels (class_fmm.py:167, 174 iscomplex; the H0
                                  kernel of examples/hankel.py:31-39): G = table + i table_im; NULL
                                  for a real kernel.  Plans of a complex kernel take real densities
                                  and give complex potentials (interleaved re, im doubles); the
                                  direct sums and kernel forms take the two parts as two real
                                  kernels. */
} kf_kernel_desc;

typedef struct {
    int32_t levels;      /* reference Tree.levels */
    int32_t key_depth;   /* Morton key resolution (levels available) */
    int64_t n_points;
    int64_t n_boxes;     /* all levels, fake boxes included */
    double root_min;     /* reference Tree.xmin (= ymin) */
    double root_width;   /* reference Tree.xmax - Tree.xmin */
} kf_tree_info;

typedef struct {
    int64_t p2p_pairs;        /* near-field kernel evaluations per right-hand side */
    int64_t p2m_evals;        /* upward kernel evaluations */
    int64_t l2p_evals;        /* downward kernel evaluations at the points */
    int64_t m2l_pairs;        /* V-list (target, source) box pairs */
    int64_t m2m_boxes;        /* boxes whose upward check potential comes from children */
    int64_t l2l_boxes;        /* boxes that receive a parent's local expansion */
    int64_t c2e_up_rows;      /* upward check-to-equivalent solves */
    int64_t c2e_down_rows;    /* downward check-to-equivalent solves */
    int64_t leaves;           /* real leaves with points */
    int32_t p;                /* equivalent points per circle */
    int32_t p_padded;         /* row stride of the expansion slabs */
} kf_plan_stats;

/* Launch structure of a plan's evaluation (results do not depend on it, except batch_form and
 * near_form, which change the order of the near-field (and P2M, L2P) sums: rounding only). */
typedef struct {
    int32_t near_grid;     /* workgroups of the near-field launch per right-hand-side chunk (each
                              loops over leaves); 0 = one per leaf */
    int32_t near_fork;     /* where an overlapped near field starts: 1 = after the upward pass
                              (overlapping the M2L and the downward pass), 2 = after the P2M
                              (overlapping the upward pass and what follows), 3 = with the P2M,
                              4 = no overlap (the near field on the evaluation's stream, each
                              pass alone on the device); 0 = the default (2) */
    int32_t m2l_pipeline;  /* 1: each level's M2L starts on a stream of its own as soon as the
                              upward pass has finished that level; 0: one M2L launch after it */
    int32_t level_rows;    /* rows (box x right-hand side) per workgroup of the per-level upward
                              and downward launches: 16 or 32 (column-split workgroups, one wave
                              per 16 expansion coefficients), 64 (whole rows per wave), 128 (two
                              64-row tiles per 8-wave workgroup sharing the operator staging;
                              the downward pass needs batches of 32 or more and takes 64
                              otherwise), 0 = by level size (16 below 24K rows, else 128 for
                              batches of 32 or more and 64 otherwise) */
    int32_t near_split;    /* per mille of the leaves (the tail of the leaf list) whose near field
                              waits until the M2L has finished and overlaps the downward pass;
                              -1 = none; 0 = the default: 100 for batches of up to 4 right-hand
                              sides unless the near field starts after the upward pass, none
                              otherwise */
    int32_t chain;         /* upward and downward passes: 1 = one launch per level (and class),
                              2 = one persistent launch per pass (work items of 16 rows from an
                              in-order queue, per-row completion flags); 0 = the default: 2 for
                              batches of up to 4 right-hand sides, 1 otherwise (plans on a
                              workspace and complex kernels: always 1) */
    int32_t batch_form;    /* near field, P2M and L2P of batches of 16, 32 or 64 right-hand sides
                              (kernel values not stored): 1 = vector ALU, each kernel value formed
                              once per 16 right-hand sides; 2 = FP64 MFMA, each value formed once
                              and contracted with all of them on the matrix cores; 0 = the
                              default (2).  The autotuner keeps the value it is given. */
    int32_t near_form;     /* near field of one right-hand side: 2 = each pair of mutual colleague
                              leaves evaluated once for both (leaves of at most 64 points, real
                              kernels, kernel values not stored), 1 = every directed pair on its
                              own; 0 = the default (2 for tabulated kernels, 1 otherwise).  The
                              autotuner keeps the value it is given. */
} kf_tuning;

/* Library version (KF_VERSION). */
int kf_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* kf_last_error(void);

/* Host FP64 K0(x) as used by the device kernels (test hook for misc/numba_special_functions.py:167-181). */
double kf_k0(double x);

/* Build the level-restricted quadtree with X-list fake children for n points.
 * Replaces tree.py:464-522 Tree.__init__ (ppl = cutoff).  d_x, d_y: device f64[n]. */
int kf_tree_build(const double* d_x, const double* d_y, int64_t n, int32_t cutoff,
                  void* stream, kf_tree** out);

/* Build n_trees independent trees at once: tree t's points are d_x/d_y[h_offsets[t] ..
 * h_offsets[t+1]) (h_offsets: host int64[n_trees + 1], h_offsets[0] = 0, every tree non-empty);
 * bounding boxes, keys, sort and gather are shared launches, the box structure of the trees is
 * built on host threads.  out: host array of n_trees handles, each the tree of tree.py:464-522
 * Tree(x_t, y_t, cutoff), destroyed individually.  A forest plan (kf_forest_create) over all of
 * them in batch order uses their points in place. */
int kf_tree_build_batch(const double* d_x, const double* d_y, const int64_t* h_offsets,
                        int32_t n_trees, int32_t cutoff, void* stream, kf_tree** out);

/* Tree summary (reference Tree.levels, Tree.xmin/xmax, Tree.N). */
int kf_tree_get_info(const kf_tree* tree, kf_tree_info* info);

/* Number of boxes at a level (reference Level.n_node, tree.py:345). */
int kf_tree_level_size(const kf_tree* tree, int32_t level, int64_t* n_node);

/* Host copy of one level in the reference Level layout (tree.py:313-420): box corners,
 * point ranges, parent/children indices within the neighbouring levels, leaf / fake_leaf /
 * Xlist flags and colleagues[n_node*9] (-1 padded).  Any output pointer may be NULL. */
int kf_tree_export_level(const kf_tree* tree, int32_t level, double* xmin, double* ymin,
                         int64_t* bot_ind, int64_t* top_ind, int64_t* parent_ind,
                         int64_t* children_ind, uint8_t* leaf, uint8_t* fake_leaf,
                         uint8_t* xlist, int64_t* colleagues);

/* Host copy of the tree-ordered points and the ordering vector (reference Tree.x, Tree.y,
 * Tree.ordv).  Any output pointer may be NULL. */
int kf_tree_export_points(const kf_tree* tree, int64_t* ordv, double* x, double* y);

/* Leaf of each of m points d_x, d_y (device f64[m]): its level and its position within the
 * level (reference tree.py:686-727 Tree.locate_point / locate_points: from the root, while the
 * box is not a leaf, go to child 2 (x > xmid) + (y > ymid); X-list leaves are leaves).
 * d_level, d_pos: device int64[m].  Stream-ordered.  Needs a tree built on the GPU. */
int kf_tree_locate(const kf_tree* tree, const double* d_x, const double* d_y, int64_t m,
                   int64_t* d_level, int64_t* d_pos, void* stream);

/* Depths of the boxes of a level that are not X-list fake children (reference Level.depths,
 * tree.py:380-383, 423-430, as the last level-restriction sweep leaves them: 0 for a leaf, else
 * 1 + the largest child depth), in level order; kf_tree_level_real gives their number (the
 * level's boxes before its fake children, which the reference's depth array does not cover). */
int kf_tree_export_depths(const kf_tree* tree, int32_t level, int64_t* depths);
int kf_tree_level_real(const kf_tree* tree, int32_t level, int64_t* n_real);

/* Device views of the tree-ordered points (valid for the tree's lifetime). */
int kf_tree_device_points(const kf_tree* tree, const double** d_x, const double** d_y,
                          const int32_t** d_ordv);

/* Box-level structure only (no device points), from caller-sorted Morton keys at depth
 * key_depth (x bit above y bit in each pair, keys[i] = interleave(ix_i, iy_i) with ix_i the
 * number of gridlines root_min + k*root_width*2^-key_depth strictly below x_i).  The
 * levels, colleagues, level restriction and X-list of tree.py:499-576 for a given point
 * ordering; an inspection hook for the box logic (exports work, plans do not). */
int kf_tree_from_sorted_keys(const uint64_t* keys, int64_t n, int32_t cutoff, double root_min,
                             double root_width, int32_t key_depth, kf_tree** out);

void kf_tree_destroy(kf_tree* tree);

/* Precompute operators (circles, C2E factors, M2M, the 40 M2L offsets per level) and the
 * per-level work lists, and allocate the expansion slabs in HBM.
 * Replaces class_fmm.py:182-298 FMM.precompute and fmm.py:249-347 (the operator part of
 * _on_the_fly_fmm) / fmm.py:530-712 fmm_planner.  The plan keeps a reference to `tree`;
 * destroy the plan before the tree.  Requires tree levels >= 2 and 4 <= p <= 64. */
int kf_plan_create(const kf_tree* tree, const kf_kernel_desc* kernel, int32_t p,
                   int32_t precision, void* stream, kf_plan** out);

/* One plan over n_trees independent trees, evaluated in the same launches: the upward and
 * downward passes of all trees share one launch per level (per distinct root width), the M2L,
 * P2M, P2P and L2P one launch each; small trees fill MFMA tiles together.  Densities and
 * potentials of a forest plan are the trees' vectors concatenated in tree order (tree i's n_i
 * values in its input order at offset n_0 + ... + n_{i-1}); kf_plan_eval / kf_plan_eval_ptrs /
 * kf_plan_expand + kf_plan_self work unchanged, kf_plan_targets needs a single-tree plan.  The
 * trees must outlive the plan.  The reference evaluates one tree per call
 * (fmm.py:741-814 planned_fmm, class_fmm.py:299-352); this is the batched form of the same. */
int kf_forest_create(const kf_tree* const* trees, int32_t n_trees, const kf_kernel_desc* kernel,
                     int32_t p, int32_t precision, void* stream, kf_plan** out);

int kf_plan_stats_get(const kf_plan* plan, kf_plan_stats* stats);

/* Inspection hook: rebuilds the plan's box trees and work lists on the host from the trees'
 * sorted keys (the host restatement of tree.py:464-576 and of the pass structure of
 * class_fmm.py:299-352) and compares them with the lists the plan built on the GPU, entry by
 * entry, and its operators with the host's (ops_host: M2L, M2M, L2L blocks and the Fourier
 * basis to a few units in the last place; the check-to-equivalent eigenvalues to 1e-13 of the
 * largest, the rounding of E's entries).  KF_OK if they agree, else KF_ERR_STATE naming the
 * first difference.  Synchronises the stream.  No reference counterpart (a self-check for
 * tests). */
int kf_plan_verify(const kf_plan* plan, void* stream);

/* Upward and downward passes for `nrhs` densities d_tau[r*ld_tau + i] (input order): fills
 * the local expansions of every box.  Replaces class_fmm.py:299-352 build_expansions. */
int kf_plan_expand(kf_plan* plan, const double* d_tau, int64_t ld_tau, int32_t nrhs,
                   void* stream);

/* Potentials at the source points, u_i = sum_{j != i} G(x_i, x_j) tau_j, in input order:
 * leaf L2P + colleague P2P with index self-exclusion.  Requires kf_plan_expand.
 * Replaces fmm.py:390-408 (downward passes 2 and 3 and the de-ordering). */
int kf_plan_self(kf_plan* plan, double* d_pot, int64_t ld_pot, void* stream);

/* kf_plan_expand + kf_plan_self.  Replaces fmm.py:741-814 planned_fmm and the pass part of
 * fmm.py:238-408 _on_the_fly_fmm. */
int kf_plan_eval(kf_plan* plan, const double* d_tau, int64_t ld_tau, int32_t nrhs,
                 double* d_pot, int64_t ld_pot, void* stream);

/* kf_plan_eval for nrhs densities at unrelated device addresses: d_taus[r] (n values, input
 * order) gives d_pots[r].  d_taus and d_pots are HOST arrays of device pointers; batches of up
 * to 64 travel in the kernel arguments, so the call copies nothing host-to-device.  A
 * density's potential is bitwise the same whatever else shares its batch.  Same computation
 * as fmm.py:741-814 planned_fmm on a (nrhs, n) tau; the pointer form serves the request loop
 * (requests own their buffers), which the reference does not have. */
int kf_plan_eval_ptrs(kf_plan* plan, const double* const* d_taus, double* const* d_pots,
                      int32_t nrhs, void* stream);

/* Read / set the launch structure (kf_tuning).  Setting discards captured graphs. */
int kf_plan_get_tuning(const kf_plan* plan, kf_tuning* tuning);
int kf_plan_set_tuning(kf_plan* plan, const kf_tuning* tuning);

/* Autotuner: times evaluations of batches of nrhs (1..64) right-hand sides under each candidate
 * launch structure (near-field grid of 0 / 4 / 8 / 16 workgroups per CU, near-field fork
 * before or after the upward pass, with the fork before it a near_split of -1 or 100 as well,
 * M2L pipelined or not, per-level launches or persistent chains), then, under the fastest, each
 * tile shape of the level launches (level_rows: the size rule, 16, 32 or 64 rows per
 * workgroup), on `stream`, and keeps the fastest.
 * Uses scratch densities (zeros) of its own; synchronises the stream.  chosen / best_ms may be
 * NULL.  No reference counterpart. */
int kf_plan_autotune(kf_plan* plan, int32_t nrhs, void* stream, kf_tuning* chosen, double* best_ms);

/* Graph replay on (enable != 0) or off: kf_plan_eval and kf_plan_eval_ptrs then capture the
 * whole evaluation of each batch width once (on a stream of the plan; the near field is a
 * parallel branch) and replay it with one hipGraphLaunch per batch, after setting the new
 * density / potential addresses in its gather and L2P nodes.  Results are bitwise those of the
 * launched evaluation.  A replay of the same width still running is waited for (host) before
 * its nodes change; growing the slabs (a wider batch than reserved) discards the graphs.  Not
 * used while profiling, nor while the plan is attached to a workspace (kf_plan_attach_workspace:
 * such plans evaluate by launches).  No reference counterpart (the reference launches nothing). */
int kf_plan_set_graphs(kf_plan* plan, int32_t enable);

/* Lays the plan out for batches of up to nrhs (1..64) right-hand sides now: tile tables for
 * every batch width up to it and slabs of that width.  Later evaluations of at most nrhs per
 * batch neither allocate device memory nor copy host-to-device.  Synchronises the stream.
 * The reference sizes its workspace once per tree (tree.py:411-420 allocate_workspace). */
int kf_plan_reserve(kf_plan* plan, int32_t nrhs, void* stream);

/* Device bytes the plan and its tree hold when laid out for batches of nrhs (1..64): the
 * admission figure of the request loop.  Never less than what the plan holds now.  A plan on
 * a workspace counts none of its expansion slabs (they are the workspace's).  No reference
 * counterpart. */
int kf_plan_footprint(const kf_plan* plan, int32_t nrhs, int64_t* bytes);

/* Workspaces: expansion slabs (check potentials, equivalent densities, densities and near
 * field in tree order) shared by plans that never evaluate at the same time, e.g. the trees
 * served through one stream slot.  A resident tree then holds only its tree, operators and work
 * lists; the slabs exist once per workspace, sized for the largest need of its plans (each
 * buffer grows on demand: kf_plan_reserve on an attached plan, or its first evaluation of a
 * wider batch).  Evaluations of plans on one workspace must be ordered (one stream, or
 * synchronised); an evaluation writes every slab element before it reads it, so nothing is
 * cleared between plans.  An expansion kept by
 * kf_plan_expand for kf_plan_self / kf_plan_targets is valid until another plan on the
 * workspace evaluates (those calls then fail with KF_ERR_STATE).  Plans on a workspace use the
 * per-level upward and downward launches (kf_tuning.chain is ignored).  Destroying a workspace
 * returns its plans to slabs of their own; in-flight evaluations must have finished.  No
 * reference counterpart (the reference allocates each tree's workspace, tree.py:411-420). */
int kf_workspace_create(kf_workspace** out);
int kf_workspace_destroy(kf_workspace* ws);
/* Device bytes the workspace holds now. */
int kf_workspace_bytes(const kf_workspace* ws, int64_t* bytes);
/* Fills every slab of the workspace with `value` on `stream` (inspection hook: with NaN, any
 * read of a slab element that an evaluation has not written first shows in its results). */
int kf_workspace_fill(kf_workspace* ws, double value, void* stream);
/* Moves the plan's expansion slabs to `ws` (NULL: back to slabs of its own, allocated on next
 * use).  Frees the plan's own slabs and discards its captured graphs. */
int kf_plan_attach_workspace(kf_plan* plan, kf_workspace* ws);
/* Bytes of expansion slabs the plan needs for batches of nrhs (1..64) right-hand sides: what a
 * workspace must hold for it. */
int kf_plan_slab_bytes(const kf_plan* plan, int32_t nrhs, int64_t* bytes);

/* Potentials at m arbitrary targets inside the root box (locate + L2P + P2P); check_self
 * skips sources that coincide exactly with a target.  Every right-hand side r of the last
 * kf_plan_expand goes to d_pot[r*m + i] (nrhs*m values).
 * Replaces class_fmm.py:379-399 evaluate_to_points (and tree.py:703-727 locate_points). */
int kf_plan_targets(kf_plan* plan, const double* d_tx, const double* d_ty, int64_t m,
                    int32_t check_self, double* d_pot, void* stream);

/* Per-pass device time.  Writes up to 8 values: the summed milliseconds of the passes
 * [gather, p2m, upward, downward, l2p, p2p] over the evaluations since the previous call, the
 * number of evaluations, and the summed milliseconds of the M2L launch inside downward (0 with
 * the pipelined M2L); then resets the totals and turns event recording on or off (enable).
 * p2p (the near field; both shares of a split near field, each timed on its own) runs on the
 * plan's second stream, overlapped with p2m/upward/downward in kf_plan_eval; l2p includes the
 * wait for it.  Events are recorded without synchronising; reading synchronises on them.
 * Returns the number of values written. */
int kf_plan_profile(kf_plan* plan, int32_t enable, double* ms, int32_t cap);

void kf_plan_destroy(kf_plan* plan);

/* Direct sums.  self_mode = 1: out_j = sum_{i != tidx_j} G(s_i, t_j) tau_i with tidx the
 * source index of target j (d_tidx NULL means tidx_j = j); self_mode = 2: skip sources that
 * coincide exactly with the target (class_fmm.py:107); self_mode = 0: every source.
 * Compensated FP64 summation.  Replaces fmm.py:128-147 KA / KAS and fmm.py:170-188
 * Kernel_Apply. */
int kf_direct(const kf_kernel_desc* kernel, const double* d_sx, const double* d_sy,
              const double* d_tau, int64_t ns, const double* d_tx, const double* d_ty,
              const int64_t* d_tidx, int64_t nt, int32_t self_mode, double* d_out, void* stream);

/* kf_direct without allocating: the same kernels and results (bitwise), with the partial sums
 * in a scratch array the library holds on each GPU it runs on (KF_DIRECT_NOALLOC_MAX_DOUBLES
 * doubles, 32 MiB, from the first call on a device until the process ends).  The call makes no
 * hipMalloc / hipFree and does not synchronise: it is stream-ordered like a kernel launch.
 * Requires 2 * nt * ceil(ns / 16384) <= KF_DIRECT_NOALLOC_MAX_DOUBLES (KF_ERR_ARG otherwise)
 * and, for a KF_TABLE kernel, a table whose device copy exists (made by its first kf_direct or
 * kf_kernel_form call; KF_ERR_STATE otherwise).  Every call on a GPU uses that GPU's one
 * scratch array: calls whose kernels overlap on one GPU (different streams, or other host
 * threads) race and must be ordered by the caller.
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation). */
#define KF_DIRECT_NOALLOC_MAX_DOUBLES 4194304
int kf_direct_noalloc(const kf_kernel_desc* kernel, const double* d_sx, const double* d_sy,
                      const double* d_tau, int64_t ns, const double* d_tx, const double* d_ty,
                      const int64_t* d_tidx, int64_t nt, int32_t self_mode, double* d_out,
                      void* stream);

/* Stores the plan's kernel matrices, as the reference's fmm_planner does (fmm.py:646-711: the
 * near field, the upward P2M and the downward L2P leaf interactions, in CSR there; applied at
 * fmm.py:745-770): dense per-leaf blocks here, 8 bytes per value -- kf_plan_stats.p2p_pairs plus
 * one per point for the near field, p2m_evals and l2p_evals.  Evaluations then read the matrices
 * instead of evaluating the kernel (worth it for costly kernels: KF_TABLE, KF_MODHELM).  P2M and
 * L2P read bitwise the values they would compute; the near field stores G with its prefactor,
 * so potentials agree with the unstored plan to rounding.  A density keeps bitwise the same
 * potential at any batch width.  on = 0 frees them.  Synchronises the stream. */
int kf_plan_store_matrices(kf_plan* plan, int32_t on, void* stream);

/* A user's radial kernel G(r^2) for KF_TABLE (replaces the numba Kernel_Eval a user hands to
 * fmm.py:119 Get_Kernel_Functions; radial, as fmm.py:373-374 already assumes): piecewise
 * polynomials in u = log2(r^2) on nseg segments of width du from u0: y = (u - u0) / du,
 * s = floor(y), x = 2 (y - s) - 1, G = sum_k coef[s * (deg + 1) + k] x^k; u below u0 or past
 * u0 + nseg du takes the end value of the first or last segment; G(0) = g0.  du is a power of two
 * in [2^-10, 16] and u0 a multiple of it (so y and x are exact); deg <= 24.  A plan needs the
 * table to reach u = log2((2 W)^2), W its root box width.  coef is a host array, copied.  Creation makes no HIP call; kf_direct and kf_kernel_form upload
 * a device copy on their first use of the table, freed by kf_ktable_destroy, so destroy a table
 * only after those calls' kernels have finished (kf_kernel_form does not wait for its own).  A
 * plan keeps its own copy and needs the table only while kf_plan_create / kf_forest_create run. */
int kf_ktable_create(const double* coef, int32_t nseg, int32_t deg, double u0, double du, double g0,
                     kf_ktable** out);
void kf_ktable_destroy(kf_ktable* table);
/* G at n squared distances on the host, as the plan's operator construction evaluates it. */
int kf_ktable_eval(const kf_ktable* table, const double* r2, int64_t n, double* out);

/* Dense kernel matrix out[j*ns + i] = G(s_i, t_j).  Replaces fmm.py:119-127 KF. */
int kf_kernel_form(const kf_kernel_desc* kernel, const double* d_sx, const double* d_sy,
                   int64_t ns, const double* d_tx, const double* d_ty, int64_t nt,
                   double* d_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* KIFMM2D_H */
