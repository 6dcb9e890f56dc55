"""CPU restatement of the reference KIFMM drivers.  TEST INFRASTRUCTURE ONLY.

* ``FMMRef`` follows class_fmm.py:166-402 (precompute, build_expansions,
  evaluate_to_points) and adds ``self_potential`` = fmm.py:392-408 (leaf L2P,
  colleague P2P with the index self-exclusion, de-ordering).
* ``on_the_fly_fmm`` follows fmm.py:190-236: direct self-sum when the tree has at most
  two levels, otherwise the passes above.  It uses class_fmm's density-row maps, i.e. the
  class_fmm-correct result where fmm.py's box-indexed variant would break (source-free
  sibling groups).
* ``direct`` is the FP64 direct-sum oracle (fmm.py:128-147 KA/KAS) with compensated
  summation.

Deliberate difference: ``Check_Us`` is zeroed at the start of every build_expansions
(class_fmm.py zeroes it only at allocation, so a second call double-counts P2M).
"""
import numpy as np
import scipy.linalg

from ._clib import lib, ptr, kargs, as_u8, as_i64, as_f64, I64, F64, INT
from .tree_ref import Tree

LAPLACE = {"kind": 0}


def modified_helmholtz(k, scale=1.0 / (2.0 * np.pi)):
    return {"kind": 1, "k": float(k), "scale": float(scale)}


def user_kernel(g):
    """A user's radial kernel G = g(r^2) (g vectorised over numpy arrays), evaluated exactly
    as given -- the reference calls its Kernel_Eval for every pair (fmm.py:119-147).  Only the
    numpy paths of the oracle take it (``kernel_eval_np``, ``direct_np``, and ``FMMRef``'s
    numpy passes)."""
    return {"kind": 2, "g": g}


def complex_kernel(g):
    """A user's complex-valued radial kernel G = g(r^2) (g vectorised, complex results), as
    class_fmm.FMM(..., iscomplex=True) takes one (class_fmm.py:167, 174); numpy paths only."""
    return {"kind": 3, "g": g}


def hankel_kernel(k):
    """examples/hankel.py:31-39: G(r) = 0.25j H0^(1)(k r), with scipy's hankel1 in place of the
    example's Chebyshev interpolant of it."""
    import scipy.special

    def g(r2):
        return 0.25j * scipy.special.hankel1(0, k * np.sqrt(r2))
    return complex_kernel(g)


def is_numpy_kernel(kernel):
    return kernel["kind"] in (2, 3)


def kernel_eval_np(kernel, sx, sy, tx, ty):
    """Vectorised G(s, t) (numpy broadcasting); Laplace, scale*K0(k r) or a user's g(r^2)."""
    dx = tx - sx
    dy = ty - sy
    if kernel["kind"] == 0:
        return (-0.25 / np.pi) * np.log(dx * dx + dy * dy)
    if kernel["kind"] in (2, 3):
        return kernel["g"](dx * dx + dy * dy)
    import scipy.special
    return kernel["scale"] * scipy.special.k0(kernel["k"] * np.sqrt(dx * dx + dy * dy))


def level_information(width, theta):
    """fmm.py:12-22 / class_fmm.py:35-45: equivalent/check circles of a level."""
    dd = 0.01
    r1 = 0.5 * width * (np.sqrt(2) + dd)
    r2 = 0.5 * width * (4 - np.sqrt(2) - 2 * dd)
    return r1 * np.cos(theta), r1 * np.sin(theta), r2 * np.cos(theta), r2 * np.sin(theta), r1, r2


def kernel_form(kernel, sx, sy, tx, ty):
    """fmm.py:119-127 KF: (nt, ns) matrix out[j, i] = G(s_i, t_j)."""
    sx, sy, tx, ty = map(as_f64, (sx, sy, tx, ty))
    if is_numpy_kernel(kernel):
        return kernel_eval_np(kernel, sx[None, :], sy[None, :], tx[:, None], ty[:, None])
    out = np.empty((tx.size, sx.size))
    lib().orc_kernel_form(*kargs(kernel), ptr(sx), ptr(sy), I64(sx.size), ptr(tx), ptr(ty),
                          I64(tx.size), ptr(out))
    return out


def direct(kernel, sx, sy, tau, tx=None, ty=None, target_index=None):
    """FP64 direct sum.  Self mode (tx is None) excludes i == j like KAS (fmm.py:138-147);
    ``target_index`` gives the source index of each target for a sampled self sum."""
    sx, sy, tau = map(as_f64, (sx, sy, tau))
    self_mode = tx is None or target_index is not None
    if tx is None:
        tx, ty = sx, sy
    tx, ty = as_f64(tx), as_f64(ty)
    tid = as_i64(target_index) if target_index is not None else None
    out = np.empty(tx.size)
    lib().orc_direct(*kargs(kernel), ptr(sx), ptr(sy), ptr(tau), I64(sx.size), ptr(tx), ptr(ty),
                     I64(tx.size), ptr(tid) if tid is not None else None, INT(1 if self_mode else 0),
                     ptr(out))
    return out


def direct_np(kernel, sx, sy, tau, tx=None, ty=None):
    """Direct sum through ``kernel_eval_np`` in blocks of targets (fmm.py:128-147 KA, or KAS
    with its i != j exclusion when tx is None); for kernels the C loops do not take."""
    sx, sy, tau = map(as_f64, (sx, sy, tau))
    self_mode = tx is None
    if self_mode:
        tx, ty = sx, sy
    tx, ty = as_f64(tx), as_f64(ty)
    out = np.empty(tx.size, dtype=complex if kernel["kind"] == 3 else float)
    for a in range(0, tx.size, 256):
        b = min(a + 256, tx.size)
        with np.errstate(divide="ignore", invalid="ignore"):
            G = kernel_eval_np(kernel, sx[None, :], sy[None, :], tx[a:b, None], ty[a:b, None])
        if self_mode:
            G[np.arange(b - a), np.arange(a, b)] = 0.0
        out[a:b] = G @ tau
    return out


class FMMRef:
    """class_fmm.py:166-402."""

    def __init__(self, x, y, kernel, Nequiv=48, Ncutoff=50):
        self.x = as_f64(x)
        self.y = as_f64(y)
        self.kernel = kernel
        self.Nequiv = Nequiv
        self.Ncutoff = Ncutoff
        # user kernels (real or complex) run numpy restatements of the C loops below
        self.np_passes = is_numpy_kernel(kernel)
        self.dtype = complex if kernel["kind"] == 3 else float

    def build_tree(self):
        self.tree = Tree(self.x, self.y, self.Ncutoff)

    def precompute(self):  # class_fmm.py:182-298
        if not hasattr(self, "tree"):
            self.build_tree()
        tree, p, K = self.tree, self.Nequiv, self.kernel
        if not tree.workspace_allocated:
            tree.allocate_workspace(p, self.dtype)
        theta = np.linspace(0, 2 * np.pi, p, endpoint=False)
        self.small_xs, self.small_ys, self.large_xs, self.large_ys, self.widths = [], [], [], [], []
        for lv in tree.Levels:
            sxx, syy, lxx, lyy, _, _ = level_information(lv.width, theta)
            self.small_xs.append(sxx)
            self.small_ys.append(syy)
            self.large_xs.append(lxx)
            self.large_ys.append(lyy)
            self.widths.append(lv.width)
        # C2E LU per level (212-217)
        self.E = [kernel_form(K, self.small_xs[i], self.small_ys[i], self.large_xs[i], self.large_ys[i])
                  for i in range(tree.levels)]
        self.E2C_LUs = [scipy.linalg.lu_factor(e) for e in self.E]
        # M2M operator (218-235): children's small circles -> parent's large circle
        self.M2MC = []
        for i in range(tree.levels - 1):
            h = 0.5 * self.widths[i + 1]
            sx, sy = self.small_xs[i + 1], self.small_ys[i + 1]
            cx = np.concatenate([sx - h, sx - h, sx + h, sx + h])
            cy = np.concatenate([sy - h, sy + h, sy - h, sy + h])
            self.M2MC.append(kernel_form(K, cx, cy, self.large_xs[i], self.large_ys[i]))
        # 40 M2L blocks per level (236-250) and the collected 4p x 4p forms (251-285)
        self.M2LS = [None]
        self.CM2LS = [None]
        for i in range(1, tree.levels):
            w = self.widths[i]
            blocks = {}
            for ix in range(-3, 4):
                for iy in range(-3, 4):
                    if abs(ix) <= 1 and abs(iy) <= 1:
                        continue
                    blocks[(ix, iy)] = kernel_form(K, self.small_xs[i] + ix * w, self.small_ys[i] + iy * w,
                                                   self.small_xs[i], self.small_ys[i])
            self.M2LS.append(blocks)
            coll = np.zeros((3, 3, 4 * p, 4 * p), dtype=self.dtype)
            for kx in range(3):
                for ky in range(3):
                    if kx == 1 and ky == 1:
                        continue
                    for a in range(2):
                        for b in range(2):
                            for c in range(2):
                                for d in range(2):
                                    ox = 2 * (kx - 1) - a + c
                                    oy = 2 * (ky - 1) - b + d
                                    blk = 0.0 if (abs(ox) <= 1 and abs(oy) <= 1) else blocks[(ox, oy)]
                                    coll[kx, ky, (2 * a + b) * p:(2 * a + b + 1) * p,
                                         (2 * c + d) * p:(2 * c + d + 1) * p] = blk
            self.CM2LS.append(coll)

    def build_expansions(self, tau):  # class_fmm.py:299-352
        if self.np_passes:
            return self._build_expansions_np(tau)
        tree, p, K = self.tree, self.Nequiv, self.kernel
        L = lib()
        tau_o = as_f64(np.asarray(tau, dtype=np.float64)[tree.ordv])
        self.tau_ordered = tau_o
        # upward pass (309-319)
        for ind in reversed(range(1, tree.levels)):
            lv = tree.Levels[ind]
            lv.Check_Us[:] = 0.0
            if ind != tree.levels - 1:
                anc = tree.Levels[ind + 1]
                temp = np.ascontiguousarray(self.M2MC[ind].dot(anc.RSEQD.T).T)
                L.orc_distribute(ptr(lv.Check_Us), ptr(temp), ptr(as_i64(anc.short_parent_ind)),
                                 ptr(anc.parent_density_ind), ptr(lv.this_density_ind),
                                 I64(anc.short_parent_ind.shape[0]), I64(p))
            cu = as_u8(lv.compute_upwards)
            L.orc_upwards_pass(*kargs(K), ptr(tree.x), ptr(tree.y), ptr(lv.bot_ind), ptr(lv.top_ind),
                               ptr(cu), ptr(as_f64(self.large_xs[ind])), ptr(as_f64(self.large_ys[ind])),
                               I64(p), ptr(lv.xmid), ptr(lv.ymid), ptr(tau_o), ptr(lv.Check_Us),
                               ptr(lv.this_density_ind), I64(lv.n_node))
            if lv.n_allocate_density:
                lv.Equiv_Densities[:] = scipy.linalg.lu_solve(self.E2C_LUs[ind], lv.Check_Us.T).T
        # downward pass 1 (324-343)
        for ind in range(1, tree.levels):
            tree.Levels[ind].Local_Solutions[:] = 0.0
        for ind in range(1, tree.levels - 1):
            lv, desc = tree.Levels[ind], tree.Levels[ind + 1]
            doit = np.logical_and(np.logical_or(lv.not_leaf, lv.Xlist), np.logical_not(lv.fake_leaf))
            loc_exp = scipy.linalg.lu_solve(self.E2C_LUs[ind], lv.Local_Solutions[doit].T).T
            loc_sol = self.M2MC[ind].T.dot(loc_exp.T).T
            loc_sol = loc_sol[np.argsort(lv.children_ind[doit], kind="stable")]
            desc.Local_Solutions[:] = loc_sol.reshape(desc.Local_Solutions.shape)
            nsrc = desc.n_sib_has_source
            m2ms = np.zeros((3, 3, 4 * p, nsrc))
            for kx in range(3):
                for ky in range(3):
                    if kx == 1 and ky == 1:
                        continue
                    m2ms[kx, ky] = self.CM2LS[ind + 1][kx, ky].dot(desc.RSEQD.T)
            ci4 = (lv.children_ind / 4).astype(np.int64)  # -1 -> 0 (trunc), as class_fmm.py:342
            L.orc_add_interactions(ptr(as_u8(doit)), ptr(ci4), ptr(lv.colleagues), ptr(lv.xmid),
                                   ptr(lv.ymid), ptr(desc.Local_Solutions), ptr(m2ms), I64(p),
                                   ptr(desc.parent_density_ind), I64(nsrc), I64(lv.n_node))
        # downward pass 2a (348-350)
        for ind in range(1, tree.levels):
            lv = tree.Levels[ind]
            lv.Local_Expansions = scipy.linalg.lu_solve(self.E2C_LUs[ind], lv.Local_Solutions.T).T

    # ---- numpy restatements of the passes (ref_loops.c), for user kernels real or complex ----
    def _build_expansions_np(self, tau):
        tree, p, K = self.tree, self.Nequiv, self.kernel
        tau_o = np.asarray(tau)[tree.ordv]
        self.tau_ordered = tau_o
        for ind in reversed(range(1, tree.levels)):  # upward pass (class_fmm.py:309-319)
            lv = tree.Levels[ind]
            lv.Check_Us[:] = 0.0
            if ind != tree.levels - 1:
                anc = tree.Levels[ind + 1]
                temp = self.M2MC[ind].dot(anc.RSEQD.T).T
                # numba_distribute (class_fmm.py:12-16)
                spi, pdi = anc.short_parent_ind, anc.parent_density_ind
                for g in range(spi.shape[0]):
                    if pdi[g] >= 0:
                        lv.Check_Us[lv.this_density_ind[spi[g]]] = temp[pdi[g]]
            # numba_upwards_pass (class_fmm.py:58-71): P2M into the large circle of every box
            # that computes upward
            lx, ly = self.large_xs[ind], self.large_ys[ind]
            for i in range(lv.n_node):
                r = lv.this_density_ind[i]
                if not lv.compute_upwards[i] or r < 0:
                    continue
                b, t = lv.bot_ind[i], lv.top_ind[i]
                G = kernel_eval_np(K, tree.x[None, b:t], tree.y[None, b:t], (lx + lv.xmid[i])[:, None],
                                   (ly + lv.ymid[i])[:, None])
                lv.Check_Us[r] += G @ tau_o[b:t]
            if lv.n_allocate_density:
                lv.Equiv_Densities[:] = scipy.linalg.lu_solve(self.E2C_LUs[ind], lv.Check_Us.T).T
        for ind in range(1, tree.levels):  # downward pass 1 (class_fmm.py:324-343)
            tree.Levels[ind].Local_Solutions[:] = 0.0
        for ind in range(1, tree.levels - 1):
            lv, desc = tree.Levels[ind], tree.Levels[ind + 1]
            doit = np.logical_and(np.logical_or(lv.not_leaf, lv.Xlist), np.logical_not(lv.fake_leaf))
            loc_exp = scipy.linalg.lu_solve(self.E2C_LUs[ind], lv.Local_Solutions[doit].T).T
            loc_sol = self.M2MC[ind].T.dot(loc_exp.T).T
            loc_sol = loc_sol[np.argsort(lv.children_ind[doit], kind="stable")]
            desc.Local_Solutions[:] = loc_sol.reshape(desc.Local_Solutions.shape)
            nsrc = desc.n_sib_has_source
            m2ms = np.zeros((3, 3, 4 * p, nsrc), dtype=self.dtype)
            for kx in range(3):
                for ky in range(3):
                    if kx == 1 and ky == 1:
                        continue
                    m2ms[kx, ky] = self.CM2LS[ind + 1][kx, ky].dot(desc.RSEQD.T)
            ci4 = (lv.children_ind / 4).astype(np.int64)  # -1 -> 0 (trunc), as class_fmm.py:342
            # numba_add_interactions (class_fmm.py:18-33)
            for i in range(lv.n_node):
                if not doit[i]:
                    continue
                dii = ci4[i]
                for j in range(9):
                    ci = lv.colleagues[i, j]
                    if ci < 0 or ci == i:
                        continue
                    ddx, ddy = lv.xmid[ci] - lv.xmid[i], lv.ymid[ci] - lv.ymid[i]
                    xd, yd = int(np.sign(ddx)), int(np.sign(ddy))
                    di = desc.parent_density_ind[ci4[ci]]
                    if di < 0:
                        continue
                    blk = m2ms[xd + 1, yd + 1, :, di]
                    desc.Local_Solutions[4 * dii:4 * dii + 4] += blk.reshape(4, p)
        for ind in range(1, tree.levels):  # downward pass 2a (class_fmm.py:348-350)
            lv = tree.Levels[ind]
            lv.Local_Expansions = scipy.linalg.lu_solve(self.E2C_LUs[ind], lv.Local_Solutions.T).T

    def _self_potential_np(self):
        tree, K = self.tree, self.kernel
        sol = np.zeros(tree.N, dtype=self.dtype)
        x, y, tau = tree.x, tree.y, self.tau_ordered
        for ind in range(1, tree.levels):
            lv = tree.Levels[ind]
            for i in range(lv.n_node):
                if not lv.leaf[i] or lv.top_ind[i] <= lv.bot_ind[i]:
                    continue
                b, t = lv.bot_ind[i], lv.top_ind[i]
                # numba_downwards_pass2 (fmm.py:91-115): L2P from the large circle
                G = kernel_eval_np(K, self.large_xs[ind][None, :], self.large_ys[ind][None, :],
                                   (x[b:t] - lv.xmid[i])[:, None], (y[b:t] - lv.ymid[i])[:, None])
                sol[b:t] += G @ lv.Local_Expansions[i]
                # neighbour P2P with the index self-exclusion (fmm.py:30-64)
                acc = np.zeros(t - b, dtype=self.dtype)
                for j in range(9):
                    ci = lv.colleagues[i, j]
                    if ci < 0:
                        continue
                    b2, t2 = lv.bot_ind[ci], lv.top_ind[ci]
                    with np.errstate(divide="ignore", invalid="ignore"):
                        G = kernel_eval_np(K, x[None, b2:t2], y[None, b2:t2], x[b:t, None], y[b:t, None])
                    if ci == i:
                        G[np.arange(t - b), np.arange(t - b)] = 0.0
                    acc += G @ tau[b2:t2]
                sol[b:t] += acc
        out = np.empty(tree.N, dtype=self.dtype)
        out[tree.ordv] = sol
        return out

    def _evaluate_to_points_np(self, x, y, check_self):
        tree, K = self.tree, self.kernel
        inds, locs = tree.locate_points(x, y)
        pot = np.zeros(x.size, dtype=self.dtype)
        for n in range(x.size):
            lv = tree.Levels[inds[n]]
            g = locs[n]
            # class_fmm.py:73-113 (L2P from the leaf's local expansion, then its colleagues)
            G = kernel_eval_np(K, self.large_xs[inds[n]], self.large_ys[inds[n]], x[n] - lv.xmid[g],
                               y[n] - lv.ymid[g])
            acc = G @ lv.Local_Expansions[g]
            for j in range(9):
                ci = lv.colleagues[g, j]
                if ci < 0:
                    continue
                b2, t2 = lv.bot_ind[ci], lv.top_ind[ci]
                sx, sy = tree.x[b2:t2], tree.y[b2:t2]
                keep = np.ones(t2 - b2, dtype=bool)
                if check_self:
                    keep = np.logical_not(np.logical_and(x[n] - sx == 0.0, y[n] - sy == 0.0))
                with np.errstate(divide="ignore", invalid="ignore"):
                    G = kernel_eval_np(K, sx[keep], sy[keep], x[n], y[n])
                acc += G @ self.tau_ordered[b2:t2][keep]
            pot[n] = acc
        return pot

    def self_potential(self):
        """fmm.py:390-408: L2P into every real leaf, colleague P2P with index exclusion,
        then de-order.  Requires build_expansions."""
        if self.np_passes:
            return self._self_potential_np()
        tree, p, K = self.tree, self.Nequiv, self.kernel
        L = lib()
        sol = np.zeros(tree.N)
        for ind in range(1, tree.levels):
            lv = tree.Levels[ind]
            le = as_f64(lv.Local_Expansions)
            L.orc_downwards_pass2(*kargs(K), ptr(tree.x), ptr(tree.y), ptr(lv.bot_ind), ptr(lv.top_ind),
                                  ptr(as_u8(lv.leaf)), ptr(as_f64(self.large_xs[ind])),
                                  ptr(as_f64(self.large_ys[ind])), I64(p), ptr(lv.xmid), ptr(lv.ymid),
                                  ptr(le), ptr(sol), I64(lv.n_node))
        for ind in range(1, tree.levels):
            lv = tree.Levels[ind]
            L.orc_neighbor_interactions(*kargs(K), ptr(tree.x), ptr(tree.y), ptr(as_u8(lv.leaf)),
                                        ptr(lv.bot_ind), ptr(lv.top_ind), ptr(self.tau_ordered),
                                        ptr(lv.colleagues), ptr(sol), I64(lv.n_node))
        out = np.empty(tree.N)
        out[tree.ordv] = sol
        return out

    def evaluate_to_points(self, x, y, check_self=False):  # class_fmm.py:379-399
        tree, p, K = self.tree, self.Nequiv, self.kernel
        L = lib()
        x = as_f64(np.ravel(x))
        y = as_f64(np.ravel(y))
        if self.np_passes:
            return self._evaluate_to_points_np(x, y, check_self)
        inds, locs = tree.locate_points(x, y)
        f = tree.flat()
        exp = np.zeros((f["off"][-1], p))
        for ind in range(1, tree.levels):
            exp[f["off"][ind]:f["off"][ind + 1]] = tree.Levels[ind].Local_Expansions
        lx = as_f64(np.stack(self.large_xs))
        ly = as_f64(np.stack(self.large_ys))
        pot = np.zeros(x.size)
        L.orc_target_l2p(*kargs(K), ptr(x), ptr(y), I64(x.size), ptr(inds), ptr(locs), ptr(lx), ptr(ly),
                         I64(p), ptr(f["xmid"]), ptr(f["ymid"]), ptr(f["off"]), ptr(exp), ptr(pot))
        L.orc_target_p2p(*kargs(K), ptr(x), ptr(y), I64(x.size), ptr(tree.x), ptr(tree.y), ptr(inds),
                         ptr(locs), ptr(f["bot"]), ptr(f["top"]), ptr(f["coll"]), ptr(f["off"]),
                         ptr(self.tau_ordered), ptr(pot), INT(1 if check_self else 0))
        return pot


def on_the_fly_fmm(x, y, tau, Nequiv, Ncutoff, kernel):
    """fmm.py:190-236 -> (potential, tree)."""
    tree = Tree(x, y, Ncutoff)
    if tree.levels <= 2:
        return (direct_np(kernel, x, y, tau) if is_numpy_kernel(kernel) else direct(kernel, x, y, tau)), tree
    f = FMMRef(x, y, kernel, Nequiv, Ncutoff)
    f.tree = tree
    f.precompute()
    f.build_expansions(tau)
    return f.self_potential(), tree
