"""CPU restatement of the reference quadtree (pykifmm2d/tree.py).  TEST INFRASTRUCTURE ONLY.

Follows tree.py step by step: root box rule (464-498, including the ``ymax = y.min()``
quirk at 478), level-by-level splitting with the in-place cycle partition (128-190,
229-258), colleague tagging (260-294, 349-375), depths (423-430), the level-restriction
fixed point (533-545 with 432-444 and 296-311), the X-list fake children (546-557 with
446-458) and the per-level flags / density-row maps (558-576).  The numba loops are in
oracle/ref_loops.c.
"""
import numpy as np

from ._clib import lib, ptr, as_u8, as_i64, as_f64, I64, F64, INT

TREE_SEARCH_CROSSOVER = 100  # tree.py:12


class Level:
    """tree.py:313-420."""

    def __init__(self, xmin, ymin, width, bot_ind, top_ind, parent_ind):
        self.xmin = as_f64(xmin)
        self.ymin = as_f64(ymin)
        self.width = width
        self.half_width = 0.5 * width
        self.bot_ind = as_i64(bot_ind)
        self.top_ind = as_i64(top_ind)
        self.parent_ind = as_i64(parent_ind)
        n = self.xmin.shape[0]
        self.leaf = np.ones(n, dtype=bool)
        self.fake_leaf = np.zeros(n, dtype=bool)
        self.children_ind = -np.ones(n, dtype=np.int64)
        self._derive()

    def _derive(self):  # tree.py:339-346
        self.ns = self.top_ind - self.bot_ind
        self.xmid = self.xmin + self.half_width
        self.ymid = self.ymin + self.half_width
        self.xmax = self.xmin + self.width
        self.ymax = self.ymin + self.width
        self.n_node = self.xmin.shape[0]
        self.short_parent_ind = self.parent_ind[::4]

    def tag_colleagues(self, ancestor=None):  # tree.py:349-375
        self.colleagues = -np.ones((self.n_node, 9), dtype=np.int64)
        dist = 1.5 * self.width
        L = lib()
        if self.n_node < TREE_SEARCH_CROSSOVER or ancestor is None:
            # the cKDTree branch (tree.py:362-368) is only reachable for a level without
            # an ancestor, i.e. the one-box root level, where it coincides with this one.
            L.orc_tag_colleagues_direct(ptr(self.xmid), ptr(self.ymid), I64(self.n_node),
                                        F64(dist), ptr(self.colleagues))
        else:
            L.orc_loop_colleagues(ptr(self.xmid), ptr(self.ymid), I64(self.n_node), F64(dist),
                                  ptr(self.parent_ind), ptr(ancestor.colleagues),
                                  ptr(ancestor.children_ind), ptr(self.colleagues))

    def get_depths(self, descendant):  # tree.py:380-383
        self.depths = np.zeros(self.n_node, dtype=np.int64)
        if descendant is not None:
            leaf = as_u8(self.leaf)
            lib().orc_get_depths(ptr(self.depths), ptr(leaf), ptr(self.children_ind),
                                 ptr(descendant.depths), I64(self.n_node))

    def add_nodes(self, xmin, ymin, bot_ind, top_ind, parent_ind, xlist):  # tree.py:384-395
        m = xmin.shape[0]
        self.xmin = np.concatenate([self.xmin, xmin])
        self.ymin = np.concatenate([self.ymin, ymin])
        self.bot_ind = np.concatenate([self.bot_ind, bot_ind])
        self.top_ind = np.concatenate([self.top_ind, top_ind])
        self.parent_ind = np.concatenate([self.parent_ind, parent_ind])
        self.leaf = np.concatenate([self.leaf, np.full(m, not xlist)])
        self.fake_leaf = np.concatenate([self.fake_leaf, np.full(m, bool(xlist))])
        self.children_ind = np.concatenate([self.children_ind, -np.ones(m, dtype=np.int64)])
        self._derive()

    def get_Xlist(self):  # tree.py:396-398
        self.Xlist = np.zeros(self.n_node, dtype=bool)
        out = np.zeros(self.n_node, dtype=np.uint8)
        leaf = as_u8(self.leaf)
        lib().orc_get_xlist(ptr(self.depths), I64(self.depths.shape[0]), ptr(self.colleagues),
                            ptr(leaf), ptr(out))
        self.Xlist = out.astype(bool)

    def add_null_Xlist(self):
        self.Xlist = np.zeros(self.n_node, dtype=bool)

    def allocate_workspace(self, p, dtype=float):  # tree.py:411-420 (the definition that wins)
        self.Local_Solutions = np.zeros((self.n_node, p), dtype=dtype)
        self.Check_Us = np.zeros((self.n_allocate_density, p), dtype=dtype)
        self.Equiv_Densities = np.zeros((self.n_allocate_density, p), dtype=dtype)
        self.RSEQD = self.Equiv_Densities.reshape(self.n_allocate_density // 4, 4 * p)


def _divide(level, x, y, ordv, tosplit, children_start, xlist):
    """tree.py:128-190 via the C loop; returns the new children's arrays."""
    tosplit = as_u8(tosplit)
    m = 4 * int(tosplit.sum())
    xmin = np.empty(m)
    ymin = np.empty(m)
    bot = np.empty(m, dtype=np.int64)
    top = np.empty(m, dtype=np.int64)
    par = np.empty(m, dtype=np.int64)
    leaf = as_u8(level.leaf)
    rc = lib().orc_divide_and_reorder(
        ptr(x), ptr(y), ptr(ordv), ptr(tosplit), F64(level.half_width), ptr(level.xmid),
        ptr(level.ymid), ptr(level.bot_ind), ptr(level.top_ind), ptr(leaf), ptr(xmin), ptr(ymin),
        ptr(bot), ptr(top), ptr(par), ptr(level.children_ind), I64(children_start),
        INT(1 if xlist else 0), I64(level.n_node))
    if rc != 0:
        raise RuntimeError("oracle partition failed (%d)" % rc)
    level.leaf = leaf.astype(bool)
    return xmin, ymin, bot, top, par


class Tree:
    """tree.py:460-589, 703-710."""

    def __init__(self, x, y, ppl):
        self.x = as_f64(x).copy()
        self.y = as_f64(y).copy()
        self.points_per_leaf = ppl
        xmin, xmax = self.x.min(), self.x.max()
        ymin = self.y.min()
        ymax = self.y.min()  # tree.py:478 uses y.min() for ymax (kept as in the reference)
        mmin = int(np.floor(min(xmin, ymin)))
        mmax = int(np.ceil(max(xmax, ymax)))
        self.xmin = self.ymin = mmin
        self.xmax = self.ymax = mmax
        self.N = self.x.shape[0]
        self.ordv = np.arange(self.N, dtype=np.int64)
        self.workspace_allocated = False
        self.Levels = [Level(np.array([float(mmin)]), np.array([float(mmin)]), float(mmax - mmin),
                             np.array([0]), np.array([self.N]), np.array([-1]))]
        if self.N > ppl:
            cur = self.Levels[0]
            while True:  # tree.py:499-506 / get_new_level 229-258
                to_split = cur.ns > ppl
                xs, ys, b, t, p = _divide(cur, self.x, self.y, self.ordv, to_split, 0, False)
                new = Level(xs, ys, cur.half_width, b, t, p)
                self.Levels.append(new)
                cur = new
                if not np.any(new.ns > ppl):
                    break
        self.levels = len(self.Levels)
        self.tag_colleagues()
        self.gather_depths()
        self.level_restrict()
        self.split_Xlist()
        for lv in self.Levels:
            lv.not_leaf = np.logical_not(lv.leaf)
        self.leafs = [lv.leaf for lv in self.Levels]
        self.xmids = [lv.xmid for lv in self.Levels]
        self.ymids = [lv.ymid for lv in self.Levels]
        self.children_inds = [lv.children_ind for lv in self.Levels]

    def tag_colleagues(self):  # tree.py:523-529
        for i, lv in enumerate(self.Levels):
            lv.tag_colleagues(None if i == 0 else self.Levels[i - 1])

    def gather_depths(self):  # tree.py:582-585
        for i in reversed(range(self.levels)):
            lv = self.Levels[i]
            lv.get_depths(None if i == self.levels - 1 else self.Levels[i + 1])

    def _split_bad_leaves(self, lv, desc, bads, xlist):  # tree.py:296-311
        xs, ys, b, t, p = _divide(lv, self.x, self.y, self.ordv, bads, desc.n_node, xlist)
        desc.add_nodes(xs, ys, b, t, p, xlist)
        desc.tag_colleagues(lv)

    def level_restrict(self):  # tree.py:533-545
        new_nodes = 1
        while new_nodes > 0:
            new_nodes = 0
            for ind in range(1, self.levels - 2):
                lv, desc = self.Levels[ind], self.Levels[ind + 1]
                bads = np.zeros(lv.n_node, dtype=np.uint8)
                leaf = as_u8(lv.leaf)
                lib().orc_get_bads(ptr(lv.depths), ptr(lv.colleagues), ptr(leaf), ptr(bads),
                                   I64(lv.n_node))
                nb = int(bads.sum())
                self._split_bad_leaves(lv, desc, bads, False)
                self.gather_depths()
                new_nodes += nb

    def split_Xlist(self):  # tree.py:546-576
        for ind in range(self.levels):
            lv = self.Levels[ind]
            if ind == 0 or ind == self.levels - 1:
                lv.add_null_Xlist()
            else:
                lv.get_Xlist()
                if lv.Xlist.sum() > 0:
                    self._split_bad_leaves(lv, self.Levels[ind + 1], lv.Xlist, True)
            if ind > 0:
                cu = np.logical_or(np.logical_and(lv.leaf, np.logical_not(lv.Xlist)), lv.fake_leaf)
                ns = lv.top_ind - lv.bot_ind
                lv.compute_upwards = np.logical_and(cu, ns > 0)
                lv.has_source = ns > 0
                lv.sib_has_source = np.add.reduceat(lv.has_source, np.arange(0, lv.n_node, 4)) > 0
                lv.n_sib_has_source = int(np.sum(lv.sib_has_source))
                lv.allocate_density = np.repeat(lv.sib_has_source, 4)
                lv.n_allocate_density = int(np.sum(lv.allocate_density))
                lv.parent_density_ind = -np.ones(lv.n_node // 4, dtype=np.int64)
                lv.parent_density_ind[lv.sib_has_source] = np.arange(lv.n_sib_has_source)
                lv.this_density_ind = -np.ones(lv.n_node, dtype=np.int64)
                lv.this_density_ind[lv.allocate_density] = np.arange(lv.n_allocate_density)

    def allocate_workspace(self, p, dtype=float):  # tree.py:586-589
        for lv in self.Levels[1:]:
            lv.allocate_workspace(p, dtype)
        self.workspace_allocated = True

    # flattened views used by the target-evaluation loops
    def flat(self):
        off = np.zeros(self.levels + 1, dtype=np.int64)
        for i, lv in enumerate(self.Levels):
            off[i + 1] = off[i] + lv.n_node
        cat = lambda name, dt: np.ascontiguousarray(
            np.concatenate([getattr(lv, name) for lv in self.Levels]).astype(dt))
        return dict(off=off, leaf=cat("leaf", np.uint8), xmid=cat("xmid", np.float64),
                    ymid=cat("ymid", np.float64), children=cat("children_ind", np.int64),
                    bot=cat("bot_ind", np.int64), top=cat("top_ind", np.int64),
                    coll=np.ascontiguousarray(np.concatenate([lv.colleagues for lv in self.Levels])))

    def locate_points(self, x, y):  # tree.py:703-710, 712-727
        x = as_f64(np.ravel(x))
        y = as_f64(np.ravel(y))
        f = self.flat()
        inds = np.zeros(x.size, dtype=np.int64)
        locs = np.zeros(x.size, dtype=np.int64)
        lib().orc_locate_points(ptr(x), ptr(y), I64(x.size), ptr(f["leaf"]), ptr(f["xmid"]),
                                ptr(f["ymid"]), ptr(f["children"]), ptr(f["off"]), ptr(inds), ptr(locs))
        return inds, locs
