"""Quadtree built on the GPU; drop-in for the reference ``pykifmm2d.tree.Tree``.

``Tree(x, y, ppl)`` (reference tree.py:464) sorts the points by exact Morton key on the
device and derives the level-restricted tree with X-list fake children; ``Levels`` exposes
the reference ``Level`` attributes (tree.py:313-420, 558-576) as host numpy arrays, copied
from the device structure on first access.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .device import _torch, as_device_f64, stream_handle


class Level:
    """Host view of one tree level with the reference attribute names (tree.py:313-420)."""

    def __init__(self, arrays, width):
        for k, v in arrays.items():
            setattr(self, k, v)
        self.width = width
        self.half_width = 0.5 * width
        self.ns = self.top_ind - self.bot_ind
        self.xmid = self.xmin + self.half_width
        self.ymid = self.ymin + self.half_width
        self.xmax = self.xmin + width
        self.ymax = self.ymin + width
        self.n_node = self.xmin.shape[0]
        self.short_parent_ind = self.parent_ind[::4]
        self.not_leaf = np.logical_not(self.leaf)

    def _flags(self):
        """tree.py:558-576: compute_upwards, has_source and the density-row maps."""
        cu = np.logical_or(np.logical_and(self.leaf, np.logical_not(self.Xlist)), self.fake_leaf)
        self.compute_upwards = np.logical_and(cu, self.ns > 0)
        self.has_source = self.ns > 0
        self.sib_has_source = np.add.reduceat(self.has_source, np.arange(0, self.n_node, 4)) > 0
        self.n_sib_has_source = int(self.sib_has_source.sum())
        self.allocate_density = np.repeat(self.sib_has_source, 4)
        self.n_allocate_density = int(self.allocate_density.sum())
        self.parent_density_ind = -np.ones(self.n_node // 4, dtype=np.int64)
        self.parent_density_ind[self.sib_has_source] = np.arange(self.n_sib_has_source)
        self.this_density_ind = -np.ones(self.n_node, dtype=np.int64)
        self.this_density_ind[self.allocate_density] = np.arange(self.n_allocate_density)


def export_levels(handle, root_width):
    """All levels of a kf_tree handle as Level views."""
    L = lib()
    info = _lib.TreeInfo()
    check(L.kf_tree_get_info(handle, ctypes.byref(info)))
    out = []
    for lev in range(info.levels):
        n = ctypes.c_int64()
        check(L.kf_tree_level_size(handle, lev, ctypes.byref(n)))
        n = n.value
        a = dict(xmin=np.empty(n), ymin=np.empty(n), bot_ind=np.empty(n, np.int64),
                 top_ind=np.empty(n, np.int64), parent_ind=np.empty(n, np.int64),
                 children_ind=np.empty(n, np.int64), leaf=np.empty(n, np.uint8),
                 fake_leaf=np.empty(n, np.uint8), Xlist=np.empty(n, np.uint8),
                 colleagues=np.empty((n, 9), np.int64))
        order = ["xmin", "ymin", "bot_ind", "top_ind", "parent_ind", "children_ind", "leaf", "fake_leaf",
                 "Xlist", "colleagues"]
        check(L.kf_tree_export_level(handle, lev, *[a[k].ctypes.data_as(ctypes.c_void_p) for k in order]))
        for k in ("leaf", "fake_leaf", "Xlist"):
            a[k] = a[k].astype(bool)
        nr = ctypes.c_int64()
        check(L.kf_tree_level_real(handle, lev, ctypes.byref(nr)))
        a["depths"] = np.empty(nr.value, np.int64)
        check(L.kf_tree_export_depths(handle, lev, a["depths"].ctypes.data_as(ctypes.c_void_p)))
        lv = Level(a, root_width * 2.0 ** -lev)
        if lev > 0:
            lv._flags()
        out.append(lv)
    return out


class Tree:
    """Level-restricted quadtree (reference tree.py:460-589), built on cuda:0."""

    def __init__(self, x, y, ppl, bbox=None):
        self._handle = None
        dx = as_device_f64(x)
        dy = as_device_f64(y)
        if dx.numel() != dy.numel():
            raise ValueError("x and y must have the same length")
        h = ctypes.c_void_p()
        check(lib().kf_tree_build(ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()), int(dx.numel()),
                                  int(ppl), stream_handle(), ctypes.byref(h)))
        self._adopt(h, dx, dy, ppl)

    @classmethod
    def batch(cls, points, ppl):
        """Trees of several point sets built together (shared launches): a list of Tree, in
        the order of `points` (a list of (x, y))."""
        torch = _torch()
        xs = [as_device_f64(x).reshape(-1) for x, _ in points]
        ys = [as_device_f64(y).reshape(-1) for _, y in points]
        for a, b in zip(xs, ys):
            if a.numel() != b.numel():
                raise ValueError("x and y must have the same length")
        dx, dy = torch.cat(xs), torch.cat(ys)
        off = np.concatenate([[0], np.cumsum([a.numel() for a in xs])]).astype(np.int64)
        k = len(points)
        hs = (ctypes.c_void_p * k)()
        check(lib().kf_tree_build_batch(ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()),
                                        off.ctypes.data_as(ctypes.c_void_p), k, int(ppl), stream_handle(), hs))
        out = []
        for t in range(k):
            tr = cls.__new__(cls)
            tr._handle = None
            tr._adopt(ctypes.c_void_p(hs[t]), dx[off[t]:off[t + 1]], dy[off[t]:off[t + 1]], ppl)
            out.append(tr)
        return out

    def _adopt(self, h, dx, dy, ppl):
        self.points_per_leaf = int(ppl)
        self.N = int(dx.numel())
        self._handle = h
        self._dx, self._dy = dx, dy  # keep the inputs alive for the caller's convenience
        info = _lib.TreeInfo()
        check(lib().kf_tree_get_info(h, ctypes.byref(info)))
        self.levels = int(info.levels)
        self.key_depth = int(info.key_depth)
        self.n_boxes = int(info.n_boxes)
        self.xmin = self.ymin = int(info.root_min)
        self.xmax = self.ymax = int(info.root_min + info.root_width)
        self._Levels = None
        self._points = None
        self.workspace_allocated = False

    @property
    def handle(self):
        return self._handle

    @property
    def Levels(self):
        if self._Levels is None:
            self._Levels = export_levels(self._handle, float(self.xmax - self.xmin))
        return self._Levels

    def _export_points(self):
        if self._points is None:
            ordv = np.empty(self.N, np.int64)
            xs = np.empty(self.N)
            ys = np.empty(self.N)
            check(lib().kf_tree_export_points(self._handle, ordv.ctypes.data_as(ctypes.c_void_p),
                                              xs.ctypes.data_as(ctypes.c_void_p),
                                              ys.ctypes.data_as(ctypes.c_void_p)))
            self._points = (ordv, xs, ys)
        return self._points

    @property
    def ordv(self):
        return self._export_points()[0]

    @property
    def x(self):
        return self._export_points()[1]

    @property
    def y(self):
        return self._export_points()[2]

    @property
    def leafs(self):
        return [lv.leaf for lv in self.Levels]

    @property
    def xmids(self):
        return [lv.xmid for lv in self.Levels]

    @property
    def ymids(self):
        return [lv.ymid for lv in self.Levels]

    @property
    def children_inds(self):
        return [lv.children_ind for lv in self.Levels]

    def check_bounding_box(self, x, y):
        """tree.py:680-685, as the reference writes it: all(x < xmin) and all(x > xmax) and
        all(y < ymin) and all(y > ymax), which holds only for empty inputs (the reference calls
        it and ignores the result)."""
        x = np.asarray(x)
        y = np.asarray(y)
        return bool(np.all(x < self.xmin) and np.all(x > self.xmax) and np.all(y < self.ymin)
                    and np.all(y > self.ymax))

    def locate_points(self, x, y):
        """tree.py:703-710 with numba_locate_points 712-727: for each point, the level and the
        position within the level of its leaf (X-list leaves count as leaves), found on the GPU
        (kf_tree_locate).  Returns (inds, locs) shaped like x."""
        torch = _torch()
        shape = np.shape(x)
        dx = as_device_f64(x).reshape(-1).contiguous()
        dy = as_device_f64(y).reshape(-1).contiguous()
        self.check_bounding_box(x, y)
        m = dx.numel()
        inds = torch.empty(m, dtype=torch.int64, device=dx.device)
        locs = torch.empty(m, dtype=torch.int64, device=dx.device)
        check(lib().kf_tree_locate(self._handle, ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()), m,
                                   ctypes.c_void_p(inds.data_ptr()), ctypes.c_void_p(locs.data_ptr()),
                                   stream_handle()))
        return inds.cpu().numpy().reshape(shape), locs.cpu().numpy().reshape(shape)

    def locate_point(self, x, y):
        """tree.py:686-702: (level, position) of the leaf holding the point (x, y)."""
        inds, locs = self.locate_points(np.array([float(x)]), np.array([float(y)]))
        return int(inds[0]), int(locs[0])

    def print_structure(self):
        """tree.py:674-679."""
        for ind, lv in enumerate(self.Levels):
            print("Level", ind + 1, "of", self.levels, "has", lv.n_node, "nodes.")

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                lib().kf_tree_destroy(h)
            except Exception:
                pass
            self._handle = None
