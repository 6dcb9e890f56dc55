"""Many independent FMM problems evaluated together.

    F = Forest([(x0, y0), (x1, y1), ...], kernels.Laplace(), p=32, cutoff=50)
    u = F.eval(tau)          # tau: the trees' densities concatenated (or a list, one per tree)

Each problem is the reference's planned FMM on its own points (fmm.py:530-712 fmm_planner,
fmm.py:741-814 planned_fmm); the forest builds the trees together (one segmented bounding
box, key, sort and gather launch sequence for all of them, box structures on host threads)
and plans them as one: every pass of every tree runs in the same launches (one per level for
the upward and downward passes, one each for P2M, M2L, P2P and L2P), and targets of small
trees share MFMA tiles.  Trees with at most two levels are
evaluated by direct sum, as fmm.py:229-231 does.
"""
import ctypes

import numpy as np

from . import _lib, fmm
from . import kernels as _kernels
from ._lib import check, lib
from .device import _torch, as_device_f64, empty_f64, ptr, stream_handle
from .tree import Tree


class Forest:
    def __init__(self, points, kernel=None, p=32, cutoff=50, precision="fp64", batch_build=True):
        kernel = _kernels.as_descriptor(kernel if kernel is not None else _kernels.Laplace())
        if kernel.iscomplex:
            raise TypeError("forests take real kernels (a complex kernel plans one tree at a time)")
        self.kernel = kernel
        self.p = int(p)
        self.cutoff = int(cutoff)
        if batch_build:
            self.trees = Tree.batch(points, self.cutoff)
        else:
            self.trees = [Tree(x, y, self.cutoff) for x, y in points]
        self.sizes = np.array([t.N for t in self.trees], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        self.N = int(self.offsets[-1])
        self.planned = [i for i, t in enumerate(self.trees) if t.levels > 2]
        self.direct = [i for i, t in enumerate(self.trees) if t.levels <= 2]
        self._h = None
        if self.planned:
            arr = (ctypes.c_void_p * len(self.planned))(*[self.trees[i].handle.value for i in self.planned])
            h = ctypes.c_void_p()
            d = kernel.desc()
            check(lib().kf_forest_create(arr, len(self.planned), ctypes.byref(d), self.p,
                                         fmm._precision(precision), stream_handle(), ctypes.byref(h)))
            self._h = h
        self.n_planned = int(sum(self.sizes[i] for i in self.planned))

    @property
    def handle(self):
        return self._h

    def autotune(self, nrhs=1):
        """kf_plan_autotune on the forest plan: launch structure and tile shapes timed on the
        forest itself; returns (tuning dict, ms per evaluation)."""
        t = _lib.Tuning()
        ms = ctypes.c_double()
        check(lib().kf_plan_autotune(self._h, int(nrhs), stream_handle(), ctypes.byref(t), ctypes.byref(ms)))
        return {f: getattr(t, f) for f, _ in _lib.Tuning._fields_}, float(ms.value)

    def stats(self):
        s = _lib.PlanStats()
        check(lib().kf_plan_stats_get(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in _lib.PlanStats._fields_}

    def store_matrices(self, on=True):
        """Keep the planned trees' near-field, P2M and L2P kernel matrices in HBM (see
        fmm._Plan.store_matrices); the direct-sum trees are unaffected."""
        if self._h is not None:
            check(lib().kf_plan_store_matrices(self._h, 1 if on else 0, stream_handle()))

    def eval(self, tau):
        """Potentials of every tree at its own points.  tau: (N,) or (nrhs, N), the trees'
        densities concatenated in tree order (device or host), or a list of per-tree arrays;
        the result has the same form."""
        torch = _torch()
        as_list = isinstance(tau, (list, tuple))
        if as_list:
            if len(tau) != len(self.trees):
                raise ValueError("one density per tree")
            dt = torch.cat([as_device_f64(t).reshape(-1) for t in tau])
        else:
            dt = as_device_f64(tau)
        if dt.numel() % self.N:
            raise ValueError("densities must have %d values per right-hand side" % self.N)
        nrhs = dt.numel() // self.N
        dt = dt.reshape(nrhs, self.N)
        pot = empty_f64((nrhs, self.N))
        if self._h is not None:
            if not self.direct:
                src, dst = dt, pot
            else:
                src = torch.cat([dt[:, self.offsets[i]:self.offsets[i + 1]] for i in self.planned], dim=1).contiguous()
                dst = empty_f64((nrhs, self.n_planned))
            check(lib().kf_plan_eval(self._h, ptr(src), self.n_planned, nrhs, ptr(dst), self.n_planned,
                                     stream_handle()))
            if self.direct:
                o = 0
                for i in self.planned:
                    n = int(self.sizes[i])
                    pot[:, self.offsets[i]:self.offsets[i] + n] = dst[:, o:o + n]
                    o += n
        for i in self.direct:
            t = self.trees[i]
            a, b = int(self.offsets[i]), int(self.offsets[i + 1])
            for r in range(nrhs):
                pot[r, a:b] = fmm._direct(self.kernel, t._dx, t._dy, dt[r, a:b].contiguous(), t._dx, t._dy, 1)
        if as_list:
            return [fmm._out(pot[:, self.offsets[i]:self.offsets[i + 1]].reshape(-1) if nrhs == 1 else
                             pot[:, self.offsets[i]:self.offsets[i + 1]], tau[i]) for i in range(len(self.trees))]
        one = tau.dim() == 1 if hasattr(tau, "dim") else np.ndim(tau) == 1
        return fmm._out(pot.reshape(-1) if one else pot, tau)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().kf_plan_destroy(h)
            except Exception:
                pass
            self._h = None
