"""GPU drop-in for the reference ``pykifmm2d.class_fmm`` module (class_fmm.py:166-402).

    fmm = FMM(x, y, kernel_functions, numba_functions, Nequiv=48, Ncutoff=50)
    fmm.precompute()               # tree + operators + work lists, resident in HBM
    fmm.build_expansions(tau)      # upward / downward passes
    u = fmm.evaluate_to_points(tx, ty, check_self=True)

Differences from the reference, all on inputs where the reference misbehaves: repeated
``build_expansions`` calls are independent (the reference double-counts P2M, its
Check_Us is never re-zeroed); a one-level tree (N <= Ncutoff) is evaluated by direct sum
instead of failing.

``iscomplex=True`` (class_fmm.py:167, 174, 187: complex workspaces) takes complex densities
and complex kernels (``kernels.ComplexTabulated``, or a Kernel_Eval returning complex values
such as examples/hankel.py's 0.25j H0(k r)): the real and imaginary parts of the density are
expanded together as two right-hand sides of one batched evaluation and recombined at the
targets, u = u(Re tau) + i u(Im tau).
"""
import ctypes
import time

import numpy as np

from . import _lib
from ._lib import check, lib
from .device import as_device_f64, empty_f64, ptr, stream_handle, to_host
from .fmm import (Get_Kernel_Functions, Kernel_Form, _direct, _is_torch, _out, _Plan, _precision,  # noqa: F401
                  get_level_information, kernel_of, prepare_numba_functions_planned)
from .fmm import prepare_numba_functions_on_the_fly as _otf
from .tree import Tree


def prepare_numba_functions_on_the_fly(Kernel_Eval):
    """class_fmm.py:55-113: (upward, target L2P, target P2P) pass handles."""
    a, b, c = _otf(Kernel_Eval)
    a.what, b.what, c.what = ("numba_upwards_pass", "numba_target_local_expansion_evaluation",
                              "numba_target_neighbor_evaluation")
    return a, b, c


def get_print_function(verbose):
    return print if verbose else (lambda *a, **k: None)


class FMM:
    def __init__(self, x, y, kernel_functions, numba_functions, Nequiv=48, Ncutoff=50, iscomplex=False,
                 verbose=False, precision="fp64"):
        self.iscomplex = bool(iscomplex)
        self.x = x
        self.y = y
        self.Nequiv = int(Nequiv)
        self.Ncutoff = int(Ncutoff)
        self.kernel_functions = kernel_functions
        self.numba_functions = numba_functions
        self.kernel = kernel_of(kernel_functions, numba_functions)
        if self.kernel.iscomplex and not self.iscomplex:
            raise TypeError("a complex kernel needs FMM(..., iscomplex=True) (class_fmm.py:167, 174)")
        self.dtype = np.complex128 if self.iscomplex else np.float64
        self.verbose = verbose
        self.print = get_print_function(verbose)
        self.precision = _precision(precision)
        self._plan = None
        self._tau = None

    def build_tree(self):
        st = time.time()
        self.tree = Tree(self.x, self.y, self.Ncutoff)
        self.print("....Tree formed in:             {:0.1f}".format((time.time() - st) * 1000))

    def precompute(self):
        if not hasattr(self, "tree"):
            self.build_tree()
        st = time.time()
        if self.tree.levels >= 2:
            self._plan = _Plan(self.tree, self.kernel, self.Nequiv, self.precision)
        self.print("....Time for precomputations:   {:0.2f}".format(1000 * (time.time() - st)))

    @property
    def plan(self):
        return self._plan

    def build_expansions(self, tau):
        if not hasattr(self, "tree"):
            self.precompute()
        st = time.time()
        self.tau = tau
        n = self.tree.N
        if self.iscomplex:
            # [real; imag] as two right-hand sides
            if hasattr(tau, "data_ptr"):
                import torch
                t = tau.reshape(-1)
                parts = torch.stack([t.real, t.imag]) if t.is_complex() else torch.stack([t, torch.zeros_like(t)])
                self._tau = as_device_f64(parts)
            else:
                t = np.ravel(np.asarray(tau, dtype=np.complex128))
                self._tau = as_device_f64(np.stack([t.real, t.imag]))
        else:
            if (hasattr(tau, "is_complex") and tau.is_complex()) or np.iscomplexobj(tau):
                raise TypeError("complex densities need FMM(..., iscomplex=True)")
            self._tau = as_device_f64(tau)
        self._nrhs = 2 if self.iscomplex else 1
        if self._tau.numel() != self._nrhs * n:
            raise ValueError("tau must have one entry per source point")
        if self._plan is not None:
            check(lib().kf_plan_expand(self._plan.handle, ptr(self._tau), n, self._nrhs, stream_handle()))
        self.print("....Time for expansions:        {:0.2f}".format(1000 * (time.time() - st)))

    def evaluate_to_points(self, x, y, check_self=False):
        """class_fmm.py:379-399: potential at arbitrary targets inside the root box."""
        if self._tau is None:
            raise RuntimeError("build_expansions must be called first")
        dx = as_device_f64(np.ravel(x) if not hasattr(x, "data_ptr") else x.reshape(-1))
        dy = as_device_f64(np.ravel(y) if not hasattr(y, "data_ptr") else y.reshape(-1))
        m = dx.numel()
        ck = self.kernel.iscomplex
        import torch
        pot = torch.empty((self._nrhs, m), dtype=torch.complex128 if ck else torch.float64, device="cuda")
        if self._plan is None:
            mode = 2 if check_self else 0
            for r in range(self._nrhs):
                pot[r] = _direct(self.kernel, self.tree._dx, self.tree._dy, self._tau.reshape(self._nrhs, -1)[r],
                                 dx, dy, mode)
        else:  # complex kernels: pot[r] complex (interleaved re, im doubles)
            check(lib().kf_plan_targets(self._plan.handle, ptr(dx), ptr(dy), m, 1 if check_self else 0,
                                        ptr(pot), stream_handle()))
        if self.iscomplex:
            u = pot[0] + 1j * pot[1]  # u(Re tau) + i u(Im tau)
            return u if _is_torch(x) else to_host(u)
        return _out(pot[0], x)

    def evaluate_to_point(self, x, y):
        """class_fmm.py:353-378 (single target, no self check)."""
        pot = self.evaluate_to_points(np.array([float(x)]), np.array([float(y)]))
        return np.asarray(pot, dtype=self.dtype)
