"""GPU drop-in for the reference ``pykifmm2d.fmm`` module.

Same entry points, argument order and return values as the reference (fmm.py):
  Get_Kernel_Functions(Kernel_Eval)            -> (KF, KA, KAS)          fmm.py:119-149
  prepare_numba_functions_on_the_fly(K)        -> 3-tuple of pass handles fmm.py:29-117
  prepare_numba_functions_planned(K)           -> 2-tuple                 fmm.py:426-493
  Kernel_Form(KF, sx, sy, tx, ty, out)                                    fmm.py:151-168
  Kernel_Apply(KA, KAS, sx, sy, tau, tx, ty, out)                         fmm.py:170-188
  on_the_fly_fmm(x, y, tau, Nequiv, Ncutoff, kernel_functions, numba_functions, verbose)
                                               -> (potential, tree)       fmm.py:190-236
  fmm_planner(x, y, Nequiv, Ncutoff, kernel_functions, numba_functions, verbose) -> FMM_Plan
  planned_fmm(fmm_plan, tau)                   -> potential               fmm.py:530-814
``Kernel_Eval`` is a kernel descriptor from ``pykifmm2d_amd.kernels``.  Arrays may be numpy
(results come back as numpy) or torch tensors (results stay on the device).  The potential
is u_i = sum_{j != i} G(x_i, x_j) tau_j in input order.
"""
import ctypes
import time

import numpy as np

from . import _lib
from ._lib import check, lib
from .device import as_device_f64, empty_f64, ptr, stream_handle, to_host
from .kernels import as_descriptor
from .tree import Tree


def get_level_information(node_width, theta):
    """Equivalent / check circles of a level (fmm.py:12-22)."""
    dd = 0.01
    r1 = 0.5 * node_width * (np.sqrt(2) + dd)
    r2 = 0.5 * node_width * (4 - np.sqrt(2) - 2 * dd)
    return r1 * np.cos(theta), r1 * np.sin(theta), r2 * np.cos(theta), r2 * np.sin(theta), r1, r2


def get_print_function(verbose):
    return print if verbose else (lambda *a, **k: None)


def _is_torch(a):
    try:
        import torch
        return isinstance(a, torch.Tensor)
    except ImportError:
        return False


def _out(result, like):
    return result if _is_torch(like) else to_host(result)


def _store(out, result):
    if out is None:
        return None
    if _is_torch(out):
        out.copy_(result.reshape(out.shape).to(out.dtype))
    else:
        out[...] = to_host(result).reshape(out.shape)
    return out


class _KernelCall:
    """GPU kernel-matrix / direct-sum callable bound to a kernel descriptor."""

    def __init__(self, kernel, what):
        self.kernel = as_descriptor(kernel)
        self.what = what

    def __repr__(self):
        return "<%s %r>" % (self.what, self.kernel)


def _kernel_form(kernel, sx, sy, tx, ty):
    """(nt, ns) device matrix G(s_i, t_j); a complex kernel's two parts as two real forms."""
    if kernel.iscomplex:
        import torch
        re, im = (_kernel_form(k, sx, sy, tx, ty) for k in kernel.parts())
        return torch.complex(re, im)
    dsx, dsy, dtx, dty = map(as_device_f64, (sx, sy, tx, ty))
    res = empty_f64((dtx.numel(), dsx.numel()))
    d = kernel.desc()
    check(lib().kf_kernel_form(ctypes.byref(d), ptr(dsx), ptr(dsy), dsx.numel(), ptr(dtx), ptr(dty),
                               dtx.numel(), ptr(res), stream_handle()))
    return res


class _KF(_KernelCall):
    def __call__(self, sx, sy, tx, ty, out):
        """out[j, i] = G(s_i, t_j) (fmm.py:120-126)."""
        return _store(out, _kernel_form(self.kernel, sx, sy, tx, ty))


def _is_complex(a):
    return a.is_complex() if _is_torch(a) else np.iscomplexobj(a)


def _split(tau):
    """(real part, imaginary part) of a complex density, host or device."""
    if _is_torch(tau):
        return tau.real, tau.imag
    t = np.asarray(tau)
    return t.real, t.imag


def _direct(kernel, sx, sy, tau, tx, ty, mode):
    """Direct sum on the device.  A complex kernel runs its two parts as real kernels; a complex
    density its two parts as real densities (the sums are linear)."""
    if kernel.iscomplex or _is_complex(tau):
        import torch
        if _is_complex(tau):
            tr, ti = _split(tau)
            return _direct(kernel, sx, sy, tr, tx, ty, mode) + 1j * _direct(kernel, sx, sy, ti, tx, ty, mode)
        re, im = kernel.parts()
        return torch.complex(_direct(re, sx, sy, tau, tx, ty, mode), _direct(im, sx, sy, tau, tx, ty, mode))
    dsx, dsy, dtau, dtx, dty = map(as_device_f64, (sx, sy, tau, tx, ty))
    res = empty_f64((dtx.numel(),))
    d = kernel.desc()
    check(lib().kf_direct(ctypes.byref(d), ptr(dsx), ptr(dsy), ptr(dtau), dsx.numel(), ptr(dtx), ptr(dty),
                          None, dtx.numel(), mode, ptr(res), stream_handle()))
    return res


class _KA(_KernelCall):
    def __call__(self, sx, sy, tx, ty, tau, out):
        """out_j = sum_i G(s_i, t_j) tau_i (fmm.py:129-136)."""
        return _store(out, _direct(self.kernel, sx, sy, tau, tx, ty, 0))


class _KAS(_KernelCall):
    def __call__(self, sx, sy, tau, out):
        """out_j = sum_{i != j} G(s_i, s_j) tau_i (fmm.py:139-147)."""
        return _store(out, _direct(self.kernel, sx, sy, tau, sx, sy, 1))


def Get_Kernel_Functions(Kernel_Eval):
    """(KF, KA, KAS) for a kernel descriptor (fmm.py:119-149)."""
    return _KF(Kernel_Eval, "KF"), _KA(Kernel_Eval, "KA"), _KAS(Kernel_Eval, "KAS")


class _PassHandle(_KernelCall):
    """Stands in for a numba pass function; the passes themselves run in libkifmm2d."""

    def __call__(self, *args, **kwargs):
        raise TypeError("%s runs inside libkifmm2d and is not callable on its own" % self.what)


def prepare_numba_functions_on_the_fly(Kernel_Eval):
    """fmm.py:29-117: (neighbour, upward, downward) pass handles."""
    return (_PassHandle(Kernel_Eval, "evaluate_neighbor_interactions"),
            _PassHandle(Kernel_Eval, "numba_upwards_pass"),
            _PassHandle(Kernel_Eval, "numba_downwards_pass2"))


def prepare_numba_functions_planned(Kernel_Eval):
    """fmm.py:426-493: (neighbour, upward) builder handles."""
    return (_PassHandle(Kernel_Eval, "build_neighbor_interactions"),
            _PassHandle(Kernel_Eval, "build_upwards_pass"))


def kernel_of(kernel_functions, numba_functions):
    """The kernel descriptor carried by the factory tuples (they must agree)."""
    ks = {id(getattr(f, "kernel", None)): getattr(f, "kernel", None)
          for f in tuple(kernel_functions) + tuple(numba_functions)}
    ks.pop(id(None), None)
    if not ks:
        raise TypeError("kernel_functions / numba_functions must come from this package's factories")
    descs = list(ks.values())
    first = descs[0]
    for d in descs[1:]:
        if d.key() != first.key():
            raise ValueError("kernel_functions and numba_functions were built for different kernels")
    return first


def Kernel_Form(KF, sx, sy, tx=None, ty=None, out=None):
    """fmm.py:151-168 (zero diagonal for a self matrix)."""
    isself = tx is None or ty is None or (sx is tx and sy is ty)
    if tx is None or ty is None:
        tx, ty = sx, sy
    m = _kernel_form(KF.kernel, sx, sy, tx, ty)
    if isself:
        m.fill_diagonal_(0.0)
    if out is not None:
        return _store(out, m)
    return _out(m, sx)


def Kernel_Apply(KA, KAS, sx, sy, tau, tx=None, ty=None, out=None):
    """fmm.py:170-188: direct sum, self-excluding when the targets are the sources."""
    isself = tx is None or ty is None or (sx is tx and sy is ty)
    if tx is None or ty is None:
        tx, ty = sx, sy
    res = _direct(KAS.kernel if isself else KA.kernel, sx, sy, tau, tx, ty, 1 if isself else 0)
    if out is not None:
        return _store(out, res)
    return _out(res, tau)


def _torch_f64():
    import torch
    return torch.float64


class Workspace:
    """Expansion slabs shared by plans that never evaluate at the same time (kf_workspace_create):
    attach the plans with _Plan.attach_workspace and order their evaluations (one stream)."""

    def __init__(self):
        self._h = None
        h = ctypes.c_void_p()
        check(lib().kf_workspace_create(ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def nbytes(self):
        """Device bytes the workspace holds now."""
        b = ctypes.c_int64()
        check(lib().kf_workspace_bytes(self._h, ctypes.byref(b)))
        return int(b.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().kf_workspace_destroy(h)
            except Exception:
                pass
            self._h = None


class _Plan:
    """Owns a kf_plan handle for a tree."""

    def __init__(self, tree, kernel, p, precision=_lib.KF_PREC_FP64):
        self.tree = tree
        self.kernel = kernel
        self.p = int(p)
        self._h = None
        self._ws = None
        h = ctypes.c_void_p()
        d = kernel.desc()
        check(lib().kf_plan_create(tree.handle, ctypes.byref(d), self.p, precision, stream_handle(),
                                   ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def stats(self):
        s = _lib.PlanStats()
        check(lib().kf_plan_stats_get(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in _lib.PlanStats._fields_}

    def eval(self, tau):
        """Potentials at the sources for tau (n,) or (nrhs, n); complex for a complex kernel
        (whose plan takes real densities: a complex tau runs as its two parts)."""
        if _is_complex(tau):
            tr, ti = _split(tau)
            return self.eval(tr) + 1j * self.eval(ti)
        dt = as_device_f64(tau)
        n = self.tree.N
        nrhs = 1 if dt.dim() == 1 else dt.shape[0]
        if dt.numel() != nrhs * n:
            raise ValueError("tau must have length %d per right-hand side" % n)
        if self.kernel.iscomplex:
            import torch
            pot = torch.empty(tuple(dt.shape), dtype=torch.complex128, device="cuda")
        else:
            pot = empty_f64(tuple(dt.shape))
        check(lib().kf_plan_eval(self._h, ptr(dt), n, nrhs, ptr(pot), n, stream_handle()))
        return pot

    def eval_ptrs(self, taus, pots, stream=None):
        """pots[r] = potential of taus[r]: device float64 tensors of n values each, at any
        addresses; enqueued on `stream` (a torch stream; default the current one)."""
        k = len(taus)
        if k != len(pots) or k == 0:
            raise ValueError("need as many potential buffers as densities (at least one)")
        n = self.tree.N
        import torch
        pot_dtype = torch.complex128 if self.kernel.iscomplex else torch.float64
        for t in taus:
            if t.numel() != n or not t.is_cuda or t.dtype != _torch_f64() or not t.is_contiguous():
                raise ValueError("densities must be contiguous float64 cuda tensors of length %d" % n)
        for t in pots:
            if t.numel() != n or not t.is_cuda or t.dtype != pot_dtype or not t.is_contiguous():
                raise ValueError("potentials must be contiguous %s cuda tensors of length %d" % (pot_dtype, n))
        arr = ctypes.c_void_p * k
        ti = arr(*[t.data_ptr() for t in taus])
        po = arr(*[t.data_ptr() for t in pots])
        check(lib().kf_plan_eval_ptrs(self._h, ti, po, k, stream_handle(stream)))

    def tuning(self):
        t = _lib.Tuning()
        check(lib().kf_plan_get_tuning(self._h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in _lib.Tuning._fields_}

    def set_tuning(self, near_grid=0, near_fork=0, m2l_pipeline=0, level_rows=0, near_split=0, chain=0,
                   batch_form=0, near_form=0):
        t = _lib.Tuning(int(near_grid), int(near_fork), int(m2l_pipeline), int(level_rows), int(near_split),
                        int(chain), int(batch_form), int(near_form))
        check(lib().kf_plan_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, nrhs=1, stream=None):
        """Time the candidate launch structures for batches of nrhs and keep the fastest;
        returns (tuning dict, milliseconds per evaluation)."""
        t = _lib.Tuning()
        ms = ctypes.c_double()
        check(lib().kf_plan_autotune(self._h, int(nrhs), stream_handle(stream), ctypes.byref(t), ctypes.byref(ms)))
        return {f: getattr(t, f) for f, _ in _lib.Tuning._fields_}, float(ms.value)

    def store_matrices(self, on=True, stream=None):
        """Keep the kernel matrices in HBM, as the reference's fmm_planner does (fmm.py:646-711:
        near field, P2M and L2P; dense per-leaf blocks here): 8 bytes per value
        (stats()['p2p_pairs'] + N + p2m_evals + l2p_evals); evaluations read them instead of
        evaluating the kernel, which pays for costly kernels (tables, K0).  on=False frees them."""
        check(lib().kf_plan_store_matrices(self._h, 1 if on else 0, stream_handle(stream)))

    def set_graphs(self, enable=True):
        """Replay each batch width's evaluation as a captured HIP graph (one launch per batch)."""
        check(lib().kf_plan_set_graphs(self._h, 1 if enable else 0))

    def reserve(self, nrhs, stream=None):
        """Lay the plan out for batches of up to nrhs (1..64) now (no allocation later)."""
        check(lib().kf_plan_reserve(self._h, int(nrhs), stream_handle(stream)))

    def attach_workspace(self, ws):
        """Keep the expansion slabs in Workspace `ws` (None: back to slabs of the plan's own)."""
        check(lib().kf_plan_attach_workspace(self._h, ws.handle if ws is not None else None))
        self._ws = ws  # the workspace outlives the plan's use of it

    def slab_bytes(self, nrhs):
        """Bytes of expansion slabs a batch of nrhs (1..64) needs: what a workspace holds for it."""
        b = ctypes.c_int64()
        check(lib().kf_plan_slab_bytes(self._h, int(nrhs), ctypes.byref(b)))
        return int(b.value)

    def footprint(self, nrhs):
        """Device bytes of the plan and its tree laid out for batches of nrhs (1..64)."""
        b = ctypes.c_int64()
        check(lib().kf_plan_footprint(self._h, int(nrhs), ctypes.byref(b)))
        return int(b.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().kf_plan_destroy(h)
            except Exception:
                pass
            self._h = None


PRECISIONS = {"fp64": _lib.KF_PREC_FP64, "split": _lib.KF_PREC_SPLIT, "split16": _lib.KF_PREC_SPLIT16}


def _precision(precision):
    try:
        return PRECISIONS[precision]
    except KeyError:
        raise ValueError("precision must be 'fp64', 'split' or 'split16'") from None


def on_the_fly_fmm(x, y, tau, Nequiv, Ncutoff, kernel_functions, numba_functions, verbose=False,
                   precision="fp64"):
    """On-the-fly KIFMM (fmm.py:190-236) -> (potential, tree).

    precision="split" runs M2L in FP32 with per-level scaling (everything else FP64); relative
    error about 1e-7 instead of about 1e-14 at p = 64."""
    my_print = get_print_function(verbose)
    my_print("\nBeginning FMM")
    kernel = kernel_of(kernel_functions, numba_functions)
    st = time.time()
    tree = Tree(x, y, Ncutoff)
    my_print("....Tree formed in:            {:0.1f}".format((time.time() - st) * 1000))
    if tree.levels <= 2:
        pot = _direct(kernel, x, y, tau, x, y, 1)  # fmm.py:229-231
    else:
        plan = _Plan(tree, kernel, Nequiv, _precision(precision))
        pot = plan.eval(tau)
        tree._plan = plan
    my_print("FMM completed in               {:0.1f}".format((time.time() - st) * 1000))
    return _out(pot, tau), tree


class FMM_Plan:
    """fmm.py:512-528: a resident tree, operators and work lists in HBM."""

    def __init__(self, tree, plan, kernel, Nequiv, verbose):
        self.tree = tree
        self.plan = plan
        self.kernel = kernel
        self.Nequiv = Nequiv
        self.verbose = verbose
        self.theta = np.linspace(0, 2 * np.pi, Nequiv, endpoint=False)

    def extract(self):
        return self.tree, self.plan, self.kernel, self.Nequiv, self.verbose


def fmm_planner(x, y, Nequiv, Ncutoff, kernel_functions, numba_functions, verbose=False, precision="fp64",
                store_matrices=False):
    """fmm.py:530-712 (precision: see on_the_fly_fmm).  The plan is matrix-free by default;
    store_matrices=True keeps the near-field, P2M and L2P kernel matrices, as the reference's
    plan does (fmm.py:646-711), for kernels costly enough that reading beats evaluating them."""
    my_print = get_print_function(verbose)
    my_print("\nPlanning FMM")
    kernel = kernel_of(kernel_functions, numba_functions)
    st = time.time()
    tree = Tree(x, y, Ncutoff)
    my_print("....Tree formed in:              {:0.1f}".format((time.time() - st) * 1000))
    plan = _Plan(tree, kernel, Nequiv, _precision(precision)) if tree.levels >= 2 else None
    if plan is not None and store_matrices:
        plan.store_matrices(True)
    return FMM_Plan(tree, plan, kernel, Nequiv, verbose)


def planned_fmm(fmm_plan, tau):
    """fmm.py:741-814."""
    tree, plan, kernel, _, verbose = fmm_plan.extract()
    get_print_function(verbose)("Executing FMM")
    if plan is None:
        dx, dy, _ = _tree_inputs(tree)
        return _out(_direct(kernel, dx, dy, tau, dx, dy, 1), tau)
    return _out(plan.eval(tau), tau)


def _tree_inputs(tree):
    return tree._dx, tree._dy, None
