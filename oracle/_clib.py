"""Loader for the oracle's C loops (oracle/ref_loops.c).  TEST INFRASTRUCTURE ONLY.

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may use
this package; the product path never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SRC = os.path.join(_HERE, "ref_loops.c")
_SO = os.path.join(_HERE, "liboracle.so")

_lib = None


def build(force=True):
    """Compile ref_loops.c with gcc/OpenMP into oracle/liboracle.so."""
    if not force and os.path.exists(_SO) and os.path.getmtime(_SO) >= os.path.getmtime(_SRC):
        return _SO
    tmp = _SO + ".tmp%d" % os.getpid()
    cmd = ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=gnu11", _SRC, "-o", tmp, "-lm"]
    subprocess.check_call(cmd)
    os.replace(tmp, _SO)
    return _SO


def _k0_pointer():
    """Address of scipy's compiled Cephes k0 (double (double, int skip_dispatch))."""
    import scipy.special.cython_special as cs
    cap = cs.__pyx_capi__["k0"]
    get = ctypes.pythonapi.PyCapsule_GetPointer
    get.restype = ctypes.c_void_p
    get.argtypes = [ctypes.py_object, ctypes.c_char_p]
    name = ctypes.pythonapi.PyCapsule_GetName
    name.restype = ctypes.c_char_p
    name.argtypes = [ctypes.py_object]
    return get(cap, name(cap))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.orc_set_k0.argtypes = [ctypes.c_void_p]
        L.orc_set_k0(_k0_pointer())
        _lib = L
    return _lib


def ptr(a):
    """Data pointer of a C-contiguous numpy array; the returned ctypes object keeps the array
    alive for the duration of the call (temporaries such as ptr(as_u8(mask)) are safe)."""
    assert a.flags.c_contiguous, "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


I64 = ctypes.c_int64
F64 = ctypes.c_double
INT = ctypes.c_int


def kargs(kernel):
    """(kind, k, scale) ctypes triple for a kernel description dict (built-in kinds 0 and 1;
    a user kernel, kind 2, goes through the numpy paths)."""
    if kernel["kind"] not in (0, 1):
        raise ValueError("the oracle's C loops take the Laplace and modified Helmholtz kernels only")
    return INT(kernel["kind"]), F64(kernel.get("k", 0.0)), F64(kernel.get("scale", 1.0))


def as_u8(b):
    return np.ascontiguousarray(b, dtype=np.uint8)


def as_i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
