"""Kernel descriptors (the GPU build cannot JIT a Python Green's function).

In the reference a kernel is a numba function ``Kernel_Eval(sx, sy, tx, ty)``
(examples/laplace.py:30-33); here the factories in ``fmm`` / ``class_fmm`` take one of these
descriptors in its place.  Both kernels are radially symmetric, as the reference requires.
"""
import math

from . import _lib


class KernelDescriptor:
    kind = None

    def desc(self):
        return _lib.KernelDesc(self.kind, getattr(self, "k", 0.0), getattr(self, "scale", 1.0))


class Laplace(KernelDescriptor):
    """G(s, t) = -(1/4pi) log |t - s|^2  (reference kernels/laplace.py:5-8)."""
    kind = _lib.KF_LAPLACE

    def __repr__(self):
        return "Laplace()"


class ModifiedHelmholtz(KernelDescriptor):
    """G(s, t) = scale * K0(k |t - s|)  (reference kernels/modified_helmholtz.py:7-9 has
    scale = 1; its examples use 1/(2pi), the default here)."""
    kind = _lib.KF_MODHELM

    def __init__(self, k, scale=1.0 / (2.0 * math.pi)):
        if not k > 0:
            raise ValueError("k must be positive")
        self.k = float(k)
        self.scale = float(scale)

    def __repr__(self):
        return "ModifiedHelmholtz(k=%r, scale=%r)" % (self.k, self.scale)


laplace_eval = Laplace()


def generate_modified_helmholtz_functions(k, scale=1.0):
    """Descriptor for the reference module's kernel (kernels/modified_helmholtz.py:6-9)."""
    return ModifiedHelmholtz(k, scale)


def as_descriptor(kernel):
    if isinstance(kernel, KernelDescriptor):
        return kernel
    raise TypeError("kernel must be a pykifmm2d_amd.kernels descriptor (Laplace() or "
                    "ModifiedHelmholtz(k, scale)); numba kernel functions cannot run on the GPU")
