"""Kernel descriptors.

In the reference a kernel is a numba function ``Kernel_Eval(sx, sy, tx, ty)``
(examples/laplace.py:30-33) that its loops call for every pair.  Here the factories in ``fmm`` /
``class_fmm`` take a descriptor: ``Laplace()`` and ``ModifiedHelmholtz(k, scale)`` are compiled
into the GPU kernels; any other radial kernel -- a function of r^2, or a reference-style
``Kernel_Eval`` -- becomes a ``Tabulated`` kernel, sampled once into piecewise polynomials in
log2(r^2) that the GPU evaluates.  The reference's passes already assume a radial kernel
(fmm.py:373-374 form one M2L matrix per offset class and level).
"""
import ctypes
import math
import weakref

import numpy as np

from . import _lib


class KernelDescriptor:
    kind = None
    iscomplex = False

    def desc(self):
        return _lib.KernelDesc(self.kind, getattr(self, "k", 0.0), getattr(self, "scale", 1.0), None)

    def key(self):
        """Identity of the kernel (two descriptors with equal keys compute the same G)."""
        return (type(self).__name__, getattr(self, "k", 0.0), getattr(self, "scale", 1.0))


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


class Tabulated(KernelDescriptor):
    """A user's radial kernel G(r^2), tabulated for the GPU (KF_TABLE).

    ``g`` maps an array of squared distances r^2 > 0 to G (numpy float64, vectorised).  u =
    log2(r^2) is cut into segments of width ``du`` (a power of two, 1/4 by default) from
    ``u_min`` to ``u_max``: by default -1022 (r^2 = 2^-1022, the smallest normal double) and
    256, less any end segments where g or its polynomial is not finite.  On each segment G is
    interpolated at Chebyshev points (17, or deg + 1 if more) and the Chebyshev series is cut
    after degree ``deg`` -- by default the lowest degree whose error is within twice the best (1
    for log r^2, which is linear in u; 13 for K0) -- and stored as a polynomial in the
    segment's local x in [-1, 1].  Outside [2^u_min, 2^u_max] G is held at the end values;
    G(0) = ``g0`` (default g(0), which may be infinite, or nan where g raises: the FMM never
    evaluates coincident points of one set, like the reference's i != j sums).

    The fit is checked at construction, at points between the interpolation nodes: the
    largest error on a segment, relative to the largest |G| on it (or to 2^-52 times the largest
    |G| down to r / 2^10, where G has decayed below that), must stay below ``tol``
    (``fit_error`` holds it).  The default grid (5112 segments) reproduces log r^2 and K0(k r)
    to a few units in the 15th digit.  A plan needs the table to reach (2 W)^2, W the root box
    width.
    """
    kind = _lib.KF_TABLE

    def __init__(self, g, g0=None, u_min=None, u_max=None, du=0.25, deg=None, tol=1e-13, name=None):
        du = float(du)
        m, e = math.frexp(du)
        if m != 0.5 or not (-9 <= e <= 5):
            raise ValueError("du must be a power of two in [2^-10, 16]")
        s0 = math.floor((-1022.0 if u_min is None else float(u_min)) / du)
        s1 = math.ceil((256.0 if u_max is None else float(u_max)) / du)
        if s1 <= s0:
            raise ValueError("u_max must exceed u_min")
        if deg is not None and not 0 <= int(deg) <= 24:
            raise ValueError("deg must be in [0, 24]")
        self.g = g
        self.name = name or getattr(g, "__name__", "g")
        # interpolate at N + 1 Chebyshev points per segment (N = 16, or deg if larger)
        N = max(16, int(deg) if deg is not None else 0)
        n = N + 1
        t = np.cos(np.pi * (np.arange(n) + 0.5) / n)
        j = np.arange(s0, s1, dtype=np.float64)[:, None]
        vals = self._g(np.exp2((j + 0.5 * (t[None, :] + 1.0)) * du))
        # Chebyshev coefficients a_k = (2/n) sum_i v_i T_k(t_i), a_0 halved
        T = np.cos(np.outer(np.arange(n), np.pi * (np.arange(n) + 0.5) / n))
        with np.errstate(all="ignore"):
            a = vals @ T.T * (2.0 / n)
            a[:, 0] *= 0.5
            C = np.zeros((n, n))
            for k in range(n):
                C[:k + 1, k] = np.polynomial.chebyshev.cheb2poly(np.eye(k + 1)[k])
            mono = a @ C.T
        # check between the nodes: 2 N + 3 points per segment, ends included
        xc = np.cos(np.pi * np.arange(2 * n + 1) / (2 * n))
        ref = self._g(np.exp2((j + 0.5 * (xc[None, :] + 1.0)) * du))
        # a default end drops the segments where g or its polynomial is not finite there (1/r^2
        # near r^2 = 2^-1022: finite values whose coefficient sums overflow)
        bad = ~(np.isfinite(vals).all(1) & np.isfinite(ref).all(1) & np.isfinite(mono).all(1))
        lo, hi = 0, bad.size
        while u_min is None and lo < hi and bad[lo]:
            lo += 1
        while u_max is None and hi > lo and bad[hi - 1]:
            hi -= 1
        if hi == lo or bad[lo:hi].any():
            raise ValueError("g is not finite on the table's range [2^%g, 2^%g]; narrow u_min/u_max"
                             % ((s0 + lo) * du, (s0 + hi) * du))
        vals, ref, a = vals[lo:hi], ref[lo:hi], a[lo:hi]
        s0, s1 = s0 + lo, s0 + hi
        self.du, self.s0, self.nseg = du, s0, s1 - s0
        self.u_min, self.u_max = s0 * du, s1 * du
        seg = np.max(np.abs(ref), axis=1)
        # the scale of a segment: its largest |G|, or the rounding unit of the largest |G| from
        # it down to r / 2^10 where a fast-decaying G has fallen below that (an error there is
        # lost in the rounding of nearer pairs' values)
        w = int(round(20.0 / du))
        near = np.lib.stride_tricks.sliding_window_view(np.concatenate([np.zeros(w), seg]), w + 1).max(axis=1)
        scale = np.maximum(seg, 2.0 ** -52 * near)

        def truncated(d):
            """Monomial coefficients in x of the series cut after T_d, and its error."""
            C = np.zeros((d + 1, d + 1))
            for k in range(d + 1):
                C[:k + 1, k] = np.polynomial.chebyshev.cheb2poly(np.eye(k + 1)[k])
            coef = np.ascontiguousarray(a[:, :d + 1] @ C.T)
            err = np.max(np.abs(np.polynomial.polynomial.polyval(xc, coef.T) - ref), axis=1)
            rel = np.where(scale > 1e-290, err / np.maximum(scale, 1e-290), err)
            return coef, float(np.max(rel))

        if deg is not None:
            self.coef, self.fit_error = truncated(int(deg))
        else:
            # the lowest degree within twice the best error (rounding sets the floor)
            fits = [truncated(d) for d in range(1, N + 1)]
            best = min(f[1] for f in fits)
            self.coef, self.fit_error = next(f for f in fits if f[1] <= 2.0 * best)
        self.deg = self.coef.shape[1] - 1
        if not self.fit_error <= tol:
            raise ValueError("table fit error %.3g exceeds tol %.3g; use a smaller du or a larger deg"
                             % (self.fit_error, tol))
        if g0 is None:
            try:
                with np.errstate(all="ignore"):
                    g0 = float(np.asarray(g(np.zeros(1)), dtype=np.float64).reshape(-1)[0])
            except (ArithmeticError, ValueError):  # e.g. math.log(0): G(0) undefined
                g0 = math.nan
        self.g0 = float(g0)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().kf_ktable_create(self.coef.ctypes.data_as(ctypes.c_void_p), self.nseg, self.deg,
                                               self.u_min, du, self.g0, ctypes.byref(h)))
        self._h = h
        self._destroy = _lib.lib().kf_ktable_destroy  # held for __del__ at interpreter exit

    def _g(self, r2):
        v = np.asarray(self.g(r2.reshape(-1)), dtype=np.float64)
        return np.broadcast_to(v, (r2.size,)).reshape(r2.shape)

    @classmethod
    def from_kernel_eval(cls, f, **kw):
        """Table of a reference-style kernel function ``f(sx, sy, tx, ty)`` (scalar, as the
        reference's numba kernels, e.g. kernels/laplace.py:5-8).  Raises if f is not radial."""
        def g(r2):
            return np.array([f(0.0, 0.0, math.sqrt(v), 0.0) for v in np.asarray(r2, dtype=np.float64)])
        rng = np.random.default_rng(7)
        for _ in range(16):
            sx, sy, tx, ty = rng.uniform(-2.0, 2.0, 4)
            r2 = (tx - sx) ** 2 + (ty - sy) ** 2
            a, b = float(f(sx, sy, tx, ty)), float(g(np.array([r2]))[0])
            if not abs(a - b) <= 1e-12 * max(1.0, abs(a)):
                raise ValueError("kernel %r is not radial: G(s, t) must depend on |t - s| only"
                                 % getattr(f, "__name__", f))
        kw.setdefault("name", getattr(f, "__name__", "Kernel_Eval"))
        t = cls(g, **kw)
        t.g = None  # the sampler (and f) is needed only while the table is fitted
        return t

    def desc(self):
        return _lib.KernelDesc(self.kind, 0.0, 1.0, self._h)

    def key(self):
        return ("Tabulated", id(self))

    def eval_r2(self, r2):
        """The table at squared distances r2 (host evaluation, as the plan's operators use)."""
        r2 = np.ascontiguousarray(r2, dtype=np.float64)
        out = np.empty_like(r2)
        _lib.check(_lib.lib().kf_ktable_eval(self._h, r2.ctypes.data_as(ctypes.c_void_p), r2.size,
                                             out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._destroy(h)
            self._h = None

    def __repr__(self):
        return "Tabulated(%s, du=%g, deg=%d, r2 in [2^%g, 2^%g])" % (self.name, self.du, self.deg,
                                                                    self.u_min, self.u_max)


class _SampleCache:
    """g evaluated once per sample array: the real and imaginary tables of a complex kernel are
    fitted on the same grid, in the same order."""

    def __init__(self, g):
        self.g = g
        self.memo = {}

    def __call__(self, r2):
        a = np.ascontiguousarray(r2, dtype=np.float64)
        key = (a.shape, a.tobytes())
        v = self.memo.get(key)
        if v is None:
            with np.errstate(all="ignore"):
                v = np.asarray(self.g(a), dtype=np.complex128)
            self.memo[key] = v
        return v


class ComplexTabulated(KernelDescriptor):
    """A user's complex-valued radial kernel G(r^2) = Gr + i Gi, tabulated for the GPU as two
    tables on one grid (``Tabulated``; the plan evaluates both parts from one log).  This is
    class_fmm's complex FMM (class_fmm.py:167, 174: ``iscomplex``) with a complex Kernel_Eval,
    e.g. the Helmholtz kernel 0.25j H0(k r) of examples/hankel.py:31-39.

    ``g`` maps squared distances to complex values (vectorised).  The table stops at
    ``u_max`` (log2 r^2; default 8, r <= 16): an oscillating kernel cannot be fitted on the
    real tables' default range, and a plan needs (2 W)^2, W the root box width.  Plans of a
    complex kernel take real densities and p <= 32 and give complex potentials; direct sums and
    kernel forms run the two parts as real kernels."""
    kind = _lib.KF_TABLE
    iscomplex = True

    def __init__(self, g, g0=None, u_min=None, u_max=8.0, du=0.25, deg=None, tol=1e-13, name=None):
        cache = _SampleCache(g)
        self.name = name or getattr(g, "__name__", "g")
        g0r = g0i = None
        if g0 is not None:
            g0r, g0i = complex(g0).real, complex(g0).imag
        kw = dict(u_min=u_min, u_max=u_max, du=du, deg=deg, tol=tol)
        self.re = Tabulated(lambda r2: cache(r2).real, g0=g0r, name=self.name + ".re", **kw)
        self.im = Tabulated(lambda r2: cache(r2).imag, g0=g0i, name=self.name + ".im", **kw)
        self.re.g = self.im.g = None  # the samples are needed only while the tables are fitted
        self.g = g

    def parts(self):
        """(real part, imaginary part) as real kernel descriptors."""
        return self.re, self.im

    def desc(self):
        return _lib.KernelDesc(self.kind, 0.0, 1.0, self.re._h, self.im._h)

    def key(self):
        return ("ComplexTabulated", id(self))

    def eval_r2(self, r2):
        return self.re.eval_r2(r2) + 1j * self.im.eval_r2(r2)

    @classmethod
    def from_kernel_eval(cls, f, **kw):
        """Tables of a complex reference-style kernel function ``f(sx, sy, tx, ty)`` (scalar,
        as examples/hankel.py:37-39).  Raises if f is not radial."""
        def g(r2):
            return np.array([complex(f(0.0, 0.0, math.sqrt(v), 0.0)) for v in np.asarray(r2, dtype=np.float64)])
        rng = np.random.default_rng(7)
        for _ in range(16):
            sx, sy, tx, ty = rng.uniform(-2.0, 2.0, 4)
            r2 = (tx - sx) ** 2 + (ty - sy) ** 2
            a, b = complex(f(sx, sy, tx, ty)), complex(g(np.array([r2]))[0])
            if not abs(a - b) <= 1e-12 * max(1.0, abs(a)):
                raise ValueError("kernel %r is not radial: G(s, t) must depend on |t - s| only"
                                 % getattr(f, "__name__", f))
        kw.setdefault("name", getattr(f, "__name__", "Kernel_Eval"))
        t = cls(g, **kw)
        t.g = None
        return t

    def __repr__(self):
        return "ComplexTabulated(%s, du=%g, deg=%d/%d, r2 in [2^%g, 2^%g])" % (
            self.name, self.re.du, self.re.deg, self.im.deg, self.re.u_min, self.re.u_max)


def hankel_h0(k):
    """G(r) = 0.25j H0^(1)(k r), the Helmholtz kernel of examples/hankel.py:31-39 (scipy's hankel1
    sampled into the tables)."""
    import scipy.special

    def g(r2):
        return 0.25j * scipy.special.hankel1(0, k * np.sqrt(r2))
    return ComplexTabulated(g, name="0.25j*H0(%g r)" % k)


def _returns_complex(f):
    try:
        v = f(0.0, 0.0, 1.0, 0.0)
    except Exception:
        return False
    return isinstance(v, complex) or np.iscomplexobj(v)


laplace_eval = Laplace()


def ood2_eval(sx, sy, tx, ty):
    """G = 1 / |t - s|^2, the reference's kernels/one_over_d2.py:5-8, as the scalar function it
    is there: handed to the factories, it is tabulated like any radial Kernel_Eval."""
    return 1.0 / ((tx - sx) ** 2 + (ty - sy) ** 2)


def generate_modified_helmholtz_functions(k, scale=1.0):
    """Descriptor for the reference module's kernel (kernels/modified_helmholtz.py:6-9)."""
    return ModifiedHelmholtz(k, scale)


# tables of Kernel_Eval functions by the identity of the function (for a bound method: of its
# function and its object, since each attribute access makes a new method object); an entry
# lives as long as those objects, or for the process if one of them cannot be weakly referenced
_TABLES = {}


def _identity(kernel):
    func, obj = getattr(kernel, "__func__", None), getattr(kernel, "__self__", None)
    if func is not None and obj is not None:
        return (id(func), id(obj)), (func, obj)
    return (id(kernel),), (kernel,)


def as_descriptor(kernel):
    """A descriptor for ``kernel``: descriptors pass through; a reference-style
    ``Kernel_Eval(sx, sy, tx, ty)`` callable is tabulated once (``Tabulated.from_kernel_eval``)
    and the table reused for later calls with the same function (the same bound method
    included) while it lives."""
    if isinstance(kernel, KernelDescriptor):
        return kernel
    if callable(kernel):
        key, owners = _identity(kernel)
        hit = _TABLES.get(key)
        if hit is not None:
            return hit[1]
        t = (ComplexTabulated if _returns_complex(kernel) else Tabulated).from_kernel_eval(kernel)
        try:
            for o in owners:
                weakref.ref(o)
        except TypeError:  # held strongly, so the ids in the key stay theirs
            _TABLES[key] = (owners, t)
            return t
        for o in owners:
            weakref.finalize(o, _TABLES.pop, key, None)
        _TABLES[key] = (None, t)
        return t
    raise TypeError("kernel must be a pykifmm2d_amd.kernels descriptor or a radial "
                    "Kernel_Eval(sx, sy, tx, ty) function")
