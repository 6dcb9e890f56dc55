"""Tabulated user kernels (KF_TABLE) on the host: the table format through the C-ABI, the fit of
``kernels.Tabulated`` against closed forms, reference-style kernel functions, and the oracle's
numpy path for user kernels.  No GPU: creating and evaluating a table make no HIP call."""
import ctypes
import math

import numpy as np
import pytest
import scipy.special

from oracle import fmm_ref
from pykifmm2d_amd import _lib, kernels

L = _lib.lib()


def make_table(coef, nseg, deg, u0, du, g0):
    c = np.ascontiguousarray(coef, dtype=np.float64)
    h = ctypes.c_void_p()
    rc = L.kf_ktable_create(c.ctypes.data_as(ctypes.c_void_p), nseg, deg, u0, du, g0, ctypes.byref(h))
    return rc, h


def eval_table(h, r2):
    r2 = np.ascontiguousarray(r2, dtype=np.float64)
    out = np.empty_like(r2)
    _lib.check(L.kf_ktable_eval(h, r2.ctypes.data_as(ctypes.c_void_p), r2.size, out.ctypes.data_as(ctypes.c_void_p)))
    return out


def test_table_format_known_polynomials():
    """Three quadratics on u in [-4, 0), [0, 4), [4, 8): at r^2 = 2^u for integer u the local
    x = 2 ((u + 4) / 4 - s) - 1 is exact, so each value is the quadratic's closed form."""
    coef = [[1.0, 2.0, 3.0], [-1.0, 0.5, 0.25], [7.0, -3.0, 1.0]]
    rc, h = make_table(coef, 3, 2, -4.0, 4.0, 123.0)
    assert rc == _lib.KF_OK
    try:
        def want(u):
            if u < -4:
                s, x = 0, -1.0
            elif u >= 8:
                s, x = 2, 1.0
            else:
                y = (u + 4) / 4.0
                s = int(math.floor(y))
                x = 2.0 * (y - s) - 1.0
            a, b, c = coef[s]
            return a + b * x + c * x * x
        us = list(range(-9, 13))
        got = eval_table(h, np.exp2(np.array(us, dtype=np.float64)))
        np.testing.assert_array_equal(got, [want(u) for u in us])
        # G(0) = g0; tiny and huge r^2 take the end values
        np.testing.assert_array_equal(eval_table(h, np.array([0.0, 1e-300, 1e300])),
                                      [123.0, want(-100), want(100)])
    finally:
        L.kf_ktable_destroy(h)


@pytest.mark.parametrize("args", [
    dict(nseg=0), dict(deg=-1), dict(deg=25), dict(du=0.3), dict(du=32.0), dict(du=2.0 ** -11),
    dict(u0=0.25), dict(u0=float("nan")), dict(coef_bad=float("inf"))])
def test_table_create_rejects_bad_arguments(args):
    a = dict(nseg=2, deg=1, u0=-1.0, du=0.5, g0=0.0)
    bad = args.pop("coef_bad", None)
    a.update(args)
    n = max(a["nseg"], 1) * (max(a["deg"], 0) + 1)
    coef = np.ones(n)
    if bad is not None:
        coef[-1] = bad
    rc, h = make_table(coef, a["nseg"], a["deg"], a["u0"], a["du"], a["g0"])
    assert rc == -1 and not h.value  # KF_ERR_ARG
    assert L.kf_last_error()


def lap_g(r2):
    return (-0.25 / np.pi) * np.log(r2)


def k0_g(k, scale):
    return lambda r2: scale * scipy.special.k0(k * np.sqrt(r2))


def test_tabulated_laplace_matches_closed_form():
    T = kernels.Tabulated(lap_g)
    assert T.fit_error < 1e-14
    r2 = np.exp2(np.random.default_rng(1).uniform(-80.0, 40.0, 200000))
    g = lap_g(r2)
    # relative to |G|, or to 0.01 (|G| at r^2 = 2^0.18) where G crosses zero at r = 1
    err = np.abs(T.eval_r2(r2) - g) / np.maximum(np.abs(g), 1e-2)
    print("laplace table max err", err.max())
    assert err.max() < 1e-14


@pytest.mark.parametrize("k", [0.5, 10.0, 300.0])
def test_tabulated_k0_matches_scipy(k):
    scale = 1.0 / (2.0 * np.pi)
    T = kernels.Tabulated(k0_g(k, scale))
    assert T.fit_error < 5e-14
    r = np.exp(np.random.default_rng(2).uniform(np.log(1e-12 / k), np.log(40.0 / k), 200000))
    g = k0_g(k, scale)(r * r)
    got = T.eval_r2(r * r)
    # relative to |G|, or to 2^-52 |G| a factor 2^10 closer where K0 has fallen below that
    # (e^-(k r): 1e-13 of its value at k r = 1 by k r = 30), as the fit is checked
    err = np.abs(got - g) / np.abs(g)
    print("k0 table max rel err (k r <= 10)", err[k * r <= 10.0].max())
    assert err[k * r <= 10.0].max() < 2e-14
    # (where K0 falls by e^-(0.19 k r) across one segment, the monomial evaluation loses up to
    # that factor in ulps at the segment's low end: 5e-14 at k r = 22, 1e-28 absolute at 32)
    ref = np.maximum(np.abs(g), 2.0 ** -52 * k0_g(k, scale)((r / 1024.0) ** 2))
    print("k0 table max err, floored scale", (np.abs(got - g) / ref).max())
    assert (np.abs(got - g) / ref)[k * r <= 25.0].max() < 1e-13
    assert (np.abs(got - g) / ref).max() < 1e-12


def test_tabulated_rejects_poor_fit_and_bad_range():
    with pytest.raises(ValueError, match="fit error"):
        kernels.Tabulated(lambda r2: np.cos(40.0 * np.log(r2)), deg=4)
    with pytest.raises(ValueError, match="power of two"):
        kernels.Tabulated(lap_g, du=0.3)
    # a default end drops the segments where g is not finite; an explicit one refuses them, and so
    # does any range with a non-finite value inside
    step = lambda r2: np.where(r2 > 2.0 ** 100, np.inf, 1.0)
    assert kernels.Tabulated(step).u_max == 100.0
    with pytest.raises(ValueError, match="not finite"):
        kernels.Tabulated(step, u_max=256.0)
    assert kernels.Tabulated(step, u_max=100.0).fit_error < 1e-14
    with pytest.raises(ValueError, match="not finite"):
        kernels.Tabulated(lambda r2: np.where(np.abs(np.log2(r2) - 3.1) < 0.1, np.nan, 1.0))


def reference_style_laplace(sx, sy, tx, ty):
    # the reference's kernels/laplace.py:5-8 form: a scalar function of two points
    dx = tx - sx
    dy = ty - sy
    return -math.log(dx * dx + dy * dy) * 0.25 / math.pi


def test_reference_style_kernel_function():
    D = kernels.as_descriptor(reference_style_laplace)
    assert isinstance(D, kernels.Tabulated)
    assert kernels.as_descriptor(reference_style_laplace) is D  # tabulated once
    r2 = np.exp2(np.linspace(-60.0, 20.0, 5001))
    np.testing.assert_allclose(D.eval_r2(r2), lap_g(r2), rtol=2e-14, atol=1e-16)
    assert math.isnan(D.g0)  # math.log(0) raises: G(0) undefined
    assert kernels.Tabulated(lap_g).g0 == np.inf  # numpy's log 0 = -inf, times -1/(4 pi)

    def skewed(sx, sy, tx, ty):
        return math.log((tx - sx) ** 2 + 2.0 * (ty - sy) ** 2)
    with pytest.raises(ValueError, match="not radial"):
        kernels.as_descriptor(skewed)
    with pytest.raises(TypeError):
        kernels.as_descriptor(3.0)
    # the cache holds a table only while its function lives
    import gc

    def short_lived(sx, sy, tx, ty):
        return math.log((tx - sx) ** 2 + (ty - sy) ** 2)
    kernels.as_descriptor(short_lived)
    n = len(kernels._TABLES)
    del short_lived
    gc.collect()
    assert len(kernels._TABLES) == n - 1


class _BoundKernel:
    def ev(self, sx, sy, tx, ty):
        return reference_style_laplace(sx, sy, tx, ty)


class _SlotKernel:
    __slots__ = ()  # no __weakref__: the instance cannot be weakly referenced

    def __call__(self, sx, sy, tx, ty):
        return reference_style_laplace(sx, sy, tx, ty)


def test_kernel_functions_from_bound_methods_and_slot_callables():
    """The factories tabulate a Kernel_Eval once per function identity: a bound method (a new
    object on each attribute access) and a callable that cannot be weakly referenced give the
    same table to Get_Kernel_Functions and prepare_numba_functions_*, which kernel_of needs."""
    import gc

    from pykifmm2d_amd import fmm
    obj, slot = _BoundKernel(), _SlotKernel()
    for f in (obj.ev, slot):
        kf = fmm.Get_Kernel_Functions(f)
        for prep in (fmm.prepare_numba_functions_on_the_fly, fmm.prepare_numba_functions_planned):
            nf = prep(f)
            d = fmm.kernel_of(kf, nf)
            assert isinstance(d, kernels.Tabulated)
            assert d is kernels.as_descriptor(f)
    n = len(kernels._TABLES)
    del obj, kf, nf, d
    gc.collect()
    assert len(kernels._TABLES) == n - 1  # the bound method's entry goes with its object


def test_descriptor_struct_matches_header():
    """kf_kernel_desc: int32 kind, double k, double scale, const kf_ktable* table, table_im."""
    assert [f for f, _ in _lib.KernelDesc._fields_] == ["kind", "k", "scale", "table", "table_im"]
    assert ctypes.sizeof(_lib.KernelDesc) == 40
    assert _lib.KernelDesc.table.offset == 24 and _lib.KernelDesc.table_im.offset == 32
    d = kernels.Laplace().desc()
    assert d.kind == _lib.KF_LAPLACE and not d.table and not d.table_im
    T = kernels.Tabulated(lap_g)
    assert T.desc().kind == _lib.KF_TABLE and T.desc().table == T._h.value and not T.desc().table_im


def test_complex_table_of_hankel_h0():
    """The complex kernel 0.25j H0(k r) of examples/hankel.py:31-39 as two tables: both parts
    within 2e-14 of scipy's hankel1 (relative to |G|) from r = 1e-8 to 16, G(0) not finite; the
    descriptor carries both tables; a complex scalar Kernel_Eval handed to the factories is
    tabulated the same way (and a real one is refused as complex)."""
    import scipy.special
    K = kernels.hankel_h0(2.0)
    assert K.iscomplex and K.re.u_max == 8.0 and K.im.u_max == 8.0
    rng = np.random.default_rng(5)
    r = np.concatenate([np.exp(rng.uniform(np.log(1e-8), np.log(16.0), 4000)), [1e-8, 1.0, 16.0]])
    ref = 0.25j * scipy.special.hankel1(0, 2.0 * r)
    got = K.eval_r2(r * r)
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 2e-14
    d = K.desc()
    assert d.kind == _lib.KF_TABLE and d.table == K.re._h.value and d.table_im == K.im._h.value

    def h0_eval(sx, sy, tx, ty):
        return 0.25j * scipy.special.hankel1(0, 2.0 * math.sqrt((tx - sx) ** 2 + (ty - sy) ** 2))
    D = kernels.as_descriptor(h0_eval)
    assert isinstance(D, kernels.ComplexTabulated) and D is kernels.as_descriptor(h0_eval)
    assert np.max(np.abs(D.eval_r2(r * r) - got)) == 0.0
    assert not isinstance(kernels.as_descriptor(lambda sx, sy, tx, ty: 1.0 + tx * 0.0), kernels.ComplexTabulated)


def test_oracle_user_kernel_direct_matches_c_loops():
    """The oracle's numpy direct sum for user kernels against its C loops (built-in kernels)."""
    rng = np.random.default_rng(3)
    x, y, tau = rng.random(1500), rng.random(1500), rng.standard_normal(1500)
    tx, ty = rng.random(300), rng.random(300)
    for g, ref in [(lap_g, fmm_ref.LAPLACE), (k0_g(7.0, 0.3), fmm_ref.modified_helmholtz(7.0, 0.3))]:
        U = fmm_ref.user_kernel(g)
        d = fmm_ref.direct(ref, x, y, tau)
        assert np.abs(fmm_ref.direct_np(U, x, y, tau) - d).max() / np.abs(d).max() < 1e-13
        d = fmm_ref.direct(ref, x, y, tau, tx, ty)
        assert np.abs(fmm_ref.direct_np(U, x, y, tau, tx, ty) - d).max() / np.abs(d).max() < 1e-13
    with pytest.raises(ValueError):
        fmm_ref.direct(fmm_ref.user_kernel(lap_g), x, y, tau)


def test_svd_fmm_and_utils_modules():
    """The reference's examples import pykifmm2d.svd_fmm and pykifmm2d.utils (examples/laplace.py:
    1-27); the drop-ins expose the same names (svd_fmm's are fmm's: same exact C2E inverse)."""
    from pykifmm2d_amd import fmm, svd_fmm, utils
    for name in ("on_the_fly_fmm", "fmm_planner", "planned_fmm", "Get_Kernel_Functions",
                 "prepare_numba_functions_on_the_fly", "prepare_numba_functions_planned", "Kernel_Form",
                 "Kernel_Apply", "get_level_information"):
        assert getattr(svd_fmm, name) is getattr(fmm, name)
    np.random.seed(1)
    x, y = utils.random2(1000, -1.0, 3.0)
    assert x.shape == y.shape == (1000,) and x.min() >= -1.0 and y.max() < 3.0 and x.std() > 1.0


def test_one_over_d2_kernel_tabulated():
    """kernels.ood2_eval (reference kernels/one_over_d2.py:5-8) through the factories' tabulation."""
    D = kernels.as_descriptor(kernels.ood2_eval)
    assert isinstance(D, kernels.Tabulated) and math.isnan(D.g0)
    r2 = np.exp2(np.random.default_rng(4).uniform(-40.0, 20.0, 100000))
    got = D.eval_r2(r2)
    assert np.abs(got * r2 - 1.0).max() < 5e-14
