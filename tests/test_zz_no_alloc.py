"""kf_direct_noalloc (the direct sum with its partial sums in the library's static per-device
scratch) against kf_direct, through the C-ABI: bitwise equal on two calls in a row, the second
smaller than the first, and sizes beyond the scratch refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from pykifmm2d_amd import _lib, kernels, workloads


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _both(torch, desc, x, y, tau, nt, self_mode):
    L = _lib.lib()
    dev = torch.device("cuda:0")
    sx = torch.from_numpy(x).to(dev)
    sy = torch.from_numpy(y).to(dev)
    st = torch.from_numpy(tau).to(dev)
    tx, ty = sx[:nt].contiguous(), sy[:nt].contiguous()
    ref = torch.full((nt,), np.nan, dtype=torch.float64, device=dev)
    out = torch.full((nt,), np.nan, dtype=torch.float64, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = (_vp(sx), _vp(sy), _vp(st), x.size, _vp(tx), _vp(ty), None, nt, self_mode)
    _lib.check(L.kf_direct(ctypes.byref(desc), *args, _vp(ref), stream))
    _lib.check(L.kf_direct_noalloc(ctypes.byref(desc), *args, _vp(out), stream))
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy()


@pytest.mark.parametrize("kern", [kernels.Laplace(), kernels.ModifiedHelmholtz(10.0)])
def test_direct_noalloc_matches_direct(kern):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    desc = kern.desc()
    # first call: 70k sources (5 chunks of 16384) at 30k targets; second: smaller, so the
    # scratch holds the first call's partials beyond what the second writes
    for n, nt, mode in ((70000, 30000, 1), (20000, 5000, 2)):
        x, y, tau = workloads.clustered(n, seed=n)
        ref, out = _both(torch, desc, x, y, tau, nt, mode)
        assert np.isfinite(ref).all()
        np.testing.assert_array_equal(out, ref)


def test_direct_noalloc_refuses_sizes_beyond_its_scratch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    L = _lib.lib()
    desc = kernels.Laplace().desc()
    dev = torch.device("cuda:0")
    # 2 * nt * ceil(ns / 16384) = 2 * 1,100,000 * 2 > 4,194,304: refused before any launch
    ns, nt = 20000, 1_100_000
    s = torch.zeros(ns, dtype=torch.float64, device=dev)
    t = torch.zeros(nt, dtype=torch.float64, device=dev)
    out = torch.zeros(nt, dtype=torch.float64, device=dev)
    rc = L.kf_direct_noalloc(ctypes.byref(desc), _vp(s), _vp(s), _vp(s), ns, _vp(t), _vp(t), None, nt, 0,
                             _vp(out), None)
    assert rc == -1  # KF_ERR_ARG
    assert b"scratch" in L.kf_last_error()
