"""The C-ABI library loads on a machine without a GPU and exports every declared symbol."""
import ctypes
import os
import re

from pykifmm2d_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "kifmm2d.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kf_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported_and_bound():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, "no ctypes signature for %s" % n
    assert set(_lib.SIGNATURES) == set(names)


def test_version_and_host_k0():
    import scipy.special
    L = _lib.lib()
    assert L.kf_version() == 100
    for x in (1e-8, 0.3, 1.0, 1.7, 2.5, 9.0, 40.0):
        assert abs(L.kf_k0(x) - scipy.special.k0(x)) <= 2e-15 * scipy.special.k0(x)


def test_error_path_sets_message():
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.kf_tree_from_sorted_keys(None, 0, 10, 0.0, 1.0, 30, ctypes.byref(h))
    assert rc != 0
    assert L.kf_last_error()
