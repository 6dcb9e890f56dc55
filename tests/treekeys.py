"""Exact Morton keys in numpy for the tree tests (independent of the device key kernel)."""
import numpy as np


def root_box(x, y):
    """Reference root box rule (tree.py:475-484), ymax taken as y.min()."""
    mmin = float(np.floor(min(x.min(), y.min())))
    mmax = float(np.ceil(max(x.max(), y.min())))
    return mmin, mmax - mmin


def grid_index(v, lo, W, D):
    h = W * 2.0 ** -D
    top = (1 << D) - 1
    m = np.clip(np.ceil((v - lo) / h).astype(np.int64) - 1, 0, top)
    for _ in range(4):
        dec = (m > 0) & ~(v > lo + m * h)
        m = m - dec
        inc = (m < top) & (v > lo + (m + 1) * h)
        m = m + inc
    return m.astype(np.uint64)


def spread(v):
    v = v.astype(np.uint64) & np.uint64(0xFFFFFFFF)
    for s, mask in ((16, 0x0000FFFF0000FFFF), (8, 0x00FF00FF00FF00FF), (4, 0x0F0F0F0F0F0F0F0F),
                    (2, 0x3333333333333333), (1, 0x5555555555555555)):
        v = (v | (v << np.uint64(s))) & np.uint64(mask)
    return v


def sorted_keys(x, y, D=30):
    mmin, W = root_box(x, y)
    k = (spread(grid_index(x, mmin, W, D)) << np.uint64(1)) | spread(grid_index(y, mmin, W, D))
    order = np.argsort(k, kind="stable")
    return np.ascontiguousarray(k[order]), order, mmin, W
