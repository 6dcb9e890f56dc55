"""Box-level tree logic (host part of libkifmm2d) vs the oracle's restatement of tree.py.

The library builds levels / colleagues / level restriction / X-list from sorted Morton keys;
here the keys come from numpy (tests/treekeys.py) so the check runs without a GPU.  The
oracle builds its tree with the reference's in-place cycle partition (tree.py:86-190), so
point ranges, level order, flags and colleague slots must agree exactly.
"""
import ctypes

import numpy as np
import pytest

from oracle.tree_ref import Tree as RefTree
from pykifmm2d_amd import _lib, workloads
from pykifmm2d_amd.tree import export_levels
from tests.treekeys import sorted_keys

FIELDS = ["xmin", "ymin", "bot_ind", "top_ind", "parent_ind", "children_ind", "leaf", "fake_leaf",
          "Xlist", "colleagues"]


def host_tree(x, y, cutoff, D=30):
    keys, order, mmin, W = sorted_keys(x, y, D)
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().kf_tree_from_sorted_keys(keys.ctypes.data_as(ctypes.c_void_p), keys.size, cutoff,
                                                   mmin, W, D, ctypes.byref(h)))
    try:
        levels = export_levels(h, W)
    finally:
        _lib.lib().kf_tree_destroy(h)
    return levels, order


def compare(x, y, cutoff):
    mine, order = host_tree(x, y, cutoff)
    ref = RefTree(x, y, cutoff)
    assert len(mine) == ref.levels
    for l, (a, b) in enumerate(zip(mine, ref.Levels)):
        assert a.n_node == b.n_node, "level %d size" % l
        for f in FIELDS + ["depths"]:
            np.testing.assert_array_equal(np.asarray(getattr(a, f)), np.asarray(getattr(b, f)),
                                          err_msg="level %d field %s" % (l, f))
        if l > 0:
            for f in ("compute_upwards", "this_density_ind", "parent_density_ind"):
                np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg="level %d %s" % (l, f))
    # leaf membership: the same point sets per leaf (within-leaf order may differ)
    for l, (a, b) in enumerate(zip(mine, ref.Levels)):
        for i in np.nonzero(a.leaf & (a.ns > 0))[0][:200]:
            s1 = set(order[a.bot_ind[i]:a.top_ind[i]].tolist())
            s2 = set(ref.ordv[b.bot_ind[i]:b.top_ind[i]].tolist())
            assert s1 == s2
    return mine


@pytest.mark.parametrize("n,cutoff", [(10000, 50), (3000, 10), (257, 64)])
def test_uniform_matches_oracle(n, cutoff):
    x, y, _ = workloads.uniform(n, seed=1)
    compare(x, y, cutoff)


def test_clustered_matches_oracle():
    x, y, _ = workloads.clustered(60000, seed=2)
    levels = compare(x, y, 64)
    assert len(levels) >= 8
    assert any(lv.fake_leaf.any() for lv in levels)  # X-list fake children occur


def test_annulus_and_curve_match_oracle():
    x, y, _ = workloads.annulus(20000, seed=3)
    compare(x, y, 64)
    cx, cy = workloads.star_curve(20000, 0.7)
    compare(cx, cy, 50)


def test_points_on_midlines_and_box_edges():
    # grid points exactly on dyadic midlines (strict ">" sends them to the low child) and
    # points at x = mmax (classified high)
    g = np.arange(0, 33) / 32.0
    x, y = np.meshgrid(g, g)
    compare(x.ravel(), y.ravel(), 4)


def test_small_and_single_level():
    x, y, _ = workloads.uniform(40, seed=5)
    levels = compare(x, y, 50)
    assert len(levels) == 1
    levels = compare(x, y, 3)
    assert len(levels) > 2


def test_level_restriction_invariant():
    """No leaf has a colleague with grandchildren (doc.tex:68) and X-list = leaf with a
    non-leaf real colleague, checked directly on the library's tree."""
    x, y, _ = workloads.clustered(30000, seed=7)
    levels, _ = host_tree(x, y, 32)
    depth = [np.zeros(lv.n_node, np.int64) for lv in levels]
    for l in reversed(range(len(levels))):
        lv = levels[l]
        for i in range(lv.n_node):
            if not lv.leaf[i] and not lv.fake_leaf[i]:
                c = lv.children_ind[i]
                depth[l][i] = 1 + depth[l + 1][c:c + 4].max()
    for l, lv in enumerate(levels):
        real = ~lv.fake_leaf
        for i in np.nonzero(lv.leaf & real)[0]:
            cols = [c for c in lv.colleagues[i] if c >= 0 and not lv.fake_leaf[c]]
            assert all(depth[l][c] <= 1 for c in cols)
            if 0 < l < len(levels) - 1:
                assert lv.Xlist[i] == any(depth[l][c] > 0 for c in cols)
