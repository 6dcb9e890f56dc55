"""The box structure built on the GPU (tree_dev.hip) against the oracle's restatement of tree.py and
against the library's host builder (tree_host.cpp, from the same sorted keys).

Bars: every exported level array bit-exact (box order, point ranges, parent / children, leaf /
fake / X-list flags, colleague slots), the same per-level sizes; batch builds equal to single
builds tree by tree.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle.tree_ref import Tree as RefTree
from pykifmm2d_amd import _lib, workloads
from pykifmm2d_amd.tree import Tree, export_levels
from tests.treekeys import sorted_keys

FIELDS = ["xmin", "ymin", "bot_ind", "top_ind", "parent_ind", "children_ind", "leaf", "fake_leaf",
          "Xlist", "colleagues"]


def host_levels(x, y, cutoff, D):
    keys, order, mmin, W = sorted_keys(x, y, D)
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().kf_tree_from_sorted_keys(keys.ctypes.data_as(ctypes.c_void_p), keys.size, cutoff,
                                                   mmin, W, D, ctypes.byref(h)))
    try:
        return export_levels(h, W)
    finally:
        _lib.lib().kf_tree_destroy(h)


def same_levels(a_levels, b_levels, what):
    assert len(a_levels) == len(b_levels), what + ": levels"
    for l, (a, b) in enumerate(zip(a_levels, b_levels)):
        assert a.n_node == b.n_node, "%s: level %d size %d vs %d" % (what, l, a.n_node, b.n_node)
        for f in FIELDS + ["depths"]:
            np.testing.assert_array_equal(np.asarray(getattr(a, f)), np.asarray(getattr(b, f)),
                                          err_msg="%s: level %d %s" % (what, l, f))


def case_points(case):
    if case == "uniform":
        x, y, _ = workloads.uniform(30000, seed=11)
        return x, y, 40
    if case == "clustered":
        x, y, _ = workloads.clustered(60000, seed=12)
        return x, y, 32
    if case == "clustered_deep":
        x, y, _ = workloads.clustered(40000, seed=13, n_clusters=3, frac=0.95)
        return x, y, 8
    if case == "annulus":
        x, y, _ = workloads.annulus(30000, seed=14)
        return x, y, 20
    if case == "curve":
        x, y = workloads.star_curve(20000, 0.3)
        return x, y, 16
    if case == "grid":
        g = np.arange(0, 65) / 64.0
        x, y = [a.ravel() for a in np.meshgrid(g, g)]
        return x, y, 3
    if case == "tiny":
        x, y, _ = workloads.uniform(70, seed=15)
        return x, y, 64
    raise ValueError(case)


@pytest.mark.parametrize("case", ["uniform", "clustered", "clustered_deep", "annulus", "curve", "grid", "tiny"])
def test_device_tree_matches_oracle(case):
    x, y, c = case_points(case)
    print("device tree vs oracle:", case, x.size, c, flush=True)
    t = Tree(x, y, c)
    r = RefTree(x, y, c)
    same_levels(t.Levels, r.Levels, case)
    ordv = t.ordv
    for lv, rv in zip(t.Levels, r.Levels):
        for i in np.nonzero(lv.leaf & (lv.ns > 0))[0][:200]:
            assert set(ordv[lv.bot_ind[i]:lv.top_ind[i]]) == set(r.ordv[rv.bot_ind[i]:rv.top_ind[i]])


def test_device_tree_1m_clustered_matches_host_builder():
    x, y, _ = workloads.clustered(1_000_000, seed=2)
    print("device tree vs host builder: 1M clustered, cutoff 64", flush=True)
    t = Tree(x, y, 64)
    assert t.key_depth == 30
    same_levels(t.Levels, host_levels(x, y, 64, t.key_depth), "1M clustered")


def test_batch_build_equals_single_builds():
    sets = [case_points(c)[:2] for c in ("uniform", "clustered_deep", "curve", "tiny")]
    sets.append(workloads.uniform(5000, seed=16)[:2])
    print("batch build of %d trees" % len(sets), flush=True)
    trees = Tree.batch(sets, 24)
    for k, ((x, y), t) in enumerate(zip(sets, trees)):
        single = Tree(x, y, 24)
        same_levels(t.Levels, single.Levels, "tree %d" % k)
        np.testing.assert_array_equal(t.ordv, single.ordv)


def check_lists(handle):
    _lib.check(_lib.lib().kf_plan_verify(handle, None))


@pytest.mark.parametrize("case,kern,p", [("uniform", "laplace", 16), ("clustered_deep", "laplace", 64),
                                         ("annulus", "mhk", 48), ("curve", "table", 32), ("grid", "laplace", 24)])
def test_device_plan_lists_match_host_builder(case, kern, p):
    """Every work list of a plan (slab rows, upward / downward / leaf lists, interaction-list
    rows, statistics), built on the GPU, equals the host builder's from the same sorted keys, and
    its operators the host's (kf_plan_verify)."""
    from pykifmm2d_amd import fmm, kernels
    K = {"laplace": kernels.Laplace(), "mhk": kernels.ModifiedHelmholtz(10.0),
         "table": kernels.Tabulated(lambda r2: -0.25 / np.pi * np.log(r2))}[kern]
    x, y, c = case_points(case)
    print("device plan lists vs host:", case, kern, p, flush=True)
    plan = fmm._Plan(Tree(x, y, c), K, p)
    check_lists(plan.handle)
    plan32 = fmm._Plan(Tree(x, y, c), K, p, _lib.KF_PREC_SPLIT)
    check_lists(plan32.handle)


def test_device_plan_lists_1m_and_forest():
    from pykifmm2d_amd import fmm, kernels
    from pykifmm2d_amd.forest import Forest
    x, y, _ = workloads.clustered(1_000_000, seed=2)
    print("device plan lists vs host: 1M clustered", flush=True)
    plan = fmm._Plan(Tree(x, y, 64), kernels.Laplace(), 64)
    check_lists(plan.handle)
    sets = [case_points(c)[:2] for c in ("uniform", "clustered_deep", "curve")]
    sets.append(workloads.uniform(5000, seed=17)[:2])
    print("device plan lists vs host: forest of %d trees" % len(sets), flush=True)
    F = Forest(sets, kernels.Laplace(), p=16, cutoff=24)
    check_lists(F.handle)
    F2 = Forest(sets, kernels.Laplace(), p=16, cutoff=24, batch_build=False)  # trees of separate builds
    check_lists(F2.handle)


@pytest.mark.parametrize("case", ["clustered_deep", "curve", "grid"])
def test_locate_points_matches_oracle(case):
    """Tree.locate_points / locate_point (tree.py:686-727) on the GPU: the same (level, position)
    as the oracle's walk, for random points, the tree's own points and points on midlines."""
    x, y, c = case_points(case)
    t = Tree(x, y, c)
    r = RefTree(x, y, c)
    rng = np.random.default_rng(21)
    lo, hi = float(t.xmin), float(t.xmax)
    qx = np.concatenate([rng.uniform(lo, hi, 5000), x[:3000], lo + (hi - lo) * rng.integers(0, 65, 500) / 64.0])
    qy = np.concatenate([rng.uniform(lo, hi, 5000), y[:3000], lo + (hi - lo) * rng.integers(0, 65, 500) / 64.0])
    inds, locs = t.locate_points(qx, qy)
    rinds, rlocs = r.locate_points(qx, qy)
    np.testing.assert_array_equal(inds, rinds)
    np.testing.assert_array_equal(locs, rlocs)
    assert t.locate_point(qx[7], qy[7]) == (int(rinds[7]), int(rlocs[7]))
    assert inds.shape == qx.shape
    assert not t.check_bounding_box(qx, qy)  # the reference's formula, true only for empty input
    assert t.check_bounding_box(np.empty(0), np.empty(0))


def test_pool_overflow_retries_on_star_curves(monkeypatch):
    """Points on star curves need up to 5.5x the boxes of the split-by-count phase; with the
    first box pool forced to B1 + 4096 (KF_TREE_POOL=1) the level restriction overflows it and
    the build retries with larger pools.  The trees must equal the host builder's, single and
    batched, and a plan over them must pass its list and operator checks."""
    from pykifmm2d_amd import fmm, kernels
    curves = workloads.star_curves(64, seed=5)
    pick = [curves[43], curves[41], curves[44]]  # 18k, 42k, 20k points
    monkeypatch.setenv("KF_TREE_POOL", "1")
    for k, (x, y) in enumerate(pick):
        print("pool overflow: curve %d, %d points" % (k, x.size), flush=True)
        t = Tree(x, y, 50)
        same_levels(t.Levels, host_levels(x, y, 50, t.key_depth), "curve %d" % k)
        plan = fmm._Plan(t, kernels.Laplace(), 48)  # held: the handle lives as long as the plan
        check_lists(plan.handle)
    trees = Tree.batch(pick, 50)
    for k, ((x, y), t) in enumerate(zip(pick, trees)):
        same_levels(t.Levels, host_levels(x, y, 50, t.key_depth), "batched curve %d" % k)
