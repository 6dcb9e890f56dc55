"""GPU tests of forests: many independent trees planned and evaluated in shared launches.

Bar: every tree's potentials in the forest are bitwise equal to the tree's own single-tree plan
(same operators, same per-right-hand-side arithmetic; only the launch and tile composition
differ), and within 1e-12 of the oracle; split precision within 2e-6 of the direct sum (FP32
M2L on uniform points at p = 48 measures 1.3e-6; the clustered, annulus and curve workloads
stay below 1e-6).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import fmm_ref
from pykifmm2d_amd import fmm, forest, kernels, workloads


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _points(seed):
    """Uneven trees: uniform in the unit box (several sizes), clustered, a star curve (root
    [-2,2], a second width class), points in [0,2]^2 (a third) and one tree small enough for
    the direct-sum path."""
    rng = np.random.default_rng(seed)
    pts = []
    for n in (1000, 3500, 12000, 30000, 2200):
        pts.append((rng.random(n), rng.random(n)))
    x, y, _ = workloads.clustered(15000, seed=seed)
    pts.append((x, y))
    pts.append(workloads.star_curve(9000, 0.7))
    pts.append((2 * rng.random(5000), 2 * rng.random(5000)))
    pts.append((rng.random(40), rng.random(40)))
    return pts


@pytest.mark.parametrize("kern,ref,p", [(kernels.Laplace(), fmm_ref.LAPLACE, 32),
                                        (kernels.ModifiedHelmholtz(10.0), fmm_ref.modified_helmholtz(10.0), 48)])
def test_forest_equals_single_tree_plans(kern, ref, p):
    pts = _points(41)
    F = forest.Forest(pts, kern, p=p, cutoff=50)
    assert F.direct == [len(pts) - 1] and len(F.planned) == len(pts) - 1
    rng = np.random.default_rng(5)
    taus = [rng.standard_normal(len(x)) for x, _ in pts]
    us = F.eval(taus)
    kf = fmm.Get_Kernel_Functions(kern)
    nf = fmm.prepare_numba_functions_planned(kern)
    for (x, y), tau, u in zip(pts, taus, us):
        plan = fmm.fmm_planner(x, y, p, 50, kf, nf)
        np.testing.assert_array_equal(u, fmm.planned_fmm(plan, tau))
    # one tree of each width class against the oracle
    for i in (3, 6, 7):
        x, y = pts[i]
        uo, _ = fmm_ref.on_the_fly_fmm(x, y, taus[i], p, 50, ref)
        assert rel(us[i], uo) < 1e-12
    # the concatenated form gives the same numbers
    u_cat = F.eval(np.concatenate(taus))
    np.testing.assert_array_equal(u_cat, np.concatenate(us))


def test_forest_batched_rhs_and_split_precision():
    pts = _points(42)[:8]
    rng = np.random.default_rng(6)
    N = sum(len(x) for x, _ in pts)
    T = rng.standard_normal((5, N))
    F = forest.Forest(pts, kernels.Laplace(), p=48, cutoff=50)
    U = F.eval(T)
    assert U.shape == (5, N)
    for r in (0, 4):  # a single evaluation sums the near field in the symmetric form
        assert rel(F.eval(T[r]), U[r]) < 1e-13
    S = forest.Forest(pts, kernels.Laplace(), p=48, cutoff=50, precision="split")
    Us = S.eval(T[:2])
    off = np.concatenate([[0], np.cumsum([len(x) for x, _ in pts])])
    for i in (0, 3, 6):
        x, y = pts[i]
        d = fmm_ref.direct(fmm_ref.LAPLACE, x, y, T[1, off[i]:off[i + 1]])
        e = rel(Us[1, off[i]:off[i + 1]], d)
        print("forest split rel err", i, e)
        assert e < 2e-6


def test_batch_tree_build_matches_single_builds():
    """Trees built together are the trees built one by one: every level array, the points in
    tree order and the ordering vector."""
    from pykifmm2d_amd.tree import Tree
    pts = _points(43)
    batch = Tree.batch(pts, 50)
    for (x, y), tb in zip(pts, batch):
        ts = Tree(x, y, 50)
        assert tb.levels == ts.levels and tb.N == ts.N
        np.testing.assert_array_equal(tb.ordv, ts.ordv)
        np.testing.assert_array_equal(tb.x, ts.x)
        for la, lb in zip(tb.Levels, ts.Levels):
            for f in ("xmin", "ymin", "bot_ind", "top_ind", "parent_ind", "children_ind", "leaf", "fake_leaf",
                      "Xlist", "colleagues"):
                np.testing.assert_array_equal(np.asarray(getattr(la, f)), np.asarray(getattr(lb, f)))
    from pykifmm2d_amd._lib import KifmmError
    with pytest.raises(KifmmError):  # second tree's y range exceeds its root box
        Tree.batch([(np.random.rand(100), np.random.rand(100)), (np.array([0.1, 0.2]), np.array([0.1, 2.5]))], 50)
