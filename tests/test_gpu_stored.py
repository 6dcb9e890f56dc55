"""GPU tests of the stored kernel matrices (kf_plan_store_matrices through
fmm_planner(store_matrices=True)): near field, P2M and L2P, as the reference's planner keeps them.

The stored plan reads its kernel blocks instead of evaluating the kernel: its potentials agree
with the matrix-free plan to rounding (1e-13, max-norm relative) and with the oracle to the FP64
bar (1e-12); a density gives bitwise the same potential at every batch width and through graph
replay; freeing the matrices restores the matrix-free results bitwise.
"""
import numpy as np
import pytest
import scipy.special

pytestmark = pytest.mark.gpu

from oracle import fmm_ref
from pykifmm2d_amd import fmm, forest, kernels, workloads

LAP = kernels.Laplace()
MHK = kernels.ModifiedHelmholtz(10.0)
K0_T = kernels.Tabulated(lambda r2: scipy.special.k0(10.0 * np.sqrt(r2)) / (2.0 * np.pi), name="k0")


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("gen,kern,ref,p,nrhs", [
    ("clustered", LAP, fmm_ref.LAPLACE, 48, 3),
    ("annulus", MHK, fmm_ref.modified_helmholtz(10.0), 48, 16),
    ("annulus", K0_T, fmm_ref.modified_helmholtz(10.0), 32, 3),
])
def test_stored_near_matches_matrix_free(gen, kern, ref, p, nrhs):
    x, y, _ = getattr(workloads, gen)(15000)
    T = np.random.default_rng(21).standard_normal((nrhs, x.size))
    plan = fmm.fmm_planner(x, y, p, 50, fmm.Get_Kernel_Functions(kern), fmm.prepare_numba_functions_planned(kern))
    u0 = fmm.planned_fmm(plan, T)
    P = plan.plan
    fp0 = P.footprint(nrhs)
    P.store_matrices(True)
    st = P.stats()
    values = st["p2p_pairs"] + x.size + st["p2m_evals"] + st["l2p_evals"]
    # offsets: 2 (leaves + 1) + (P2M boxes + 1), the P2M boxes (leaves and X-list fake
    # children with sources) fewer than the tree's boxes
    offsets = (P.footprint(nrhs) - fp0) // 8 - values
    assert (P.footprint(nrhs) - fp0) % 8 == 0
    assert 2 * (st["leaves"] + 1) + 1 < offsets <= 2 * (st["leaves"] + 1) + 1 + plan.tree.n_boxes
    u1 = fmm.planned_fmm(plan, T)
    print("stored vs matrix-free", rel(u1, u0))
    assert rel(u1, u0) < 1e-13
    for r in (0, nrhs - 1):
        np.testing.assert_array_equal(u1[r], fmm.planned_fmm(plan, T[r]))
    uo, _ = fmm_ref.on_the_fly_fmm(x, y, T[0], p, 50, ref)
    assert rel(u1[0], uo) < 1e-12
    # graph replay reads the same matrix
    P.set_graphs(True)
    for _ in range(2):
        np.testing.assert_array_equal(fmm.planned_fmm(plan, T), u1)
    P.set_graphs(False)
    # the launch structures all read it the same way
    P.set_tuning(near_grid=256, near_fork=1, near_split=300)
    np.testing.assert_array_equal(fmm.planned_fmm(plan, T), u1)
    P.set_tuning()
    P.store_matrices(False)
    assert P.footprint(nrhs) == fp0
    np.testing.assert_array_equal(fmm.planned_fmm(plan, T), u0)


def test_fmm_planner_store_matrices_keyword():
    """fmm_planner(store_matrices=True) plans with the stored matrices from the start."""
    x, y, tau = workloads.clustered(12000, seed=9)
    plan = fmm.fmm_planner(x, y, 48, 50, fmm.Get_Kernel_Functions(LAP), fmm.prepare_numba_functions_planned(LAP),
                           store_matrices=True)
    st = plan.plan.stats()
    base = plan.plan.footprint(1)
    u = fmm.planned_fmm(plan, tau)
    uo, _ = fmm_ref.on_the_fly_fmm(x, y, tau, 48, 50, fmm_ref.LAPLACE)
    assert rel(u, uo) < 1e-12
    plan.plan.store_matrices(False)
    assert (base - plan.plan.footprint(1)) // 8 > st["p2p_pairs"] + x.size + st["p2m_evals"] + st["l2p_evals"]


def test_forest_stored_matrices():
    """A forest plan stores its trees' matrices in one set of blocks: within rounding of the
    matrix-free forest, bitwise the same per density at every width."""
    rng = np.random.default_rng(5)
    pts = [(rng.random(n), rng.random(n)) for n in (2500, 9000, 20000)]
    pts.append(workloads.star_curve(7000, 0.7))
    pts.append((rng.random(50), rng.random(50)))
    N = sum(len(x) for x, _ in pts)
    T = rng.standard_normal((4, N))
    F = forest.Forest(pts, MHK, p=48, cutoff=50)
    u0 = F.eval(T)
    F.store_matrices(True)
    u1 = F.eval(T)
    assert rel(u1, u0) < 1e-13
    np.testing.assert_array_equal(u1[2], F.eval(T[2]))
    F.store_matrices(False)
    np.testing.assert_array_equal(F.eval(T), u0)


def test_stored_matrices_class_fmm_expand_and_targets():
    """class_fmm's build_expansions (kf_plan_expand: the stored P2M) and evaluate_to_points
    (arbitrary targets: matrix-free) on a plan with stored matrices, against the oracle."""
    from pykifmm2d_amd import class_fmm
    x, y, tau = workloads.clustered(15000, seed=9)
    F = class_fmm.FMM(x, y, fmm.Get_Kernel_Functions(K0_T), class_fmm.prepare_numba_functions_on_the_fly(K0_T),
                      Nequiv=48, Ncutoff=50)
    F.precompute()
    F._plan.store_matrices(True)
    F.build_expansions(tau)
    R = fmm_ref.FMMRef(x, y, fmm_ref.modified_helmholtz(10.0), 48, 50)
    R.precompute()
    R.build_expansions(tau)
    assert rel(F.evaluate_to_points(x, y, check_self=True), R.evaluate_to_points(x, y, check_self=True)) < 1e-12
    rng = np.random.default_rng(3)
    tx, ty = rng.random(4000), rng.random(4000)
    assert rel(F.evaluate_to_points(tx, ty), R.evaluate_to_points(tx, ty)) < 1e-12
