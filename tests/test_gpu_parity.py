"""GPU parity: libkifmm2d (through its C-ABI, via the drop-in modules) vs the oracle.

Bars: FP64 potentials within 1e-12 relative (max-norm, scaled by max|u|) of the oracle's
restatement of the reference FMM; tree arrays bit-exact; direct sums within 1e-13.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import fmm_ref
from oracle.tree_ref import Tree as RefTree
from pykifmm2d_amd import class_fmm, fmm, kernels, workloads
from pykifmm2d_amd.tree import Tree

LAP = kernels.Laplace()
MHK = kernels.ModifiedHelmholtz(10.0)
REF_LAP = fmm_ref.LAPLACE
REF_MHK = fmm_ref.modified_helmholtz(10.0)

FIELDS = ["xmin", "ymin", "bot_ind", "top_ind", "parent_ind", "children_ind", "leaf", "fake_leaf",
          "Xlist", "colleagues"]


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


def funcs(k):
    return fmm.Get_Kernel_Functions(k), fmm.prepare_numba_functions_on_the_fly(k)


@pytest.mark.parametrize("case", ["uniform", "clustered", "annulus", "grid"])
def test_tree_matches_oracle(case):
    if case == "uniform":
        x, y, _ = workloads.uniform(20000, seed=1)
        c = 50
    elif case == "clustered":
        x, y, _ = workloads.clustered(50000, seed=2)
        c = 64
    elif case == "annulus":
        x, y, _ = workloads.annulus(20000, seed=3)
        c = 64
    else:
        g = np.arange(0, 33) / 32.0
        x, y = [a.ravel() for a in np.meshgrid(g, g)]
        c = 4
    t = Tree(x, y, c)
    r = RefTree(x, y, c)
    assert t.levels == r.levels
    for l, (a, b) in enumerate(zip(t.Levels, r.Levels)):
        for f in FIELDS:
            np.testing.assert_array_equal(np.asarray(getattr(a, f)), np.asarray(getattr(b, f)),
                                          err_msg="level %d %s" % (l, f))
    ordv, xs = t.ordv, t.x
    np.testing.assert_array_equal(np.sort(ordv), np.arange(x.size))
    np.testing.assert_array_equal(xs, x[ordv])
    for lv, rv in zip(t.Levels, r.Levels):
        for i in np.nonzero(lv.leaf & (lv.ns > 0))[0][:300]:
            assert set(ordv[lv.bot_ind[i]:lv.top_ind[i]]) == set(r.ordv[rv.bot_ind[i]:rv.top_ind[i]])


@pytest.mark.parametrize("kern,ref", [(LAP, REF_LAP), (MHK, REF_MHK)])
def test_direct_and_kernel_form(kern, ref):
    x, y, tau = workloads.uniform(3000, seed=4)
    KF, KA, KAS = fmm.Get_Kernel_Functions(kern)
    out = np.zeros(x.size)
    KAS(x, y, tau, out)
    assert rel(out, fmm_ref.direct(ref, x, y, tau)) < 1e-13
    tx, ty = x[:500] * 0.5 + 0.2, y[:500] * 0.7 + 0.1
    out2 = np.zeros(tx.size)
    KA(x, y, tx, ty, tau, out2)
    assert rel(out2, fmm_ref.direct(ref, x, y, tau, tx, ty)) < 1e-13
    M = fmm.Kernel_Form(KF, x[:300], y[:300], tx[:200], ty[:200])
    assert rel(M, fmm_ref.kernel_form(ref, x[:300], y[:300], tx[:200], ty[:200])) < 1e-14


def test_k0_device_matches_scipy():
    import scipy.special
    r = np.concatenate([np.logspace(-9, -0.5, 400), np.linspace(0.31, 60.0, 2000)])
    KF, _, _ = fmm.Get_Kernel_Functions(kernels.ModifiedHelmholtz(1.0, scale=1.0))
    M = fmm.Kernel_Form(KF, np.zeros(1), np.zeros(1), r, np.zeros_like(r))[:, 0]
    ref = scipy.special.k0(r)
    assert float(np.max(np.abs(M - ref) / ref)) < 2e-15


@pytest.mark.parametrize("n,p,c,gen,kern,ref", [
    (10000, 32, 50, "uniform", LAP, REF_LAP),
    (20000, 64, 64, "clustered", LAP, REF_LAP),
    (12000, 48, 64, "annulus", MHK, REF_MHK),
    (6000, 16, 20, "clustered", LAP, REF_LAP),
])
def test_on_the_fly_matches_oracle(n, p, c, gen, kern, ref):
    x, y, tau = getattr(workloads, gen)(n)
    kf, nf = funcs(kern)
    u, tree = fmm.on_the_fly_fmm(x, y, tau, p, c, kf, nf)
    uo, _ = fmm_ref.on_the_fly_fmm(x, y, tau, p, c, ref)
    print("rel vs oracle", rel(u, uo))
    assert rel(u, uo) < 1e-12
    if p == 64:
        d = fmm_ref.direct(ref, x, y, tau)
        assert rel(u, d) < 1e-12


def test_class_fmm_targets_match_oracle():
    x, y, tau = workloads.clustered(15000, seed=9)
    kf, _ = fmm.Get_Kernel_Functions(LAP), None
    nf = class_fmm.prepare_numba_functions_on_the_fly(LAP)
    F = class_fmm.FMM(x, y, kf, nf, Nequiv=48, Ncutoff=50)
    F.precompute()
    F.build_expansions(tau)
    R = fmm_ref.FMMRef(x, y, REF_LAP, 48, 50)
    R.precompute()
    R.build_expansions(tau)
    # self evaluation through the target path (coordinate self-exclusion)
    u = F.evaluate_to_points(x, y, check_self=True)
    assert rel(u, R.evaluate_to_points(x, y, check_self=True)) < 1e-12
    # arbitrary targets inside the root box
    rng = np.random.default_rng(3)
    tx, ty = rng.random(5000), rng.random(5000)
    v = F.evaluate_to_points(tx, ty)
    assert rel(v, R.evaluate_to_points(tx, ty)) < 1e-12
    # repeated build_expansions is independent of the previous call
    F.build_expansions(2.0 * tau)
    assert rel(F.evaluate_to_points(tx, ty), 2.0 * R.evaluate_to_points(tx, ty)) < 1e-12


def test_planned_multi_rhs_and_linearity():
    x, y, _ = workloads.clustered(20000, seed=5)
    rng = np.random.default_rng(8)
    T = rng.standard_normal((3, x.size))
    kf = fmm.Get_Kernel_Functions(LAP)
    nf = fmm.prepare_numba_functions_planned(LAP)
    plan = fmm.fmm_planner(x, y, 32, 50, kf, nf)
    u = fmm.planned_fmm(plan, T)
    np.testing.assert_array_equal(u, fmm.planned_fmm(plan, T))  # bitwise deterministic
    for r in range(3):
        # a single evaluation sums the near field in the symmetric form (each pair of mutual
        # colleague leaves once): the same to rounding; with the per-pair form a right-hand
        # side's arithmetic is the same at every batch width up to 8 (bitwise)
        assert rel(fmm.planned_fmm(plan, T[r]), u[r]) < 1e-13
    plan.plan.set_tuning(near_form=1)
    for r in range(3):
        np.testing.assert_array_equal(u[r], fmm.planned_fmm(plan, T[r]))
    plan.plan.set_tuning(near_form=0)
    lin = fmm.planned_fmm(plan, 2.0 * T[0] - 3.0 * T[1])
    assert rel(lin, 2.0 * u[0] - 3.0 * u[1]) < 1e-12
    uo, _ = fmm_ref.on_the_fly_fmm(x, y, T[2], 32, 50, REF_LAP)
    assert rel(u[2], uo) < 1e-12


@pytest.mark.parametrize("kern,ref,nrhs", [(LAP, REF_LAP, 20), (LAP, REF_LAP, 70), (MHK, REF_MHK, 16)])
def test_batched_rhs_match_single(kern, ref, nrhs):
    """Batches of right-hand sides (widths 16, 32 and 64 + 8, padded) give the single-RHS result
    within 1e-12 of max |u| (widths of 16 and more sum the near field, P2M and L2P on the MFMA,
    in another order than the vector forms of narrower batches)."""
    gen = workloads.annulus if kern is MHK else workloads.clustered
    x, y, _ = gen(12000)
    rng = np.random.default_rng(11)
    T = rng.standard_normal((nrhs, x.size))
    kf = fmm.Get_Kernel_Functions(kern)
    nf = fmm.prepare_numba_functions_planned(kern)
    plan = fmm.fmm_planner(x, y, 48, 64, kf, nf)
    u = fmm.planned_fmm(plan, T)
    assert u.shape == T.shape
    for r in (0, nrhs // 2, nrhs - 1):
        assert rel(u[r], fmm.planned_fmm(plan, T[r])) < 1e-12
    uo, _ = fmm_ref.on_the_fly_fmm(x, y, T[nrhs - 1], 48, 64, ref)
    assert rel(u[nrhs - 1], uo) < 1e-12


@pytest.mark.parametrize("precision", ["split", "split16"])
@pytest.mark.parametrize("gen,kern,ref,p,nrhs", [("clustered", LAP, REF_LAP, 64, 1),
                                                 ("annulus", MHK, REF_MHK, 48, 1),
                                                 ("clustered", LAP, REF_LAP, 48, 16)])
def test_split_precision_error(gen, kern, ref, p, nrhs, precision):
    """FP32 M2L (split) or M2L products from FP16 halves (split16), everything else FP64: error
    against the FP64 direct sum at most 1e-6 (relative, max-norm)."""
    x, y, _ = getattr(workloads, gen)(20000)
    T = np.random.default_rng(13).standard_normal((nrhs, x.size))
    kf = fmm.Get_Kernel_Functions(kern)
    nf = fmm.prepare_numba_functions_planned(kern)
    plan = fmm.fmm_planner(x, y, p, 64, kf, nf, precision=precision)
    u = fmm.planned_fmm(plan, T if nrhs > 1 else T[0]).reshape(nrhs, -1)
    for r in (0, nrhs - 1):
        d = fmm_ref.direct(ref, x, y, T[r])
        e = rel(u[r], d)
        print(precision, "rel err", e)
        assert e < 1e-6


def test_split16_against_split_and_widths():
    """The FP16-split M2L against the FP32 one on the same plan inputs (same order of error:
    within 3e-7 of max |u| of each other), a right-hand side bitwise the same at widths 16, 32
    and 64 and within 1e-6 of its single evaluation, and through a workspace."""
    import torch
    x, y, _ = workloads.clustered(30000, seed=17)
    kf = fmm.Get_Kernel_Functions(LAP)
    nf = fmm.prepare_numba_functions_planned(LAP)
    a = fmm.fmm_planner(x, y, 64, 64, kf, nf, precision="split").plan
    b = fmm.fmm_planner(x, y, 64, 64, kf, nf, precision="split16").plan
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    T = torch.randn(64, x.size, generator=g, device="cuda", dtype=torch.float64)
    ua, ub = a.eval(T), b.eval(T)
    for r in (0, 33, 63):
        assert float((ua[r] - ub[r]).abs().max() / ua[r].abs().max()) < 3e-7
    assert torch.equal(b.eval(T[:32]), ub[:32]) and torch.equal(b.eval(T[:16]), ub[:16])
    one = b.eval(T[7])
    assert float((one - ub[7]).abs().max() / ub[7].abs().max()) < 1e-6
    ws = fmm.Workspace()
    b.attach_workspace(ws)
    assert torch.equal(b.eval(T), ub)
    b.attach_workspace(None)


@pytest.mark.parametrize("p,nrhs", [(16, 1), (32, 8), (48, 64)])
def test_split16_small_orders(p, nrhs):
    """The FP16-split M2L at the orders whose FP16 K extent is padded (P = 16, 48) or a single
    K chunk (P = 32), single-tree, against the FP32 M2L on the same plan inputs (within 3e-7 of
    max |u|) and the FP64 direct sum (within 2e-6)."""
    import torch
    x, y, tau = workloads.clustered(20000, seed=29)
    kf = fmm.Get_Kernel_Functions(LAP)
    nf = fmm.prepare_numba_functions_planned(LAP)
    a = fmm.fmm_planner(x, y, p, 64, kf, nf, precision="split").plan
    b = fmm.fmm_planner(x, y, p, 64, kf, nf, precision="split16").plan
    g = torch.Generator(device="cuda")
    g.manual_seed(p)
    T = torch.randn(nrhs, x.size, generator=g, device="cuda", dtype=torch.float64)
    ua, ub = a.eval(T).reshape(nrhs, -1), b.eval(T).reshape(nrhs, -1)
    for r in {0, nrhs - 1}:
        assert float((ua[r] - ub[r]).abs().max() / ua[r].abs().max()) < 3e-7, (p, r)
    d = fmm_ref.direct(fmm_ref.LAPLACE, x, y, T[0].cpu().numpy())
    e = rel(ub[0].cpu().numpy(), d)
    print("split16 p=%d rel err %.3g" % (p, e))
    assert e < (2e-6 if p >= 32 else 1e-4)


def test_split16_forest():
    """A forest in the FP16-split precision against the same forest in split precision."""
    from pykifmm2d_amd import forest
    rng = np.random.default_rng(31)
    pts = [(rng.random(n), rng.random(n)) for n in (3000, 12000, 700, 25000)]
    N = sum(len(px) for px, _ in pts)
    T = rng.standard_normal((3, N))
    A = forest.Forest(pts, LAP, p=32, cutoff=50, precision="split")
    B = forest.Forest(pts, LAP, p=32, cutoff=50, precision="split16")
    ua, ub = A.eval(T), B.eval(T)
    for r in range(3):
        assert np.abs(ua[r] - ub[r]).max() / np.abs(ua[r]).max() < 3e-7, r


def test_c3_form_64_rhs_split_precision():
    """The modified Helmholtz headline form at reduced size: K0(10 r)/(2 pi), points uniform in
    the annulus 0.2 < |x - c| < 0.5, p = 48, cutoff 64, 64 right-hand sides in one batched
    split-precision evaluation (FP32 M2L on 1 target x 64 right-hand sides per tile row).  The
    first, a middle and the last right-hand side against the FP64 direct sum and the oracle's
    FMM at the 1e-6 bar."""
    x, y, _ = workloads.annulus(24000, seed=31)
    T = np.random.default_rng(32).standard_normal((64, x.size))
    kf = fmm.Get_Kernel_Functions(MHK)
    nf = fmm.prepare_numba_functions_planned(MHK)
    plan = fmm.fmm_planner(x, y, 48, 64, kf, nf, precision="split")
    u = fmm.planned_fmm(plan, T)
    assert u.shape == T.shape
    for r in (0, 31, 63):
        d = fmm_ref.direct(REF_MHK, x, y, T[r])
        uo, _ = fmm_ref.on_the_fly_fmm(x, y, T[r], 48, 64, REF_MHK)
        e_d, e_o = rel(u[r], d), rel(u[r], uo)
        print("64-RHS split, rhs %d: vs direct %.3e, vs oracle FMM %.3e" % (r, e_d, e_o))
        assert e_d < 1e-6 and e_o < 1e-6


def test_small_trees_fall_back_to_direct():
    x, y, tau = workloads.uniform(120, seed=6)
    kf, nf = funcs(LAP)
    u, tree = fmm.on_the_fly_fmm(x, y, tau, 32, 50, kf, nf)
    assert tree.levels <= 2
    assert rel(u, fmm_ref.direct(REF_LAP, x, y, tau)) < 1e-13


def test_rejects_points_outside_reference_root_box():
    # tree.py:478 uses y.min() for the upper bound; y values above it are rejected
    x = np.array([0.1, 0.2, 0.3, 0.4])
    y = np.array([0.1, 0.2, 1.5, 2.5])
    from pykifmm2d_amd._lib import KifmmError
    with pytest.raises(KifmmError):
        Tree(x, y, 1)


def test_class_fmm_complex_densities():
    """iscomplex=True: complex densities as two right-hand sides; real and imaginary parts give
    bitwise the potentials of two real solves, and match the oracle."""
    x, y, _ = workloads.clustered(12000, seed=19)
    rng = np.random.default_rng(4)
    a, b = rng.standard_normal(x.size), rng.standard_normal(x.size)
    kf = fmm.Get_Kernel_Functions(LAP)
    nf = class_fmm.prepare_numba_functions_on_the_fly(LAP)
    C = class_fmm.FMM(x, y, kf, nf, Nequiv=48, Ncutoff=50, iscomplex=True)
    C.precompute()
    C.build_expansions(a + 1j * b)
    tx, ty = rng.random(3000), rng.random(3000)
    u = C.evaluate_to_points(tx, ty)
    assert np.iscomplexobj(u) and u.shape == (3000,)
    F = class_fmm.FMM(x, y, kf, nf, Nequiv=48, Ncutoff=50)
    F.precompute()
    F.build_expansions(a)
    ua = F.evaluate_to_points(tx, ty)
    F.build_expansions(b)
    ub = F.evaluate_to_points(tx, ty)
    np.testing.assert_array_equal(u.real, ua)
    np.testing.assert_array_equal(u.imag, ub)
    R = fmm_ref.FMMRef(x, y, REF_LAP, 48, 50)
    R.precompute()
    R.build_expansions(b)
    assert rel(u.imag, R.evaluate_to_points(tx, ty)) < 1e-12
    with pytest.raises(TypeError):
        F.build_expansions(a + 1j * b)


@pytest.mark.parametrize("kern,precision,tol", [(LAP, "fp64", 1e-12), (MHK, "fp64", 1e-12), (LAP, "split", 1e-6)])
def test_batched_mfma_forms(kern, precision, tol):
    """Batches of 16, 32 and 64 right-hand sides form the near field, P2M and L2P on the FP64
    MFMA by default: a right-hand side's potential is bitwise the same at each of these widths,
    and within tol of max |u| of the vector forms (batch_form=1) and of single evaluations; the
    stored-matrix plan (vector forms over stored values) agrees to the same bar."""
    import torch
    gen = workloads.annulus if kern is MHK else workloads.clustered
    x, y, _ = gen(25000, seed=29)
    plan = fmm.fmm_planner(x, y, 48, 64, fmm.Get_Kernel_Functions(kern),
                           fmm.prepare_numba_functions_planned(kern), precision=precision).plan
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    T = torch.randn(64, x.size, generator=g, device="cuda", dtype=torch.float64)
    u64 = plan.eval(T)
    assert torch.equal(plan.eval(T[:32]), u64[:32])
    assert torch.equal(plan.eval(T[:16]), u64[:16])
    assert torch.equal(plan.eval(T[16:32]), u64[16:32])

    def trel(a, b):
        return float((a - b).abs().max() / b.abs().max())

    plan.set_tuning(batch_form=1)
    assert plan.tuning()["batch_form"] == 1
    v64 = plan.eval(T)
    for r in (0, 17, 63):
        e = trel(u64[r], v64[r])
        print(precision, "MFMA vs vector forms, rhs", r, e)
        assert e < tol
    plan.set_tuning(batch_form=0)
    for r in (5, 40):
        assert trel(u64[r], plan.eval(T[r])) < tol
    plan.store_matrices(True)
    w = plan.eval(T[:16])
    assert trel(w[3], u64[3]) < tol
    plan.store_matrices(False)
    assert torch.equal(plan.eval(T[:16]), u64[:16])


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_launch_tunings_give_identical_results(precision):
    """Near-field grid, near-field fork point and split, the pipelined M2L and the rows per
    workgroup of the level launches change only the launch structure: every combination gives
    bitwise the default result; the autotuner picks one."""
    import torch
    x, y, _ = workloads.clustered(30000, seed=23)
    kf = fmm.Get_Kernel_Functions(LAP)
    nf = fmm.prepare_numba_functions_planned(LAP)
    plan = fmm.fmm_planner(x, y, 48, 64, kf, nf, precision=precision).plan
    T = torch.randn(3, x.size, device="cuda", dtype=torch.float64)
    ref, ref1 = plan.eval(T), plan.eval(T[1])
    for chain in (1, 2):
        for grid in (0, 64, 1024):
            for fork in (0, 1, 2, 3, 4):
                for pipe in (0, 1):
                    for split in (-1, 0, 300, 1000):
                        plan.set_tuning(grid, fork, pipe, near_split=split, chain=chain)
                        assert plan.tuning() == {"near_grid": grid, "near_fork": fork, "m2l_pipeline": pipe,
                                                 "level_rows": 0, "near_split": split, "chain": chain,
                                                 "batch_form": 0, "near_form": 0}
                        assert torch.equal(plan.eval(T), ref), (grid, fork, pipe, split, chain)
                        assert torch.equal(plan.eval(T[1]), ref1), (grid, fork, pipe, split, chain)
    for rows in (16, 32, 64, 128):
        plan.set_tuning(level_rows=rows)
        assert torch.equal(plan.eval(T), ref), rows
        assert torch.equal(plan.eval(T[1]), ref1), rows
    # a wide batch: the 8-wave level kernels (128 rows, two boxes or a tile pair per workgroup)
    T64 = torch.randn(64, x.size, device="cuda", dtype=torch.float64)
    plan.set_tuning(level_rows=16)
    ref64 = plan.eval(T64)
    for rows in (0, 64, 128):
        plan.set_tuning(level_rows=rows)
        assert torch.equal(plan.eval(T64), ref64), rows
    # the persistent chains at every batch width and over the targets of all boxes
    plan.set_tuning(chain=2)
    T8 = torch.randn(8, x.size, device="cuda", dtype=torch.float64)
    plan.set_tuning(chain=1)
    ref8 = plan.eval(T8)
    plan.set_tuning(chain=2)
    assert torch.equal(plan.eval(T8), ref8)
    from pykifmm2d_amd import _lib
    _lib.check(_lib.lib().kf_plan_verify(plan.handle, None))  # no chain wait gave up
    best, ms = plan.autotune(nrhs=3)
    print("autotune", best, ms)
    assert ms > 0 and best["near_fork"] in (1, 2, 3, 4) and best["m2l_pipeline"] in (0, 1)
    assert plan.tuning() == best
    assert torch.equal(plan.eval(T), ref)


@pytest.mark.parametrize("gen,kern,ref,p,cutoff", [
    (workloads.clustered, LAP, REF_LAP, 64, 64),
    (workloads.uniform, LAP, REF_LAP, 32, 50),
    (workloads.annulus, MHK, REF_MHK, 48, 64),
    (workloads.clustered, LAP, REF_LAP, 32, 100),  # leaves above 64 points: the per-pair form
])
def test_symmetric_near_field(gen, kern, ref, p, cutoff):
    """One right-hand side: the near field with each pair of mutual colleague leaves evaluated
    once (near_form 2; the default for tabulated kernels) against every directed pair on its own
    (near_form 1): within 1e-13 of max |u| (bitwise where a leaf exceeds 64 points, which keeps
    the per-pair form), the same under every launch structure, and within 1e-12 of the oracle."""
    import torch
    x, y, tau = gen(30000, seed=37)
    plan = fmm.fmm_planner(x, y, p, cutoff, fmm.Get_Kernel_Functions(kern),
                           fmm.prepare_numba_functions_planned(kern)).plan
    t = torch.as_tensor(tau, device="cuda")
    plan.set_tuning(near_form=2)
    u = plan.eval(t)
    plan.set_tuning(near_form=1)
    v = plan.eval(t)
    if cutoff > 64:
        assert torch.equal(u, v)
    else:
        assert float((u - v).abs().max() / v.abs().max()) < 1e-13
    for grid, fork, split in ((1024, 0, 0), (0, 1, -1), (0, 0, 500)):
        plan.set_tuning(near_grid=grid, near_fork=fork, near_split=split, near_form=2)
        assert torch.equal(plan.eval(t), u), (grid, fork, split)
    uo, _ = fmm_ref.on_the_fly_fmm(x, y, tau, p, cutoff, ref)
    assert rel(u.cpu().numpy(), uo) < 1e-12
