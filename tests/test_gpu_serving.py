"""GPU tests of the request loop: resident trees, continuous batching, admission by footprint.

Bars: a request's potential is the same when it is evaluated alone and when it shares a batch
with others, and the plan's own single evaluation: bitwise in a batch of one, within 1e-12 (FP64)
or 1e-6 (split precision) of max |u| otherwise (a single evaluation sums the near field in the
symmetric form, batches of 16 and more sum the near field, P2M and L2P on the MFMA); FP64 served
potentials within 1e-10 (p = 48) of the FP64 direct sum, split precision within 1e-6.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import fmm_ref
from pykifmm2d_amd import fmm, kernels, serving, workloads


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


def same(a, b, tol=1e-12):
    """Bitwise equal, or (either side from a batch of 16 or more) within tol of max |b|."""
    import torch
    return torch.equal(a, b) or float((a - b).abs().max()) <= tol * float(b.abs().max())


def _densities(server, tids, n_requests, seed):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    rng = np.random.default_rng(seed)
    which = rng.integers(0, len(tids), n_requests)
    taus = [torch.randn(server.tree_info(tids[w])["n"], generator=g, device="cuda", dtype=torch.float64)
            for w in which]
    return which, taus


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("split", 1e-6)])
def test_results_same_alone_and_in_the_mix(precision, tol):
    import torch
    curves = workloads.star_curves(6, seed=21, n_min=2e3, n_max=2e4)
    server = serving.Server(slots=3)
    tids = [server.add_tree(x, y, p=48, cutoff=50, precision=precision) for x, y in curves]
    which, taus = _densities(server, tids, 60, seed=7)
    # all arrive at once: every tree's requests share batches
    mixed = [server.submit(tids[w], t) for w, t in zip(which, taus)]
    done = server.drain()
    assert len(done) == len(mixed) and all(q.done for q in mixed)
    assert max(q.batch_size for q in mixed) > 1
    assert server.pending == 0 and server.in_flight == 0
    # the same densities one at a time
    for q, t in list(zip(mixed, taus))[:15]:
        alone = server.submit(q.tree_id, t)
        server.drain()
        assert alone.batch_size == 1
        if q.batch_size == 1:  # the same arithmetic, bitwise
            assert torch.equal(alone.out, q.out), "request result depends on its batch"
        else:
            assert same(alone.out, q.out, 1e-12 if precision == "fp64" else 1e-6), "request result depends on its batch"
    # against the plan's own evaluation and the direct sum
    q = mixed[0]
    plan = server._res[q.tree_id].plan
    assert same(plan.eval(q.tau), q.out, 1e-12 if precision == "fp64" else 1e-6)
    x, y = curves[tids.index(q.tree_id)]
    d = fmm_ref.direct(fmm_ref.LAPLACE, x, y, q.tau.cpu().numpy())
    e = rel(q.out.cpu().numpy(), d)
    print(precision, "served rel err vs direct", e)
    assert e < tol


def test_poisson_arrivals_complete_with_latency():
    curves = workloads.star_curves(4, seed=22, n_min=2e3, n_max=1e4)
    server = serving.Server(slots=2)
    tids = [server.add_tree(x, y, p=32, cutoff=50) for x, y in curves]
    _, taus = _densities(server, tids, 1, seed=1)
    pool = {tid: None for tid in tids}
    import torch
    for tid in tids:
        pool[tid] = torch.randn(server.tree_info(tid)["n"], device="cuda", dtype=torch.float64)
    reqs, wall = serving.run_arrivals(server, tids, 200, lambda i, tid: pool[tid], rate=2000.0, seed=3)
    assert len(reqs) == 200 and all(q.done for q in reqs)
    s = serving.latency_summary(reqs)
    assert s["count"] == 200 and 0.0 < s["p50_ms"] <= s["p99_ms"] <= s["max_ms"]
    # every result equals the plan's own evaluation of its density
    for tid in tids:
        plan = server._res[tid].plan
        ref = plan.eval(pool[tid])
        for q in reqs:
            if q.tree_id == tid:
                assert same(q.out, ref) and (q.batch_size > 1 or torch.equal(q.out, ref))


def test_admission_by_footprint():
    x, y = workloads.star_curves(1, seed=23, n_min=3e4, n_max=3e4)[0]
    big = serving.Server(slots=1)
    t = big.add_tree(x, y, p=48, cutoff=50)
    full = big.tree_info(t)
    assert full["width"] == 64
    # the first tree's cost: its own bytes and the slot workspace grown to its slabs
    cost64 = big.used
    assert cost64 == full["footprint"] + big.workspace_bytes
    assert big.workspace_bytes >= full["slab_bytes"]
    # a budget one byte short of the 64-wide cost reserves a narrower width
    s2 = serving.Server(memory_budget=cost64 - 1, slots=1)
    t2 = s2.add_tree(x, y, p=48, cutoff=50)
    info = s2.tree_info(t2)
    assert 1 <= info["width"] < 64 and s2.used <= s2.budget
    # batches never exceed the reserved width, and results match the unconstrained server
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    taus = [torch.randn(info["n"], generator=g, device="cuda", dtype=torch.float64)
            for _ in range(2 * info["width"] + 1)]
    a = [s2.submit(t2, tau) for tau in taus]
    s2.drain()
    assert max(q.batch_size for q in a) <= info["width"]
    b = [big.submit(t, tau) for tau in taus]
    big.drain()
    for qa, qb in zip(a, b):
        assert same(qa.out, qb.out)
    # a slot cap below the 64-wide slabs narrows the width; the workspaces stay within slots x cap
    cap = full["slab_bytes"] // 3
    s4 = serving.Server(slots=2, slot_slab_bytes=cap)
    t4 = s4.add_tree(x, y, p=48, cutoff=50)
    assert s4.tree_info(t4)["width"] == 16 and s4.workspace_bytes <= 2 * cap
    # nothing fits: refused, budget unchanged
    s3 = serving.Server(memory_budget=1 << 20, slots=1)
    with pytest.raises(serving.AdmissionError):
        s3.add_tree(x, y, p=48, cutoff=50)
    assert s3.used == 0 and s3.trees == []
    # an idle tree can be removed and its budget returns (the last planned tree takes the
    # workspaces' memory with it)
    big.remove_tree(t)
    assert big.used == 0
    # the footprint of a width is exact: it is what reserving that width leaves allocated
    tree = fmm.Tree(x, y, 50)
    plan = fmm._Plan(tree, kernels.Laplace(), 48)
    f1, f16 = plan.footprint(1), plan.footprint(16)
    assert f1 < f16
    plan.reserve(16)
    assert plan.footprint(16) == f16 and plan.footprint(1) == f16  # slabs never shrink


def test_workspace_shared_by_plans():
    """Plans on one workspace (kf_plan_attach_workspace), evaluated alternately at several
    widths, give bitwise their own-slab results (a plan with graph replay enabled evaluates by
    launches while it is attached); the workspace holds the
    largest slabs, the plans' footprints none; an expansion kept for later calls is refused once
    another plan has evaluated on the workspace."""
    import torch
    from pykifmm2d_amd import _lib
    from pykifmm2d_amd.tree import Tree
    sets = [workloads.clustered(30000, seed=41)[:2], workloads.annulus(12000, seed=42)[:2],
            workloads.uniform(5000, seed=43)[:2]]
    kers = [kernels.Laplace(), kernels.ModifiedHelmholtz(10.0), kernels.Laplace()]
    own = [fmm._Plan(Tree(x, y, 40), k, 32) for (x, y), k in zip(sets, kers)]
    shared = [fmm._Plan(Tree(x, y, 40), k, 32) for (x, y), k in zip(sets, kers)]
    ws = fmm.Workspace()
    for pl in shared:
        pl.attach_workspace(ws)
    shared[2].set_graphs(True)
    f_own = shared[0].footprint(8)
    assert f_own < own[0].footprint(8) and f_own + shared[0].slab_bytes(8) == own[0].footprint(8)
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    for rnd, k in enumerate((1, 8, 3, 64, 1, 16)):
        for i in (0, 2, 1) if rnd % 2 else (1, 0, 2):
            n = sets[i][0].size
            T = torch.randn(k, n, generator=g, device="cuda", dtype=torch.float64)
            if rnd >= 2:  # NaN in every slab: an element read before it is written would show
                _lib.check(_lib.lib().kf_workspace_fill(ws.handle, float("nan"), None))
            assert torch.equal(shared[i].eval(T), own[i].eval(T)), (rnd, k, i)
    assert ws.nbytes >= max(pl.slab_bytes(64) for pl in shared[:2])
    # an expansion kept for kf_plan_self is overwritten by another plan's evaluation
    x0 = torch.randn(sets[0][0].size, device="cuda", dtype=torch.float64)
    pot = torch.empty_like(x0)
    _lib.check(_lib.lib().kf_plan_expand(shared[0].handle, ctypes.c_void_p(x0.data_ptr()), x0.numel(), 1, None))
    shared[1].eval(torch.randn(sets[1][0].size, device="cuda", dtype=torch.float64))
    rc = _lib.lib().kf_plan_self(shared[0].handle, ctypes.c_void_p(pot.data_ptr()), x0.numel(), None)
    assert rc != 0 and "workspace" in _lib.lib().kf_last_error().decode()
    # detached, a plan goes back to slabs of its own
    shared[0].attach_workspace(None)
    assert shared[0].footprint(1) > f_own - 1
    T = torch.randn(2, sets[0][0].size, device="cuda", dtype=torch.float64)
    assert torch.equal(shared[0].eval(T), own[0].eval(T))


def test_eval_ptrs_matches_strided():
    import torch
    x, y, _ = workloads.clustered(15000, seed=31)
    kf = fmm.Get_Kernel_Functions(kernels.Laplace())
    nf = fmm.prepare_numba_functions_planned(kernels.Laplace())
    plan = fmm.fmm_planner(x, y, 32, 50, kf, nf)
    T = torch.randn(5, x.size, device="cuda", dtype=torch.float64)
    u = fmm.planned_fmm(plan, T)
    p = plan.plan
    taus = [T[r].clone() for r in (3, 0, 4, 1, 2)]
    pots = [torch.empty_like(t) for t in taus]
    p.eval_ptrs(taus, pots)
    for r, pot in zip((3, 0, 4, 1, 2), pots):
        assert torch.equal(pot, u[r])
    with pytest.raises(ValueError):
        p.eval_ptrs(taus, pots[:2])


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_graph_replay_matches_launched_evaluation(precision):
    """Captured evaluations replayed with new addresses and widths, with launched work and
    synchronisations between the replays, give the launched results bitwise."""
    import torch
    x, y, _ = workloads.clustered(20000, seed=17)
    kf = fmm.Get_Kernel_Functions(kernels.Laplace())
    nf = fmm.prepare_numba_functions_planned(kernels.Laplace())
    eager = fmm.fmm_planner(x, y, 48, 64, kf, nf, precision=precision).plan
    graph = fmm.fmm_planner(x, y, 48, 64, kf, nf, precision=precision).plan
    graph.reserve(64)
    graph.set_graphs(True)
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    for k in (1, 3, 1, 16, 64, 5, 1):
        T = torch.randn(k, x.size, generator=g, device="cuda", dtype=torch.float64)
        outs = [torch.empty(x.size, device="cuda", dtype=torch.float64) for _ in range(k)]
        graph.eval_ptrs([T[r] for r in range(k)], outs)
        ref = eager.eval(T)
        torch.cuda.synchronize()
        for r in range(k):
            assert torch.equal(outs[r], ref[r]), (k, r)
        # launched evaluation on the graph plan between replays
        graph.set_graphs(False)
        assert torch.equal(graph.eval(T[:k]), ref)
        graph.set_graphs(True)
