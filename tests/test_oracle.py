"""Pins the oracle (CPU restatement of the reference) against mathematics, not against itself.

* direct sums: brute force in numpy on tiny inputs, math.fsum, reciprocity;
* FMM: exponential convergence to the FP64 direct sum as p grows (a dropped or mis-signed
  term would stall it), Laplace and modified Helmholtz, self and target evaluation;
* C2E matrix: symmetric circulant with the closed-form Laplace eigenvalues
  lambda_m = -(p/2pi) log(r2) [m=0] + (p/4pi) sum_k rho^k/k ([k=m] + [k=-m] mod p);
* tree: a brute-force geometric builder (split while > cutoff, balance by repeated sweeps,
  X-list by definition) on small inputs;
* K0: scipy's Cephes k0 against mpmath.
"""
import math

import numpy as np
import pytest
import scipy.special

from oracle import fmm_ref
from oracle.tree_ref import Tree
from pykifmm2d_amd import workloads

LAP = fmm_ref.LAPLACE
MHK = fmm_ref.modified_helmholtz(10.0)


def test_direct_matches_brute_force_and_fsum():
    rng = np.random.default_rng(0)
    x, y, tau = rng.random(60), rng.random(60), rng.standard_normal(60)
    for K in (LAP, MHK):
        G = fmm_ref.kernel_eval_np(K, x[None, :], y[None, :], x[:, None], y[:, None])
        np.fill_diagonal(G, 0.0)
        ref = G @ tau
        out = fmm_ref.direct(K, x, y, tau)
        assert np.abs(out - ref).max() <= 1e-14 * np.abs(ref).max()
        j = 7
        terms = [G[j, i] * tau[i] for i in range(60) if i != j]
        assert abs(out[j] - math.fsum(terms)) <= 1e-15 * max(1.0, abs(out[j]))


def test_direct_reciprocity_and_sampled_targets():
    x, y, tau = workloads.uniform(2000, seed=3)
    sig = np.random.default_rng(4).standard_normal(2000)
    u = fmm_ref.direct(LAP, x, y, tau)
    v = fmm_ref.direct(LAP, x, y, sig)
    assert abs(sig @ u - tau @ v) <= 1e-12 * abs(sig @ u)
    idx = np.array([0, 5, 1999])
    us = fmm_ref.direct(LAP, x, y, tau, x[idx], y[idx], target_index=idx)
    np.testing.assert_allclose(us, u[idx], rtol=1e-14)


def test_laplace_fmm_converges_exponentially():
    x, y, tau = workloads.uniform(10000, seed=1)
    d = fmm_ref.direct(LAP, x, y, tau)
    errs = []
    for p in (16, 24, 32, 48, 64):
        u, tree = fmm_ref.on_the_fly_fmm(x, y, tau, p, 50, LAP)
        errs.append(np.abs(u - d).max() / np.abs(d).max())
    assert tree.levels >= 5
    assert errs[-1] < 5e-14
    assert errs[2] < 1e-7
    assert all(b < a for a, b in zip(errs[:-1], errs[1:]))


def test_clustered_fmm_vs_direct():
    x, y, tau = workloads.clustered(30000, seed=2)
    u, tree = fmm_ref.on_the_fly_fmm(x, y, tau, 64, 64, LAP)
    idx = workloads.sample_targets(x.size, 3000)
    d = fmm_ref.direct(LAP, x, y, tau, x[idx], y[idx], target_index=idx)
    assert np.abs(u[idx] - d).max() / np.abs(d).max() < 1e-13
    assert any(lv.fake_leaf.any() for lv in tree.Levels)


def test_modified_helmholtz_fmm_vs_direct():
    x, y, tau = workloads.annulus(8000, seed=3)
    d = fmm_ref.direct(MHK, x, y, tau)
    u48, _ = fmm_ref.on_the_fly_fmm(x, y, tau, 48, 64, MHK)
    u64, _ = fmm_ref.on_the_fly_fmm(x, y, tau, 64, 64, MHK)
    e48 = np.abs(u48 - d).max() / np.abs(d).max()
    e64 = np.abs(u64 - d).max() / np.abs(d).max()
    assert e48 < 1e-10 and e64 < 1e-13 and e64 < e48


def test_class_target_evaluation_vs_direct():
    x, y, tau = workloads.clustered(12000, seed=9)
    F = fmm_ref.FMMRef(x, y, LAP, 64, 50)
    F.precompute()
    F.build_expansions(tau)
    rng = np.random.default_rng(5)
    tx, ty = rng.random(500), rng.random(500)
    d = fmm_ref.direct(LAP, x, y, tau, tx, ty)
    assert np.abs(F.evaluate_to_points(tx, ty) - d).max() / np.abs(d).max() < 1e-13
    us = F.evaluate_to_points(x, y, check_self=True)
    ds = fmm_ref.direct(LAP, x, y, tau)
    assert np.abs(us - ds).max() / np.abs(ds).max() < 1e-13
    # the self-potential path (index exclusion) agrees with the coordinate-exclusion path
    assert np.abs(F.self_potential() - us).max() / np.abs(ds).max() < 1e-14


@pytest.mark.parametrize("p,w", [(32, 1.0), (48, 0.5), (64, 4.0)])
def test_c2e_matrix_closed_form_eigenvalues(p, w):
    th = np.linspace(0, 2 * np.pi, p, endpoint=False)
    sx, sy, lx, ly, r1, r2 = fmm_ref.level_information(w, th)
    E = fmm_ref.kernel_form(LAP, sx, sy, lx, ly)
    assert np.abs(E - E.T).max() < 1e-13 * np.abs(E).max()
    assert np.abs(E - np.roll(np.roll(E, 1, 0), 1, 1)).max() < 1e-13 * np.abs(E).max()
    rho = r1 / r2
    lam = np.zeros(p)
    k = np.arange(1, 4000)
    for m in range(p):
        sel = ((k - m) % p == 0).astype(float) + ((k + m) % p == 0).astype(float)
        lam[m] = (p / (4 * np.pi)) * np.sum(rho ** k / k * sel) - (p / (2 * np.pi)) * np.log(r2) * (m == 0)
    ev = np.sort(np.linalg.eigvalsh(0.5 * (E + E.T)))
    assert np.abs(ev - np.sort(lam)).max() < 1e-13 * np.abs(lam).max()


# ---------------------------------------------------------------- brute-force tree -------

def brute_tree(x, y, c):
    """Geometric definition: boxes (l, ix, iy) of the root [m, m+W]^2 with the strict '>'
    midline rule; split while > c points; balance; X-list."""
    m = math.floor(min(x.min(), y.min()))
    W = math.ceil(max(x.max(), y.min())) - m

    def cell(l, v):
        # number of level-l gridlines strictly below v
        n = 2 ** l
        return np.array([sum(1 for k in range(1, n) if vi > m + k * W / n) for vi in v])

    pts = {}

    def members(l, i, j):
        key = (l, i, j)
        if key not in pts:
            cx, cy = cell(l, x), cell(l, y)
            pts[key] = set(np.nonzero((cx == i) & (cy == j))[0].tolist())
        return pts[key]

    boxes = {(0, 0, 0)}
    frontier = [(0, 0, 0)]
    while frontier:
        nxt = []
        for (l, i, j) in frontier:
            if len(members(l, i, j)) > c:
                for a in range(2):
                    for b in range(2):
                        ch = (l + 1, 2 * i + a, 2 * j + b)
                        boxes.add(ch)
                        nxt.append(ch)
        frontier = nxt
    L = max(b[0] for b in boxes) + 1

    def children(b):
        l, i, j = b
        return [(l + 1, 2 * i + a, 2 * j + bb) for a in range(2) for bb in range(2)]

    def is_leaf(b, bs):
        return children(b)[0] not in bs

    def depth(b, bs):
        return 0 if is_leaf(b, bs) else 1 + max(depth(ch, bs) for ch in children(b))

    def colleagues(b, bs):
        l, i, j = b
        return [(l, i + a, j + bb) for a in (-1, 0, 1) for bb in (-1, 0, 1) if (l, i + a, j + bb) in bs]

    changed = True
    while changed:
        changed = False
        for b in sorted(boxes):
            if 1 <= b[0] <= L - 3 and is_leaf(b, boxes):
                if any(depth(cb, boxes) >= 2 for cb in colleagues(b, boxes)):
                    boxes.update(children(b))
                    changed = True
    xl = set()
    for b in boxes:
        if 1 <= b[0] <= L - 2 and is_leaf(b, boxes):
            if any(depth(cb, boxes) > 0 for cb in colleagues(b, boxes)):
                xl.add(b)
    leaves = {b for b in boxes if is_leaf(b, boxes)}
    return boxes, leaves, xl, L, members


@pytest.mark.parametrize("seed,n,c", [(1, 60, 3), (2, 200, 6), (3, 120, 2)])
def test_tree_against_brute_force(seed, n, c):
    rng = np.random.default_rng(seed)
    # clustered so that level restriction and the X-list are exercised
    x = np.concatenate([rng.random(n // 2), 0.3 + 0.02 * rng.random(n - n // 2)])
    y = np.concatenate([rng.random(n // 2), 0.6 + 0.02 * rng.random(n - n // 2)])
    boxes, leaves, xl, L, members = brute_tree(x, y, c)
    t = Tree(x, y, c)
    assert t.levels == L
    got, got_leaf, got_xl = set(), set(), set()
    for l, lv in enumerate(t.Levels):
        w = lv.width
        for k in range(lv.n_node):
            if lv.fake_leaf[k]:
                continue
            key = (l, int(round((lv.xmin[k] - t.xmin) / w)), int(round((lv.ymin[k] - t.ymin) / w)))
            got.add(key)
            if lv.leaf[k]:
                got_leaf.add(key)
                assert set(t.ordv[lv.bot_ind[k]:lv.top_ind[k]].tolist()) == members(*key)
            if lv.Xlist[k]:
                got_xl.add(key)
                assert lv.children_ind[k] >= 0 and t.Levels[l + 1].fake_leaf[lv.children_ind[k]]
    assert got == boxes
    assert got_leaf == leaves
    assert got_xl == xl


def test_scipy_k0_is_pinned_by_mpmath():
    import mpmath
    for v in (1e-6, 0.1, 0.9, 1.5, 2.0, 3.7, 11.0, 55.0):
        assert abs(scipy.special.k0(v) - float(mpmath.besselk(0, v))) <= 3e-15 * float(mpmath.besselk(0, v))


def test_complex_kernel_numpy_passes_pinned():
    """The oracle's complex FMM (class_fmm with dtype complex, numpy restatements of its loops):
    (a) a complex multiple of the Laplace kernel, (1 + 0.5j) G, gives (1 + 0.5j) times the C-loop
    Laplace oracle at the sources and at targets (a conjugated operator, a transposed complex
    block or a dropped imaginary part fails this); (b) 0.25j H0(2 r) on a circle (the Helmholtz
    kernel of examples/hankel.py) converges to the direct sum with scipy's hankel1."""
    rng = np.random.default_rng(17)
    x, y, t = rng.random(2500), rng.random(2500), rng.standard_normal(2500)
    c = 1.0 + 0.5j
    KC = fmm_ref.complex_kernel(lambda r2: c * (-0.25 / np.pi) * np.log(r2))
    F = fmm_ref.FMMRef(x, y, KC, 24, 40)
    F.precompute()
    F.build_expansions(t)
    R = fmm_ref.FMMRef(x, y, fmm_ref.LAPLACE, 24, 40)
    R.precompute()
    R.build_expansions(t)
    u, v = F.self_potential(), R.self_potential()
    assert np.abs(u - c * v).max() < 1e-13 * np.abs(v).max()
    tx, ty = rng.random(200), rng.random(200)
    a, b = F.evaluate_to_points(tx, ty), R.evaluate_to_points(tx, ty)
    assert np.abs(a - c * b).max() < 1e-13 * np.abs(b).max()
    th = rng.random(2000) * 2 * np.pi
    cx, cy, tau = np.cos(th), np.sin(th), rng.random(2000)
    H = fmm_ref.hankel_kernel(2.0)
    d = fmm_ref.direct_np(H, cx, cy, tau)
    errs = []
    for p in (16, 32, 48):
        f = fmm_ref.FMMRef(cx, cy, H, p, 50)
        f.precompute()
        f.build_expansions(tau)
        errs.append(np.abs(f.self_potential() - d).max() / np.abs(d).max())
    assert errs[0] > errs[1] > errs[2] and errs[1] < 1e-7 and errs[2] < 1e-10, errs
