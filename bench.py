#!/usr/bin/env python
"""bench.py -- planned 2-D KIFMM evaluation throughput on one MI355X.

Workload: Laplace, N = 1,000,000 clustered points (70% in 8 Gaussians with sigma in
[1e-3, 5e-2], 30% uniform), charges N(0,1), p = 64 equivalent points, leaf cutoff 64, FP64.
A step is one planned evaluation u = FMM(tau) over all N points with the tree, operators and
work lists resident in HBM (gather tau -> P2M -> upward -> downward -> L2P + P2P -> scatter).
The error is stated against an FP64 direct sum (compensated, on the GPU) at 10,000 sampled
targets.  The tree + plan build is timed separately and reported with the on-the-fly rate.

Six more legs are reported beside the headline value:
  batched_split  the same tree, 64 right-hand sides per evaluation, FP32 M2L (split precision)
  batched_split16  the same, with the M2L as three FP16 products per FP32-accurate one
  tabulated      user kernels as tables: the same tree with the Laplace kernel given as a
                 function of r^2, and 1M points in the annulus below with K0(10 r)/(2 pi) given
                 as one (p = 48), beside the built-in kernel; each matrix-free and with the
                 stored kernel matrices (kf_plan_store_matrices)
  helmholtz_64rhs  modified Helmholtz K0(10 r)/(2 pi), N = 4,000,000 uniform in the annulus
                 0.2 < |x - (0.5, 0.5)| < 0.5, 64 right-hand sides per evaluation, p = 48,
                 cutoff 64, FP64 near field + FP32 M2L (and beside it the M2L as FP16 splits)
  forest_512     512 independent Laplace problems, N per problem log-uniform in [1e3, 2e5]
                 (about 19M points), uniform in their own unit boxes, p = 32, cutoff 50, FP64:
                 all trees built and evaluated in shared launches, beside a loop over the same
                 trees' single-tree plans
  served         64 resident star-curve trees (N per curve log-uniform in [2e3, 5e5]), Laplace
                 p = 48, cutoff 50: a Poisson stream of fresh densities at 70% of the measured
                 capacity through the continuous-batching request loop; p50/p99 latency

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (matrix and vector)
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 (matrix and vector)
PEAK_FP16_TFLOPS = 2500.0  # MI355X dense FP16 matrix
# kernel of each single-launch pass
# arithmetic of each precision mode: FP64 throughout; the M2L on FP32 MFMA (slot sums and
# every other pass FP64); the M2L as three FP16 MFMA products per FP32-accurate one
DTYPE = {"fp64": "f64", "split": "f32+f64", "split16": "f16x3+f64"}
PASS_KERNEL = {"p2p": "k_p2p", "p2m": "k_p2m", "l2p": "k_l2p", "m2l": "k_m2l"}


def pmc_traffic(kernel, workload):
    """(HBM bytes per launch of `kernel`, launch structure of that profile) from the committed
    rocprofv3 PMC summary of `workload` (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by
    what the counter reads for a plain copy of known size), or (None, None) when no summary was
    taken for this workload.  A kernel's own bytes do not depend on where the launch structure
    places it (fork point, pipelining)."""
    path = os.path.join(ROOT, "profiles", "kifmm_%s_pmc.txt" % workload)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            w = line.split()
            if len(w) == 4 and w[0] == "traffic_bytes_per_launch" and w[1] == kernel:
                return int(w[2]), w[3].split("=", 1)[1]
    return None, None
PEAK_HBM_GBS = 8000.0
# VALU issue rate: 256 CUs x 4 SIMDs, one wave64 instruction per 4 cycles (16 lanes per cycle), 2.4 GHz
PEAK_VALU_WAVE_INSTS = 256 * 4 * 2.4e9 / 4


def pmc_valu_insts(kernel, workload):
    """VALU wave instructions per launch of `kernel` from the committed PMC summary of
    `workload` (SQ_INSTS_VALU), or None."""
    path = os.path.join(ROOT, "profiles", "kifmm_%s_pmc.txt" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            w = line.split()
            if len(w) == 3 and w[0] == "valu_insts_per_launch" and w[1] == kernel:
                return int(w[2])
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--p", type=int, default=64)
    ap.add_argument("--cutoff", type=int, default=64)
    ap.add_argument("--err-samples", type=int, default=10000)
    ap.add_argument("--nrhs", type=int, default=1, help="right-hand sides per evaluation (batched)")
    ap.add_argument("--precision", choices=["fp64", "split", "split16"], default="fp64")
    ap.add_argument("--autotune", type=int, default=1,
                    help="after the headline (timed with the plan's default launch structure), run kf_plan_autotune "
                         "and time its choice too (reported beside the headline); the batched leg autotunes first")
    ap.add_argument("--tuning", default="",
                    help="near_grid,near_fork,m2l_pipeline,level_rows,near_split for the headline plan (overrides --autotune)")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the CPU oracle (port) too")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="minimum CPU baseline work")
    ap.add_argument("--batched", type=int, default=64, help="right-hand sides of the batched split leg (0: skip)")
    ap.add_argument("--helmholtz", type=int, default=1, help="run the 4M-point modified Helmholtz 64-RHS leg")
    ap.add_argument("--forest", type=int, default=512, help="trees of the forest leg (0: skip)")
    ap.add_argument("--tabulated", type=int, default=1, help="run the tabulated user-kernel leg")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="copy 256 MiB device-to-device 3 times before the run (known bytes for PMC scaling)")
    ap.add_argument("--serve", type=int, default=1, help="run the served leg")
    ap.add_argument("--serve-trees", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-load", type=float, default=0.7, help="arrival rate as a fraction of capacity")
    ap.add_argument("--serve-precision", choices=["fp64", "split", "split16"], default="split")
    ap.add_argument("--serve-slots", type=int, default=4)
    return ap.parse_args()


def direct_rel_err(K, x, y, tau, u, m, seed=12):
    """max-norm relative error of potentials u (host, input order) against the FP64 direct sum
    at m sampled targets (GPU, compensated, index self-exclusion)."""
    import ctypes
    import numpy as np
    import torch
    from pykifmm2d_amd import workloads
    from pykifmm2d_amd._lib import check, lib
    from pykifmm2d_amd.device import as_device_f64, empty_f64, ptr, stream_handle
    n = x.size
    idx = workloads.sample_targets(n, m, seed=seed)
    didx = torch.from_numpy(idx.astype(np.int64)).to("cuda")
    dx, dy, dt = as_device_f64(x), as_device_f64(y), as_device_f64(tau)
    tx, ty = as_device_f64(x[idx]), as_device_f64(y[idx])
    ref = empty_f64((idx.size,))
    d = K.desc()
    check(lib().kf_direct(ctypes.byref(d), ptr(dx), ptr(dy), ptr(dt), n, ptr(tx), ptr(ty), ptr(didx), idx.size, 1,
                          ptr(ref), stream_handle()))
    r = ref.cpu().numpy()
    uu = np.asarray(u)[idx]
    return float(np.abs(uu - r).max() / np.abs(r).max()), float(np.linalg.norm(uu - r) / np.linalg.norm(r)), int(idx.size)


def pass_roofline(handle, run_once, steps, nrhs, n, p, st, prec="fp64", flop_eval=8.0, pmc_workload=None):
    """Roofline of one planned-evaluation form.  Per-pass device times come from HIP events
    between the passes (kf_plan_profile), over `steps` calls of run_once() made after the timed
    ones (so the timed steps record no events), with every pass alone on the device.  Algorithmic
    work per evaluation of nrhs right-hand sides from the plan statistics: M2L 2p^2 per V-list
    pair, M2M 8p^2 per box with children, L2L 2p^2 per target below level 2, C2E 2p^2 per row
    (dense-equivalent), P2P / P2M / L2P flop_eval per kernel evaluation (Laplace 8, K0 10).
    The M2L's peak follows the precision: FP64 or FP32 matrix peak, and for "split16" the FP16
    matrix peak over the three FP16 products that make up each FP32-accurate one.  Returns (roofline of the dominant kernel, T_roof in ms per evaluation, pass times, pass
    flops)."""
    import ctypes
    import torch
    from pykifmm2d_amd import _lib
    from pykifmm2d_amd._lib import check, lib
    # each pass is timed alone on the device: the profiled calls run the near field on the
    # evaluation's stream instead of beside the other passes (near_fork 4), and the M2L as one
    # launch (a pipelined M2L, per level on a stream of its own, has no single launch to
    # bracket); the kernels are the same, only their overlap differs
    tuning = _lib.Tuning()
    check(lib().kf_plan_get_tuning(handle, ctypes.byref(tuning)))
    saved = (tuning.m2l_pipeline, tuning.near_fork)
    tuning.m2l_pipeline, tuning.near_fork = 0, 4
    check(lib().kf_plan_set_tuning(handle, ctypes.byref(tuning)))
    prof = (ctypes.c_double * 8)()
    lib().kf_plan_profile(handle, 1, prof, 8)       # reset totals, record events
    torch.cuda.synchronize()
    for _ in range(steps):
        run_once()
    torch.cuda.synchronize()
    lib().kf_plan_profile(handle, 0, prof, 8)       # read totals over those calls
    tuning.m2l_pipeline, tuning.near_fork = saved
    check(lib().kf_plan_set_tuning(handle, ctypes.byref(tuning)))
    nev = max(1.0, prof[6])
    pass_ms = dict(zip(["gather", "p2m", "upward", "downward", "l2p", "p2p"], [prof[i] / nev for i in range(6)]))
    # the downward pass split into its M2L launch and the per-level L2L + C2E chain
    pass_ms["m2l"] = prof[7] / nev
    pass_ms["down_chain"] = pass_ms["downward"] - pass_ms["m2l"]
    pp = float(p) * p
    flops = {  # per evaluation of nrhs right-hand sides; C2E counted dense-equivalent (2p^2 per row)
        "m2l": nrhs * 2 * pp * st["m2l_pairs"],
        "down_chain": nrhs * (2 * pp * st["l2l_boxes"] + 2 * pp * st["c2e_down_rows"]),
        "upward": nrhs * (2 * pp * 4 * st["m2m_boxes"] + 2 * pp * st["c2e_up_rows"]),
        "p2p": flop_eval * nrhs * st["p2p_pairs"],
        "l2p": flop_eval * nrhs * st["l2p_evals"],
        "p2m": flop_eval * nrhs * st["p2m_evals"],
    }
    # path-level roofline: T_roof = sum over passes of max(F / peak, B / HBM peak); bytes: points
    # and densities read once, potentials and expansion rows written once (8 p bytes per box row
    # and right-hand side in FP64)
    row_b = 8.0 * p * nrhs
    hbm_bytes = {
        "p2p": 8.0 * n * (2 + 2 * nrhs), "p2m": 8.0 * n * (2 + nrhs) + row_b * st["c2e_up_rows"],
        "l2p": 8.0 * n * (2 + 2 * nrhs) + row_b * st["leaves"],
        "upward": 2 * row_b * st["c2e_up_rows"], "m2l": row_b * (st["c2e_up_rows"] + st["c2e_down_rows"]),
        "down_chain": 2 * row_b * st["c2e_down_rows"],
    }
    peak_of = {k: PEAK_FP64_TFLOPS for k in flops}
    if prec == "split":
        peak_of["m2l"] = PEAK_FP32_TFLOPS
    elif prec == "split16":
        peak_of["m2l"] = round(PEAK_FP16_TFLOPS / 3, 1)
    t_roof = sum(max(flops[k] / (peak_of[k] * 1e12), hbm_bytes[k] / (PEAK_HBM_GBS * 1e9)) for k in flops)
    # the dominant kernel: the longest of the single-launch passes (the upward pass and the
    # downward chain are sequences of per-level launches)
    dom = max(["p2p", "m2l", "p2m", "l2p"], key=lambda k: pass_ms[k])
    achieved = flops[dom] / (pass_ms[dom] * 1e-3) / 1e12
    # batches of 16 and more (kernel values not stored) contract the near field, P2M and L2P on
    # the MFMA (k_p2p_mfma, k_p2m_mfma, k_l2p_mfma); the kernel evaluations stay on the vector ALU
    batched_mfma = nrhs >= 16 and not st.get("stored", 0) and tuning.batch_form != 1
    kname = PASS_KERNEL[dom] + ({"split": "_f32", "split16": "_f16"}.get(prec, "") if dom == "m2l" else "") + \
        ("_mfma" if batched_mfma and dom != "m2l" else "")
    traffic, prof_tuning = pmc_traffic(kname, pmc_workload) if pmc_workload else (None, None)
    # a VALU-bound pass: the share of the device's VALU issue rate its instructions take
    vinst = pmc_valu_insts(kname, pmc_workload) if pmc_workload and dom != "m2l" else None
    valu_frac = round(vinst / (pass_ms[dom] * 1e-3) / PEAK_VALU_WAVE_INSTS, 4) if vinst else None
    roof = {"bound": "mfma" if dom == "m2l" else "valu", "achieved": round(achieved, 3), "peak": peak_of[dom],
            "unit": "TFLOP/s", "frac": round(achieved / peak_of[dom], 4), "traffic": traffic, "kernel": dom,
            "kernel_name": kname,
            "traffic_source": ("profiles/kifmm_%s_pmc.txt (tuning %s)" % (pmc_workload, prof_tuning)) if traffic
            else None,
            "valu_issue_frac": valu_frac,
            "note": "P2P/P2M/L2P are FP64 vector-ALU passes (MI355X vector FP64 peak = matrix FP64 peak = "
                    "78.6 TFLOP/s); flops per kernel evaluation as defined for the path: %g (a log counts as "
                    "one); valu_issue_frac: the kernel's VALU wave instructions (SQ_INSTS_VALU, from the "
                    "committed PMC summary) per second over the device's issue rate, one wave64 instruction "
                    "per 4 cycles per SIMD" % flop_eval,
            "passes": {k: {"ms": round(pass_ms[k], 4), "tflops": round(flops[k] / max(pass_ms[k], 1e-9) / 1e9, 3)}
                       for k in flops}}
    return roof, 1e3 * t_roof, pass_ms, flops


def batched_leg(tree, x, y, K, p, nrhs, steps, warmup, autotune=True, prec="split"):
    """Split-precision planned evaluation of nrhs right-hand sides per call on the headline tree
    (prec "split": FP32 M2L; "split16": the M2L as FP16 splits on the FP16 matrix cores)."""
    import numpy as np
    import torch
    from pykifmm2d_amd import fmm
    from pykifmm2d_amd.device import empty_f64
    n = x.size
    plan = fmm._Plan(tree, K, p, fmm.PRECISIONS[prec])
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    T = torch.randn(nrhs, n, generator=g, device="cuda", dtype=torch.float64)
    tuning = plan.autotune(nrhs)[0] if autotune else plan.tuning()
    for _ in range(warmup):
        plan.eval(T)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        u = plan.eval(T)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    errs = [direct_rel_err(K, x, y, T[r].cpu().numpy(), u[r].cpu().numpy(), 2000, seed=20 + r)[0]
            for r in (0, nrhs - 1)]
    roof, t_roof_ms, _, _ = pass_roofline(plan.handle, lambda: plan.eval(T), steps, nrhs, n, p, plan.stats(),
                                          prec=prec, pmc_workload="batched_%s_clustered_1M_p64_64rhs" % prec)
    del plan
    return {"nrhs": nrhs, "precision": prec, "tuning": tuning, "ms_per_eval": round(ms, 3),
            "points_per_s": round(n * nrhs / (ms * 1e-3), 1), "rel_err_max": max(errs),
            "err_samples": 2000, "err_rhs": [0, nrhs - 1], "roofline": roof,
            "path_roofline": {"t_roof_ms": round(t_roof_ms, 4), "frac": round(t_roof_ms / ms, 4)}}


def helmholtz_leg(steps=3, warmup=1, n=4_000_000, nrhs=64):
    """Planned modified Helmholtz evaluation of 64 right-hand sides on a 4M-point annulus tree,
    split precision (FP32 M2L), and beside it the same tree with the M2L as FP16 splits."""
    import torch
    from pykifmm2d_amd import fmm, kernels, workloads
    x, y, _ = workloads.annulus(n, seed=3)
    K = kernels.ModifiedHelmholtz(10.0)
    t0 = time.perf_counter()
    tree = fmm.Tree(x, y, 64)
    plan = fmm._Plan(tree, K, 48, fmm.PRECISIONS["split"])
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    T = torch.randn(nrhs, n, generator=g, device="cuda", dtype=torch.float64)

    def run(plan, prec):
        for _ in range(warmup):
            u = plan.eval(T)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            u = plan.eval(T)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        errs = [direct_rel_err(K, x, y, T[r].cpu().numpy(), u[r].cpu().numpy(), 2000, seed=30 + r)[0]
                for r in (0, nrhs - 1)]
        st = plan.stats()
        wl = "mhelmholtz_annulus_4M_p48_64rhs" + ("_split16" if prec == "split16" else "")
        roof, t_roof_ms, _, _ = pass_roofline(plan.handle, lambda: plan.eval(T), steps, nrhs, n, 48, st, prec=prec,
                                              flop_eval=10.0, pmc_workload=wl)
        return st, {"precision": prec, "ms_per_eval": round(ms, 3), "points_per_s": round(n * nrhs / (ms * 1e-3), 1),
                    "rel_err_max": max(errs), "err_samples": 2000, "roofline": roof,
                    "path_roofline": {"t_roof_ms": round(t_roof_ms, 4), "frac": round(t_roof_ms / ms, 4)}}

    st, res = run(plan, "split")
    del plan
    plan16 = fmm._Plan(tree, K, 48, fmm.PRECISIONS["split16"])
    _, res16 = run(plan16, "split16")
    del plan16, tree
    out = {"workload": "mhelmholtz_k10_annulus_4M_x64rhs_split", "n": n, "nrhs": nrhs, "p": 48, "cutoff": 64,
           "build_s": round(build_s, 3), "m2l_pairs": st["m2l_pairs"], "p2p_pairs": st["p2p_pairs"]}
    out.update(res)
    out["split16"] = res16
    return out


def tabulated_leg(tree, x, y, tau, p, steps=10, warmup=3, autotune=True):
    """User kernels as tables (kernels.Tabulated): the headline tree with the Laplace kernel given
    as a function of r^2, and a 1M annulus with K0(10 r)/(2 pi) given as one, beside the built-in
    kernel on the same tree (errors against the FP64 direct sum of the built-in kernel)."""
    import numpy as np
    import scipy.special
    import torch
    from pykifmm2d_amd import fmm, kernels, workloads
    from pykifmm2d_amd.device import as_device_f64

    def run(plan, dt):
        if autotune:
            plan.autotune(1)
        for _ in range(warmup):
            u = plan.eval(dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            u = plan.eval(dt)
        torch.cuda.synchronize()
        return round(1e3 * (time.perf_counter() - t0) / steps, 4), u

    def timed(tr, K, pp, xx, yy, tt, ref):
        """(ms, rel err) matrix-free, then (ms, rel err) with the stored near field."""
        plan = fmm._Plan(tr, K, pp)
        dt = as_device_f64(tt)
        out = []
        for stored in (False, True):
            plan.store_matrices(stored)
            ms, u = run(plan, dt)
            out += [ms, direct_rel_err(ref, xx, yy, tt, u.cpu().numpy(), 2000, seed=40)[0]]
        del plan
        return out

    lap = kernels.Tabulated(lambda r2: (-0.25 / np.pi) * np.log(r2), name="laplace")
    ms, err, ms_s, err_s = timed(tree, lap, p, x, y, tau, kernels.Laplace())
    res = {"laplace_clustered_1M_p64": {"ms_per_eval": ms, "rel_err": err, "degree": lap.deg,
                                        "segments": lap.nseg, "stored_ms_per_eval": ms_s,
                                        "stored_rel_err": err_s}}
    ax, ay, at = workloads.annulus(1_000_000, seed=3)
    atree = fmm.Tree(ax, ay, 64)
    mhk = kernels.ModifiedHelmholtz(10.0)
    k0t = kernels.Tabulated(lambda r2: scipy.special.k0(10.0 * np.sqrt(r2)) / (2.0 * np.pi), name="k0")
    ms_b, err_b, ms_bs, err_bs = timed(atree, mhk, 48, ax, ay, at, mhk)
    ms, err, ms_s, err_s = timed(atree, k0t, 48, ax, ay, at, mhk)
    res["mhelmholtz_k10_annulus_1M_p48"] = {"ms_per_eval": ms, "rel_err": err, "degree": k0t.deg,
                                            "segments": k0t.nseg, "stored_ms_per_eval": ms_s,
                                            "stored_rel_err": err_s, "builtin_ms_per_eval": ms_b,
                                            "builtin_rel_err": err_b, "builtin_stored_ms_per_eval": ms_bs,
                                            "builtin_stored_rel_err": err_bs}
    del atree
    return res


def forest_leg(n_trees, steps=3, warmup=1, seed=4):
    """n_trees independent Laplace problems built and evaluated together, and the same trees'
    single-tree plans evaluated one after another."""
    import numpy as np
    import torch
    from pykifmm2d_amd import fmm, forest, kernels
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(1e3), np.log(2e5), n_trees))).astype(np.int64)
    pts = [(rng.random(n), rng.random(n)) for n in ns]
    K = kernels.Laplace()
    # the points resident in HBM before the build is timed (as the headline's tree build)
    dpts = [(torch.from_numpy(a).to("cuda"), torch.from_numpy(b).to("cuda")) for a, b in pts]
    warm = forest.Forest(dpts[:4], K, p=32, cutoff=50)
    del warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    F = forest.Forest(dpts, K, p=32, cutoff=50)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    N = F.N
    tuning = F.autotune(1)[0]  # launch structure and tile shapes timed on the forest itself
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    tau = torch.randn(N, generator=g, device="cuda", dtype=torch.float64)
    for _ in range(warmup):
        u = F.eval(tau)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        u = F.eval(tau)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    # the same trees one plan at a time
    plans = [fmm._Plan(t, K, 32) for t in F.trees]
    off = F.offsets
    slices = [tau[off[i]:off[i + 1]] for i in range(n_trees)]
    for pl, sl in zip(plans, slices):
        pl.eval(sl)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for pl, sl in zip(plans, slices):
            pl.eval(sl)
    torch.cuda.synchronize()
    ms_loop = 1e3 * (time.perf_counter() - t0) / steps
    uh = u.cpu().numpy()
    errs = []
    for i in (0, int(np.argmax(ns)), n_trees - 1):
        x, y = pts[i]
        errs.append(direct_rel_err(K, x, y, tau[off[i]:off[i + 1]].cpu().numpy(), uh[off[i]:off[i + 1]], 2000)[0])
    st = F.stats()
    roof, t_roof_ms, _, _ = pass_roofline(F.handle, lambda: F.eval(tau), steps, 1, F.n_planned, 32, st,
                                          pmc_workload="laplace_uniform_forest_%d" % n_trees)
    del plans, F, dpts
    return {"roofline": roof, "path_roofline": {"t_roof_ms": round(t_roof_ms, 4), "frac": round(t_roof_ms / ms, 4)},
            "workload": "laplace_uniform_forest_%d" % n_trees, "trees": n_trees, "n_total": int(N), "tuning": tuning,
            "n_min": int(ns.min()), "n_max": int(ns.max()), "p": 32, "cutoff": 50, "precision": "fp64",
            "build_s": round(build_s, 3), "ms_per_eval": round(ms, 3), "points_per_s": round(N / (ms * 1e-3), 1),
            "ms_per_eval_tree_loop": round(ms_loop, 3),
            "points_per_s_tree_loop": round(N / (ms_loop * 1e-3), 1), "rel_err_max": max(errs),
            "err_samples": 2000, "m2l_pairs": st["m2l_pairs"], "p2p_pairs": st["p2p_pairs"]}


def served_leg(args):
    """64 resident star-curve trees behind the continuous-batching request loop."""
    import numpy as np
    import torch
    from pykifmm2d_amd import kernels, serving, workloads
    K = kernels.Laplace()
    curves = workloads.star_curves(args.serve_trees, seed=5)
    t0 = time.perf_counter()
    srv = serving.Server(slots=args.serve_slots)
    tids = [srv.add_tree(cx, cy, K, p=48, cutoff=50, precision=args.serve_precision) for cx, cy in curves]
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    ns = {tid: srv.tree_info(tid)["n"] for tid in tids}
    g = torch.Generator(device="cuda")
    g.manual_seed(5)

    def fresh(count, seed):
        rng = np.random.default_rng(seed)
        which = rng.integers(0, len(tids), count)
        return [torch.randn(ns[tids[w]], generator=g, device="cuda", dtype=torch.float64) for w in which]

    # warm every resident at every batch width it will see
    for tid in tids:
        for _ in range(2):
            srv.submit(tid, torch.randn(ns[tid], generator=g, device="cuda", dtype=torch.float64))
    srv.drain()
    # capacity: a backlog of requests all at once (the loop batches each tree's queue)
    m_cap = min(args.serve_requests, 4000)
    dens = fresh(m_cap, 1)
    torch.cuda.synchronize()
    reqs, wall = serving.run_arrivals(srv, tids, m_cap, lambda i, tid: dens[i], rate=None, seed=1)
    cap_rps = m_cap / wall
    cap_pts = sum(ns[q.tree_id] for q in reqs) / wall
    # roofline: every resident plan profiled on batches of 8 right-hand sides (the request loop's
    # mean batch is about that), passes summed over the trees; the path bound of the capacity
    # run is the sum over its requests of their tree's T_roof per right-hand side
    f_sum, ms_sum, troof_rhs = {}, {}, {}
    for tid in tids:
        res = srv._res[tid]
        if res.plan is None:
            continue
        T8 = torch.randn(8, ns[tid], generator=g, device="cuda", dtype=torch.float64)
        _, troof, pm, fl = pass_roofline(res.plan.handle, lambda: res.plan.eval(T8), 2, 8, ns[tid], 48,
                                         res.plan.stats(), prec=args.serve_precision)
        troof_rhs[tid] = troof / 8
        for k in fl:
            f_sum[k] = f_sum.get(k, 0.0) + fl[k]
            ms_sum[k] = ms_sum.get(k, 0.0) + pm[k]
    dom = max(["p2p", "m2l", "p2m", "l2p"], key=lambda k: ms_sum[k])
    m2l_peak = {"split": PEAK_FP32_TFLOPS, "split16": round(PEAK_FP16_TFLOPS / 3, 1)}.get(args.serve_precision,
                                                                                          PEAK_FP64_TFLOPS)
    peak = m2l_peak if dom == "m2l" else PEAK_FP64_TFLOPS
    ach = f_sum[dom] / (ms_sum[dom] * 1e-3) / 1e12
    served_roof = {"bound": "mfma" if dom == "m2l" else "valu", "achieved": round(ach, 3), "peak": peak,
                   "unit": "TFLOP/s", "frac": round(ach / peak, 4), "traffic": None, "kernel": dom,
                   "kernel_name": PASS_KERNEL[dom] + ("_f32" if dom == "m2l" and split else ""),
                   "note": "each resident plan profiled on batches of 8 right-hand sides; passes summed over trees",
                   "passes": {k: {"ms": round(ms_sum[k], 4), "tflops": round(f_sum[k] / max(ms_sum[k], 1e-9) / 1e9, 3)}
                              for k in f_sum}}
    cap_troof_ms = sum(troof_rhs.get(q.tree_id, 0.0) for q in reqs)
    served_path = {"t_roof_ms": round(cap_troof_ms, 3), "wall_ms": round(1e3 * wall, 3),
                   "frac": round(cap_troof_ms / (1e3 * wall), 4), "over": "the capacity run's requests"}
    del reqs, dens
    # Poisson stream at a fraction of capacity
    rate = args.serve_load * cap_rps
    dens = fresh(args.serve_requests, 2)
    torch.cuda.synchronize()
    # the collector stays on for the timed stream; the objects of the set-up (the other legs,
    # the resident trees) are collected and frozen first, so a generation-2 collection scans only
    # what the stream itself allocates (unfrozen, one took about 70 ms of the loop's thread)
    gc.collect()
    gc.freeze()
    try:
        reqs, wall = serving.run_arrivals(srv, tids, args.serve_requests, lambda i, tid: dens[i], rate=rate, seed=2)
    finally:
        gc.unfreeze()
    lat = serving.latency_summary(reqs)
    pts = sum(ns[q.tree_id] for q in reqs)
    # accuracy of a few served results against the FP64 direct sum
    errs = []
    for q in reqs[:3]:
        cx, cy = curves[tids.index(q.tree_id)]
        errs.append(direct_rel_err(K, cx, cy, q.tau.cpu().numpy(), q.out.cpu().numpy(), 2000)[0])
    return {"workload": "laplace_star_curves_served", "trees": len(tids),
            "points_resident": int(sum(ns.values())), "p": 48, "cutoff": 50,
            "precision": args.serve_precision, "slots": args.serve_slots,
            "reserved_bytes": int(srv.used), "workspace_bytes": int(srv.workspace_bytes),
            "build_s": round(build_s, 2),
            "capacity_requests_per_s": round(cap_rps, 1), "capacity_points_per_s": round(cap_pts, 1),
            "rate_per_s": round(rate, 1), "requests": len(reqs), "wall_s": round(wall, 3),
            "points_per_s": round(pts / wall, 1), "p50_ms": round(lat["p50_ms"], 3),
            "p99_ms": round(lat["p99_ms"], 3), "mean_ms": round(lat["mean_ms"], 3),
            "mean_batch": round(lat["mean_batch"], 2), "rel_err_max": max(errs), "err_samples": 2000,
            "python_gc": "enabled during the timed stream (set-up objects collected and frozen first)",
            "roofline": served_roof, "path_roofline": served_path}


def cpu_baseline(x, y, tau, p, cutoff, min_seconds):
    """The oracle (CPU restatement of the reference, C/OpenMP loops + numpy BLAS/LAPACK) on
    the same workload: planned evaluation build_expansions + self_potential, repeated until
    at least min_seconds of work; points/s = N / mean evaluation time."""
    from threadpoolctl import threadpool_limits

    from oracle import fmm_ref
    F = fmm_ref.FMMRef(x, y, fmm_ref.LAPLACE, p, cutoff)
    F.precompute()
    reps, total = 0, 0.0
    # OpenMP loops on every core, BLAS on one thread: the per-level dense products are small, and
    # BLAS threads competing with the OpenMP loops made the evaluation about 9x slower (8-core
    # container, 100k points: 1.6 s with both at 8 threads, 0.17 s with BLAS on one)
    with threadpool_limits(limits=1, user_api="blas"):
        while total < min_seconds and reps < 20:
            t0 = time.perf_counter()
            F.build_expansions(tau)
            F.self_potential()
            total += time.perf_counter() - t0
            reps += 1
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    # one core: the reference's own throughput metric, points per second per core
    # (examples/laplace2.py:146), from one planned evaluation with the OpenMP loops and BLAS on one
    # thread
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        F.build_expansions(tau)
        F.self_potential()
        one = time.perf_counter() - t0
    return {"value": x.size * reps / total, "unit": "points/s", "cores": cores, "kind": "port",
            "sample": "full workload (N=%d clustered, p=%d, cutoff %d), %d planned evaluations, "
                      "%.1f s; OpenMP threads = cores, BLAS on one thread" % (x.size, p, cutoff, reps, total),
            "per_core": {"value": round(x.size / one, 1), "unit": "points/s/core", "cores": 1,
                         "sample": "one planned evaluation of the full workload on one thread, %.1f s" % one}}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs (--gpus %d) are not part "
              "of it (the sharding design is described in DESIGN.md)" % args.gpus)
        sys.exit(2)
    import numpy as np
    import torch

    from pykifmm2d_amd import fmm, kernels, workloads
    from pykifmm2d_amd.device import as_device_f64, empty_f64, ptr, stream_handle
    from pykifmm2d_amd._lib import check, lib
    import ctypes

    assert torch.cuda.is_available(), "bench.py needs cuda:0"
    n, p, c = args.n, args.p, args.cutoff
    x, y, tau = workloads.clustered(n, seed=2)
    K = kernels.Laplace()
    nrhs = args.nrhs
    if nrhs > 1:
        tau = np.random.default_rng(7).standard_normal((nrhs, n))
        tau[0] = workloads.clustered(n, seed=2)[2]
    dx, dy, dt = as_device_f64(x), as_device_f64(y), as_device_f64(tau)

    if args.copy_calibration:
        src = torch.ones(1 << 25, dtype=torch.float64, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # warm the library (module load, first launches) on a small problem
    wx, wy, wt = workloads.clustered(20000, seed=3)
    fmm._Plan(fmm.Tree(wx, wy, c), K, p).eval(wt)
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    tree = fmm.Tree(dx, dy, c)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plan = fmm._Plan(tree, K, p, fmm.PRECISIONS[args.precision])
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    tree_ms, plan_ms = 1e3 * (t1 - t0), 1e3 * (t2 - t1)
    # rebuilds of the same tree and plan (the library's block cache then holds the build's
    # buffers): the figures of a tree rebuilt for new points of the same size
    rebuild = []
    for _ in range(3):
        r0 = time.perf_counter()
        rt = fmm.Tree(dx, dy, c)
        torch.cuda.synchronize()
        r1 = time.perf_counter()
        rp = fmm._Plan(rt, K, p, fmm.PRECISIONS[args.precision])
        torch.cuda.synchronize()
        rebuild.append((1e3 * (r1 - r0), 1e3 * (time.perf_counter() - r1)))
        del rp, rt
    rebuild.sort(key=lambda tp: tp[0] + tp[1])
    tree_rb_ms, plan_rb_ms = rebuild[1]  # the median of three
    # launch structure for this batch width (kf_plan_autotune; results do not depend on it)
    if args.tuning:
        plan.set_tuning(*[int(v) for v in args.tuning.split(",")])
    tuning = plan.tuning()
    st = plan.stats()

    pot = empty_f64(tuple(dt.shape))
    stream = stream_handle()

    def step():
        check(lib().kf_plan_eval(plan.handle, ptr(dt), n, nrhs, ptr(pot), n, stream))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = 1e3 * (t1 - t0) / args.steps
    autotuned = None
    if args.autotune and not args.tuning:
        chosen, tuner_ms = plan.autotune(nrhs)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        a0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        autotuned = {"tuning": chosen, "ms_per_step": round(1e3 * (time.perf_counter() - a0) / args.steps, 4),
                     "tuner_ms": round(tuner_ms, 4)}
        plan.set_tuning(**tuning)  # per-pass times below: the headline's launch structure
    # error vs FP64 direct sum at sampled targets (GPU, compensated, index self-exclusion)
    tau0 = tau if nrhs == 1 else tau[0]
    rel_err, rel_l2, n_samp = direct_rel_err(K, x, y, tau0, pot.cpu().numpy().reshape(nrhs, n)[0],
                                             args.err_samples)

    # roofline of each pass from algorithmic work (plan statistics) and measured pass time
    workload = "laplace_clustered_1M_p64" if (n, p, c, nrhs, args.precision) == (1_000_000, 64, 64, 1, "fp64") else None
    roof, t_roof_ms, pass_ms, _ = pass_roofline(plan.handle, step, args.steps, nrhs, n, p, st,
                                                prec=args.precision, pmc_workload=workload)
    path_roof = {"t_roof_ms": round(t_roof_ms, 4), "frac": None}

    batched = batched_leg(tree, x, y, K, p, args.batched, 3, 1, bool(args.autotune)) if args.batched > 1 else None
    batched16 = batched_leg(tree, x, y, K, p, args.batched, 3, 1, bool(args.autotune),
                            prec="split16") if args.batched > 1 else None
    tab = tabulated_leg(tree, x, y, tau0, p, autotune=bool(args.autotune)) if args.tabulated and nrhs == 1 else None
    helm = helmholtz_leg() if args.helmholtz else None
    forest_res = forest_leg(args.forest) if args.forest else None
    served = served_leg(args) if args.serve else None
    base = cpu_baseline(x, y, tau0, p, c, args.cpu_seconds) if args.cpu_baseline else None
    value = n * nrhs / (ms_step * 1e-3)
    path_roof["frac"] = round(path_roof["t_roof_ms"] / ms_step, 4)
    line = {
        "metric": "FMM points/s at <=1e-10 rel err (N=1M Laplace, 1 GPU)",
        "value": round(value, 1), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {"workload": "laplace_clustered_1M_planned" + ("_x%drhs" % nrhs if nrhs > 1 else "")
                   + ("_" + args.precision if args.precision != "fp64" else ""),
                   "n": n, "nrhs": nrhs, "p": p, "cutoff": c,
                   "precision": args.precision, "levels": tree.levels, "boxes": tree.n_boxes,
                   "parallelism": "single-gpu", "tuning": tuning},
        "rel_err": rel_err, "rel_err_l2": rel_l2, "err_samples": n_samp,
        "tree_ms": round(tree_ms, 2), "plan_ms": round(plan_ms, 2),
        "rebuild": {"tree_ms": round(tree_rb_ms, 2), "plan_ms": round(plan_rb_ms, 2),
                    "on_the_fly_points_per_s": round(n / ((tree_rb_ms + plan_rb_ms + ms_step) * 1e-3), 1),
                    "note": "median of three rebuilds after the first build (warm block cache)"},
        "plan_stats": st,
        "on_the_fly_points_per_s": round(n / ((tree_ms + plan_ms + ms_step) * 1e-3), 1),
        "autotuned": autotuned,
        "roofline": roof, "path_roofline": path_roof, "cpu_baseline": base, "batched_split": batched,
        "batched_split16": batched16, "helmholtz_64rhs": helm,
        "tabulated": tab,
        "forest": forest_res,
        "served": served,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
