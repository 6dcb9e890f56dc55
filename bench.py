#!/usr/bin/env python
"""bench.py -- planned 2-D KIFMM evaluation throughput on one MI355X.

Workload: Laplace, N = 1,000,000 clustered points (70% in 8 Gaussians with sigma in
[1e-3, 5e-2], 30% uniform), charges N(0,1), p = 64 equivalent points, leaf cutoff 64, FP64.
A step is one planned evaluation u = FMM(tau) over all N points with the tree, operators and
work lists resident in HBM (gather tau -> P2M -> upward -> downward -> L2P + P2P -> scatter).
The error is stated against an FP64 direct sum (compensated, on the GPU) at 10,000 sampled
targets.  The tree + plan build is timed separately and reported with the on-the-fly rate.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (matrix and vector)
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--p", type=int, default=64)
    ap.add_argument("--cutoff", type=int, default=64)
    ap.add_argument("--err-samples", type=int, default=10000)
    ap.add_argument("--nrhs", type=int, default=1, help="right-hand sides per evaluation (batched)")
    ap.add_argument("--precision", choices=["fp64", "split"], default="fp64")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the CPU oracle (port) too")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="minimum CPU baseline work")
    return ap.parse_args()


def cpu_baseline(x, y, tau, p, cutoff, min_seconds):
    """The oracle (CPU restatement of the reference, C/OpenMP loops + numpy BLAS/LAPACK) on
    the same workload: planned evaluation build_expansions + self_potential, repeated until
    at least min_seconds of work; points/s = N / mean evaluation time."""
    from oracle import fmm_ref
    F = fmm_ref.FMMRef(x, y, fmm_ref.LAPLACE, p, cutoff)
    F.precompute()
    reps, total = 0, 0.0
    while total < min_seconds and reps < 20:
        t0 = time.perf_counter()
        F.build_expansions(tau)
        F.self_potential()
        total += time.perf_counter() - t0
        reps += 1
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"value": x.size * reps / total, "unit": "points/s", "cores": cores, "kind": "port",
            "sample": "full workload (N=%d clustered, p=%d, cutoff %d), %d planned evaluations, "
                      "%.1f s; OpenMP threads = cores" % (x.size, p, cutoff, reps, total)}


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
    st = plan.stats()

    pot = empty_f64(tuple(dt.shape))
    stream = stream_handle()

    def step():
        check(lib().kf_plan_eval(plan.handle, ptr(dt), n, nrhs, ptr(pot), n, stream))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    prof = (ctypes.c_double * 7)()
    lib().kf_plan_profile(plan.handle, 1, prof, 7)       # reset totals, record events
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lib().kf_plan_profile(plan.handle, 0, prof, 7)       # read totals over the timed steps
    ms_step = 1e3 * (t1 - t0) / args.steps
    nev = max(1.0, prof[6])
    pass_ms = dict(zip(["gather", "p2m", "upward", "downward", "l2p", "p2p"], [prof[i] / nev for i in range(6)]))

    # error vs FP64 direct sum at sampled targets (GPU, compensated, index self-exclusion)
    idx = workloads.sample_targets(n, args.err_samples)
    didx = torch.from_numpy(idx.astype(np.int64)).to("cuda")
    tx, ty = as_device_f64(x[idx]), as_device_f64(y[idx])
    ref = empty_f64((idx.size,))
    d = K.desc()
    check(lib().kf_direct(ctypes.byref(d), ptr(dx), ptr(dy), ptr(dt.reshape(-1)[:n].contiguous()), n, ptr(tx),
                          ptr(ty), ptr(didx), idx.size, 1, ptr(ref), stream))
    u = pot.cpu().numpy().reshape(nrhs, n)[0][idx]
    r = ref.cpu().numpy()
    rel_err = float(np.abs(u - r).max() / np.abs(r).max())
    rel_l2 = float(np.linalg.norm(u - r) / np.linalg.norm(r))

    # roofline of each pass from algorithmic work (plan statistics) and measured pass time
    pp = float(p) * p
    flops = {  # per evaluation of nrhs right-hand sides
        "downward": nrhs * (2 * pp * (st["m2l_pairs"] + st["l2l_boxes"]) + 4 * pp * st["c2e_down_rows"]),
        "upward": nrhs * (2 * pp * 4 * st["m2m_boxes"] + 4 * pp * st["c2e_up_rows"]),
        "p2p": 8.0 * nrhs * st["p2p_pairs"],
        "l2p": 8.0 * nrhs * st["l2p_evals"],
        "p2m": 8.0 * nrhs * st["p2m_evals"],
    }
    dom = max(flops, key=lambda k: pass_ms[k])
    achieved = flops[dom] / (pass_ms[dom] * 1e-3) / 1e12
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None, "kernel": dom,
            "passes": {k: {"ms": round(pass_ms[k], 4), "tflops": round(flops[k] / (pass_ms[k] * 1e-3) / 1e12, 3)}
                       for k in flops}}

    tau0 = tau if nrhs == 1 else tau[0]
    base = cpu_baseline(x, y, tau0, p, c, args.cpu_seconds) if args.cpu_baseline else None
    value = n * nrhs / (ms_step * 1e-3)
    split = args.precision == "split"
    if split:  # FP32 M2L/P2M/L2P and FP32 MFMA peak for the M2L pass
        roof["peak"] = 157.3 if dom == "downward" else roof["peak"]
        roof["frac"] = round(achieved / roof["peak"], 4)
    line = {
        "metric": "FMM points/s at <=1e-10 rel err (N=1M Laplace, 1 GPU)",
        "value": round(value, 1), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if split else "f64", "data": "synthetic",
        "config": {"workload": "laplace_clustered_1M_planned" + ("_x%drhs" % nrhs if nrhs > 1 else "")
                   + ("_split" if split else ""),
                   "n": n, "nrhs": nrhs, "p": p, "cutoff": c,
                   "precision": args.precision, "levels": tree.levels, "boxes": tree.n_boxes,
                   "parallelism": "single-gpu"},
        "rel_err": rel_err, "rel_err_l2": rel_l2, "err_samples": int(idx.size),
        "tree_ms": round(tree_ms, 2), "plan_ms": round(plan_ms, 2), "plan_stats": st,
        "on_the_fly_points_per_s": round(n / ((tree_ms + plan_ms + ms_step) * 1e-3), 1),
        "roofline": roof, "cpu_baseline": base,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
