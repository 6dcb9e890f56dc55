"""Request loop over resident planned trees, with continuous batching.

A Server keeps planned trees resident in HBM (tree, operators and work lists) and evaluates
the density vectors that arrive for them.  The expansion slabs a batch works in (check
potentials, equivalent densities, densities and near field in tree order) belong to the stream
slots, not to the trees: each slot has one workspace (kf_workspace_create), sized for the
largest admitted tree at its reserved width, and a batch's plan is attached to the workspace of
the slot it runs on.

* Admission of trees by device footprint.  A tree is planned, then reserved at the widest
  batch (a power of two <= max_batch) whose cost fits what is left of the memory budget: its
  own bytes plus what the slot workspaces must grow by to hold its slabs (kf_plan_reserve on
  each workspace: later evaluations allocate nothing).  A tree that does not fit even one
  right-hand side is refused with AdmissionError.
* Continuous batching.  Requests queue per tree.  Whenever one of the stream slots is free,
  the tree with the oldest waiting request that is not already on the GPU takes the slot,
  with every request it has queued up to its reserved width, in one batched evaluation
  (kf_plan_eval_ptrs: each request keeps its own buffers).  Requests that arrive while others
  run join the next batch of their tree; batches of different trees overlap on the device.
* Latency is measured on the host clock from a request's arrival to the moment the loop sees
  its batch's completion event.

A density's potential does not depend on what else shares its batch: for batches of up to 8
the evaluation's arithmetic for one right-hand side is the same at every width (bitwise the
same alone and in the mix); batches of 16 and more sum the near field, P2M and L2P on the MFMA
in another order (the same to rounding: 1e-12 of max |u| in FP64).

The reference has no request loop: each request is fmm.py:741-814 planned_fmm on one tau.
"""
import collections
import ctypes
import time

import numpy as np

from . import fmm
from . import kernels as _kernels
from ._lib import check, lib
from .device import _torch, as_device_f64, empty_f64, ptr, stream_handle
from .tree import Tree

MAX_BATCH = 64


class AdmissionError(RuntimeError):
    """The device footprint of a tree does not fit the server's memory budget."""


def choose_width(footprint, used, budget, max_batch=MAX_BATCH):
    """Widest power-of-two batch width <= max_batch whose footprint(width) fits budget - used,
    or 0 if even width 1 does not fit."""
    w = 1
    while 2 * w <= max_batch:
        w *= 2
    while w >= 1:
        if used + footprint(w) <= budget:
            return w
        w //= 2
    return 0


class Request:
    """One density vector for one resident tree; `out` holds the potential once `done`."""

    __slots__ = ("tree_id", "tau", "out", "t_arrival", "t_done", "batch_size")

    def __init__(self, tree_id, tau, out, t_arrival):
        self.tree_id = tree_id
        self.tau = tau
        self.out = out
        self.t_arrival = t_arrival
        self.t_done = None
        self.batch_size = 0

    @property
    def done(self):
        return self.t_done is not None

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


class _Resident:
    __slots__ = ("tid", "tree", "plan", "kernel", "n", "width", "footprint", "queue", "busy")


class _Slot:
    __slots__ = ("stream", "event", "resident", "requests", "ws")

    def __init__(self, torch):
        self.stream = torch.cuda.Stream()
        self.event = torch.cuda.Event()
        self.resident = None
        self.requests = None
        self.ws = fmm.Workspace()


class Server:
    """Resident trees + a continuous-batching request loop on one GPU.

    memory_budget: device bytes the resident trees and the slot workspaces may reserve
    (default 80% of the free device memory at construction).  slots: batches in flight at once
    (one HIP stream and one workspace each).  max_batch: widest batch (<= 64).  Batches run as
    launched evaluations: plans on slot workspaces do not replay captured graphs.
    slot_slab_bytes: the most a tree's slabs may take in one slot workspace (default 1.375 GiB):
    a tree whose slabs at width 64 exceed it reserves the widest width that fits, so the slots
    together hold at most slots x this."""

    SLOT_SLABS = 11 << 27  # default cap of one slot workspace: 1.375 GiB

    def __init__(self, memory_budget=None, slots=4, max_batch=MAX_BATCH, slot_slab_bytes=None):
        torch = _torch()
        self._torch = torch
        if memory_budget is None:
            free, _ = torch.cuda.mem_get_info()
            memory_budget = int(0.8 * free)
        if slots < 1:
            raise ValueError("need at least one slot")
        self.budget = int(memory_budget)
        self.slot_slab_bytes = int(self.SLOT_SLABS if slot_slab_bytes is None else slot_slab_bytes)
        self._own = 0  # bytes of the resident trees themselves (workspaces not included)
        self.max_batch = max(1, min(MAX_BATCH, int(max_batch)))
        self._slots = [_Slot(torch) for _ in range(int(slots))]
        self._res = {}
        self._waiting = set()  # trees with queued requests and no batch on the GPU
        self._next_tid = 0
        self.pending = 0  # queued, not yet launched
        self.in_flight = 0  # launched, not yet seen complete

    @property
    def used(self):
        """Device bytes reserved: the resident trees plus the slot workspaces."""
        return self._own + self.workspace_bytes

    @property
    def workspace_bytes(self):
        return sum(slot.ws.nbytes for slot in self._slots)

    # ------------------------------------------------------------------ trees ----------
    def add_tree(self, x, y, kernel=None, p=48, cutoff=50, precision="fp64"):
        """Build and plan a tree of the points (x, y) and keep it resident; returns its id.
        Raises AdmissionError when its footprint does not fit the remaining budget."""
        kernel = _kernels.as_descriptor(kernel if kernel is not None else _kernels.Laplace())
        if kernel.iscomplex:
            raise TypeError("the request loop serves real kernels")
        tree = Tree(x, y, cutoff)
        r = _Resident()
        r.tid = self._next_tid
        r.tree = tree
        r.kernel = kernel
        r.n = tree.N
        r.queue = collections.deque()
        r.busy = False
        # the tree's own arrays (sorted points, permutation) are counted in the plan footprint;
        # the input-order copies of the points it keeps add 16 bytes a point
        inputs = 16 * r.n
        if tree.levels <= 2:
            # too few levels for an expansion: direct sums (fmm.py:229-231), points only
            r.plan = None
            r.footprint = inputs + 20 * r.n
            if self.used + r.footprint > self.budget:
                raise AdmissionError("tree of %d points needs %d bytes, %d of %d left"
                                     % (r.n, r.footprint, self.budget - self.used, self.budget))
            r.width = self.max_batch
        else:
            plan = fmm._Plan(tree, kernel, p, fmm._precision(precision))
            plan.attach_workspace(self._slots[0].ws)  # its footprint then leaves the slabs out
            ws_now = [slot.ws.nbytes for slot in self._slots]

            def cost(k):  # own bytes + what every slot's workspace grows by for its slabs
                need = plan.slab_bytes(k)
                if need > self.slot_slab_bytes and k > 1:
                    return self.budget + 1  # wider than a slot holds
                return inputs + plan.footprint(k) + sum(max(0, need - b) for b in ws_now)

            used = self.used
            w = choose_width(cost, used, self.budget, self.max_batch)
            if w == 0:
                need = cost(1)
                plan.attach_workspace(None)
                del plan
                raise AdmissionError("tree of %d points needs %d bytes, %d of %d left"
                                     % (r.n, need, self.budget - used, self.budget))
            for slot in self._slots:  # every slot can run its batches
                plan.attach_workspace(slot.ws)
                plan.reserve(w)
            r.plan = plan
            r.width = w
            r.footprint = inputs + plan.footprint(w)
        self._own += r.footprint
        self._res[r.tid] = r
        self._next_tid += 1
        return r.tid

    def remove_tree(self, tid):
        """Release an idle resident tree (no queued or running requests)."""
        r = self._res[tid]
        if r.busy or r.queue:
            raise RuntimeError("tree %d still has requests" % tid)
        del self._res[tid]
        self._own -= r.footprint
        r.plan = None
        r.tree = None
        if not any(q.plan is not None for q in self._res.values()):
            # no planned tree left: the workspaces give their memory back
            for slot in self._slots:
                slot.ws = fmm.Workspace()

    def tree_info(self, tid):
        """n, reserved width, footprint (the tree's own bytes), slab_bytes (what a workspace
        holds for it at that width) and planned (False: direct sums)."""
        r = self._res[tid]
        return {"n": r.n, "width": r.width, "footprint": r.footprint,
                "slab_bytes": r.plan.slab_bytes(r.width) if r.plan is not None else 0,
                "planned": r.plan is not None}

    @property
    def trees(self):
        return list(self._res)

    # --------------------------------------------------------------- requests ----------
    def submit(self, tid, tau, out=None, t_arrival=None):
        """Queue a density for tree `tid` (n float64 values in input order, device or host);
        returns its Request.  The density must be ready on torch's current stream."""
        r = self._res[tid]
        tau = as_device_f64(tau).reshape(-1)
        if tau.numel() != r.n:
            raise ValueError("tree %d has %d points, density has %d values" % (tid, r.n, tau.numel()))
        if out is None:
            out = empty_f64((r.n,))
        elif (out.numel() != r.n or not out.is_cuda or out.dtype != self._torch.float64
              or not out.is_contiguous()):
            raise ValueError("out must be a contiguous float64 cuda tensor of %d values" % r.n)
        req = Request(tid, tau, out, time.perf_counter() if t_arrival is None else t_arrival)
        r.queue.append(req)
        if not r.busy:  # a busy tree rejoins the waiting set when its batch completes
            self._waiting.add(tid)
        self.pending += 1
        return req

    def poll(self):
        """One turn of the loop: collect finished batches, launch waiting requests into free
        slots.  Never blocks; returns the requests that completed."""
        done = []
        for slot in self._slots:
            if slot.resident is not None and slot.event.query():
                t = time.perf_counter()
                for q in slot.requests:
                    q.t_done = t
                done.extend(slot.requests)
                self.in_flight -= len(slot.requests)
                slot.resident.busy = False
                if slot.resident.queue:
                    self._waiting.add(slot.resident.tid)
                slot.resident = None
                slot.requests = None
        for slot in self._slots:
            if slot.resident is not None:
                continue
            # one batch per tree at a time: its plan's slabs hold the batch in flight
            best = None
            for tid in self._waiting:
                r = self._res[tid]
                if best is None or r.queue[0].t_arrival < best.queue[0].t_arrival:
                    best = r
            assert best is None or not best.busy
            if best is None:
                break
            self._launch(slot, best)
        return done

    def drain(self):
        """Run the loop until every submitted request is complete; returns them."""
        out = []
        while self.pending or self.in_flight:
            out.extend(self.poll())
        return out

    def _launch(self, slot, r):
        k = min(len(r.queue), r.width)
        reqs = [r.queue.popleft() for _ in range(k)]
        self._waiting.discard(r.tid)  # busy until its batch completes
        slot.stream.wait_stream(self._torch.cuda.current_stream())
        if r.plan is not None:
            if r.plan._ws is not slot.ws:
                r.plan.attach_workspace(slot.ws)
            r.plan.eval_ptrs([q.tau for q in reqs], [q.out for q in reqs], stream=slot.stream)
        else:
            d = r.kernel.desc()
            xs, ys = r.tree._dx, r.tree._dy
            for q in reqs:
                check(lib().kf_direct(ctypes.byref(d), ptr(xs), ptr(ys), ptr(q.tau), r.n, ptr(xs), ptr(ys), None,
                                      r.n, 1, ptr(q.out), stream_handle(slot.stream)))
        slot.event.record(slot.stream)
        for q in reqs:
            q.batch_size = k
        r.busy = True
        slot.resident = r
        slot.requests = reqs
        self.pending -= k
        self.in_flight += k


def latency_summary(requests):
    """p50/p99/mean/max latency in milliseconds and the mean batch size of completed requests."""
    lat = np.array([q.latency for q in requests if q.done], dtype=np.float64) * 1e3
    if lat.size == 0:
        return {"count": 0}
    return {"count": int(lat.size), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "max_ms": float(lat.max()),
            "mean_batch": float(np.mean([q.batch_size for q in requests if q.done]))}


def run_arrivals(server, tids, n_requests, density, rate=None, seed=0):
    """Drive the loop with n_requests requests, each for a tree drawn uniformly from `tids`
    with density density(i, tid) (a device tensor), arriving as a Poisson process of `rate`
    per second (rate None: all at once).  Returns (requests, wall seconds from the first
    arrival to the last completion)."""
    rng = np.random.default_rng(seed)
    which = rng.integers(0, len(tids), n_requests)
    if rate is None:
        arrivals = np.zeros(n_requests)
    else:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    reqs = []
    done = 0
    i = 0
    t0 = time.perf_counter()
    while done < n_requests:
        now = time.perf_counter() - t0
        while i < n_requests and arrivals[i] <= now:
            tid = tids[which[i]]
            reqs.append(server.submit(tid, density(i, tid), t_arrival=t0 + arrivals[i]))
            i += 1
        done += len(server.poll())
    return reqs, time.perf_counter() - t0
