"""Host-side logic of the request loop (no GPU): batch-width admission and latency summary."""
import pytest

from pykifmm2d_amd import serving


def test_choose_width_picks_widest_fitting_power_of_two():
    fp = lambda w: 1000 + 100 * w  # fixed part + per-rhs slabs
    assert serving.choose_width(fp, 0, 10**9) == 64
    assert serving.choose_width(fp, 0, 1000 + 100 * 64 - 1) == 32
    assert serving.choose_width(fp, 0, 1000 + 100 * 5) == 4
    assert serving.choose_width(fp, 350, 1000 + 100 * 5) == 1
    assert serving.choose_width(fp, 0, 1099) == 0  # not even one right-hand side
    assert serving.choose_width(fp, 0, 10**9, max_batch=24) == 16


class _Q:
    def __init__(self, a, d, b):
        self.t_arrival, self.t_done, self.batch_size = a, d, b

    done = property(lambda self: self.t_done is not None)
    latency = property(lambda self: self.t_done - self.t_arrival)


def test_latency_summary_percentiles():
    qs = [_Q(0.0, (i + 1) * 1e-3, 2) for i in range(100)] + [_Q(0.0, None, 1)]
    s = serving.latency_summary(qs)
    assert s["count"] == 100
    assert s["p50_ms"] == pytest.approx(50.5)
    assert s["p99_ms"] == pytest.approx(99.01)
    assert s["max_ms"] == pytest.approx(100.0)
    assert s["mean_batch"] == 2.0
    assert serving.latency_summary([]) == {"count": 0}


# ---- scheduling policy of the request loop, on CPU with stand-ins for streams and plans ----

class _FakeEvent:
    def __init__(self, log):
        self.log = log
        self.ready = True

    def record(self, stream):
        self.ready = False
        self.log["events"].append(self)

    def query(self):
        return self.ready


class _FakeTorch:
    float64 = "f64"

    def __init__(self):
        self.log = {"events": []}
        log = self.log

        class _Cuda:
            @staticmethod
            def Stream():
                return type("S", (), {"wait_stream": lambda self, other: None})()

            @staticmethod
            def Event():
                return _FakeEvent(log)

            @staticmethod
            def current_stream():
                return None

            @staticmethod
            def mem_get_info():
                return (1 << 40, 1 << 40)

        self.cuda = _Cuda


class _Vec:
    def __init__(self, n):
        self.n = n

    def reshape(self, *a):
        return self

    def numel(self):
        return self.n


class _FakePlan:
    def __init__(self, tid, running):
        self.tid = tid
        self.running = running
        self.batches = []
        self._ws = None
        self.moves = 0

    def attach_workspace(self, ws):
        self._ws = ws
        self.moves += 1

    def eval_ptrs(self, taus, pots, stream=None):
        assert self.tid not in self.running, "tree dispatched while its batch is in flight"
        assert self._ws is not None, "batch launched without a slot workspace"
        self.running.add(self.tid)
        self.batches.append(len(taus))


def _server_with_fakes(monkeypatch, widths):
    ft = _FakeTorch()
    monkeypatch.setattr(serving, "_torch", lambda: ft)
    monkeypatch.setattr(serving, "as_device_f64", lambda t: t)
    monkeypatch.setattr(serving, "empty_f64", lambda shape: _Vec(shape[0]))
    srv = serving.Server(memory_budget=1 << 30, slots=2)
    running = set()
    for tid, w in enumerate(widths):
        r = serving._Resident()
        r.tid, r.tree, r.kernel, r.n, r.width, r.footprint = tid, None, None, 10, w, 0
        r.plan = _FakePlan(tid, running)
        r.queue = serving.collections.deque()
        r.busy = False
        srv._res[tid] = r
    return srv, ft, running


def _complete_all(srv, running):
    for slot in srv._slots:
        if slot.resident is not None:
            slot.event.ready = True
            running.discard(slot.resident.tid)


def test_loop_runs_one_batch_per_tree_and_batches_the_backlog(monkeypatch):
    srv, ft, running = _server_with_fakes(monkeypatch, [4, 64, 64])
    a = [srv.submit(0, _Vec(10), t_arrival=0.0)]
    srv.poll()                                  # tree 0 goes out with one request
    assert srv.in_flight == 1 and srv._res[0].busy
    # more requests for the busy tree arrive; a slot is free, but the tree's batch is in flight
    a += [srv.submit(0, _Vec(10), t_arrival=1.0 + i) for i in range(6)]
    srv.poll()
    assert srv._res[0].plan.batches == [1] and srv.in_flight == 1
    b = srv.submit(1, _Vec(10), t_arrival=0.5)
    srv.poll()                                  # the free slot takes tree 1, never tree 0 again
    assert srv._res[1].busy and srv._res[0].plan.batches == [1]
    srv.poll()
    assert srv._res[0].plan.batches == [1]      # still in flight: nothing new for tree 0
    _complete_all(srv, running)
    done = srv.poll()                           # both finish; tree 0's backlog goes out, capped at width 4
    assert {id(q) for q in done} == {id(a[0]), id(b)}
    assert srv._res[0].plan.batches == [1, 4]
    _complete_all(srv, running)
    srv.poll()
    assert srv._res[0].plan.batches == [1, 4, 2]
    _complete_all(srv, running)
    rest = srv.drain()
    assert all(q.done for q in a + [b]) and len(rest) == 2
    assert [q.batch_size for q in a] == [1, 4, 4, 4, 4, 2, 2]
    assert srv.pending == 0 and srv.in_flight == 0


def test_loop_serves_oldest_waiting_tree_first(monkeypatch):
    srv, ft, running = _server_with_fakes(monkeypatch, [64, 64, 64])
    srv._slots = srv._slots[:1]
    srv.submit(2, _Vec(10), t_arrival=3.0)
    srv.submit(1, _Vec(10), t_arrival=1.0)
    srv.submit(0, _Vec(10), t_arrival=2.0)
    order = []
    while srv.pending or srv.in_flight:
        srv.poll()
        for slot in srv._slots:
            if slot.resident is not None and not slot.event.ready:
                order.append(slot.resident.tid)
        _complete_all(srv, running)
    assert order == [1, 0, 2]
