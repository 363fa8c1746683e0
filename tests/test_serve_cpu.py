"""The serving scheduler against a simulated backend (no GPU): FIFO
admission under an HBM budget and a slot limit, rejection of requests that
can never fit, isolation of a request whose setup fails, and the trace
generator."""
import time

import numpy as np

from vir_amd.serve import Request, Server, poisson_trace, summarize


class FakeHandle:
    def __init__(self, req, log):
        self.req = req
        self.t_end = time.perf_counter() + 0.002 * req.views
        self.log = log

    def poll(self):
        return time.perf_counter() >= self.t_end

    def result(self):
        return np.full(3, self.req.rid, np.float32)

    def release(self):
        self.log.append(("release", self.req.rid))


class FakeBackend:
    """Bytes = 1000 * n; duration = 2 ms per view; rid 7 fails to start."""

    def __init__(self):
        self.log = []
        self.live = 0
        self.max_live = 0

    def estimate(self, req):
        return 1000 * req.n

    def start(self, req, slot):
        if req.rid == 7:
            raise RuntimeError("setup failed")
        self.log.append(("start", req.rid, slot))
        return FakeHandle(req, self.log)


def _reqs():
    rng = np.random.default_rng(0)
    out = []
    t = 0.0
    for i in range(24):
        t += rng.exponential(0.002)
        out.append(Request(i, "fan", int(rng.choice([2, 4, 8])), int(rng.integers(1, 6)),
                           arrival=t))
    out.append(Request(24, "cone", 50, 2, arrival=t))      # larger than the budget
    return out


def test_budget_fifo_rejection_and_isolation():
    be = FakeBackend()
    srv = Server(be, hbm_budget=12_000, max_active=3)
    reqs = _reqs()
    res = srv.run(reqs)
    assert [r.rid for r in res] == list(range(25))
    assert not res[24].ok and "budget" in res[24].error
    assert not res[7].ok and "setup failed" in res[7].error
    for r in res:
        if r.rid not in (7, 24):
            assert r.ok and r.x[0] == r.rid
            assert r.latency >= 0.002 * reqs[r.rid].views - 1e-4
            assert r.queued >= 0 and r.latency >= r.queued
    assert srv.max_in_use <= 12_000
    assert 1 < srv.max_concurrent <= 3
    assert srv.in_use == 0
    # FIFO: requests start in arrival order
    started = [e[1] for e in be.log if e[0] == "start"]
    assert started == sorted(started)
    # every started request is released exactly once
    rel = [e[1] for e in be.log if e[0] == "release"]
    assert sorted(rel) == sorted(started)
    s = summarize(res, srv.elapsed)
    assert s["completed"] == 23 and s["rejected"] == 2 and s["p99_s"] >= s["p50_s"]


def test_waiting_request_does_not_overtake():
    # a big request at the head waits for memory; smaller ones behind it
    # must not be admitted before it (no starvation of large jobs)
    be = FakeBackend()
    srv = Server(be, hbm_budget=10_000, max_active=4)
    reqs = [Request(0, "fan", 6, 10, arrival=0.0), Request(1, "fan", 8, 2, arrival=0.001),
            Request(2, "fan", 1, 1, arrival=0.002), Request(3, "fan", 1, 1, arrival=0.003)]
    res = srv.run(reqs)
    assert all(r.ok for r in res)
    started = [e[1] for e in be.log if e[0] == "start"]
    assert started == [0, 1, 2, 3]
    assert res[1].queued > 0.015   # waited for request 0 to finish


class KeyedBackend(FakeBackend):
    @staticmethod
    def geometry_key(req):
        return (req.kind, req.n, req.views)


def test_grouped_admission_pulls_same_geometry_forward():
    reqs = [Request(0, "fan", 4, 3), Request(1, "fan", 2, 1), Request(2, "fan", 8, 2),
            Request(3, "fan", 4, 3), Request(4, "fan", 2, 1)]
    be = KeyedBackend()
    srv = Server(be, hbm_budget=100_000, max_active=1, group=True)
    res = srv.run(reqs)
    assert all(r.ok for r in res)
    assert [e[1] for e in be.log if e[0] == "start"] == [0, 3, 1, 4, 2]
    assert srv.grouped == 2
    # without grouping: arrival order
    be = KeyedBackend()
    Server(be, hbm_budget=100_000, max_active=1).run(reqs)
    assert [e[1] for e in be.log if e[0] == "start"] == [0, 1, 2, 3, 4]


class BatchingBackend(KeyedBackend):
    """Cone requests of one (n, views, iters) batch, up to 3 at once; rid 9
    makes its whole batch fail to start."""
    max_batch = 3

    @staticmethod
    def batch_key(req):
        return None if req.kind != "cone" else (req.n, req.views, req.iters)

    def start_batch(self, reqs, slots):
        assert 1 < len(reqs) <= self.max_batch and len(set(slots)) == len(slots)
        assert len({self.batch_key(r) for r in reqs}) == 1
        if any(r.rid == 9 for r in reqs):
            raise RuntimeError("batch setup failed")
        self.log.append(("batch", tuple(r.rid for r in reqs)))
        for r, s in zip(reqs, slots):
            self.log.append(("start", r.rid, s))
        return [FakeHandle(r, self.log) for r in reqs]


def test_batched_admission():
    reqs = [Request(0, "cone", 2, 2), Request(1, "fan", 2, 2), Request(2, "cone", 2, 2),
            Request(3, "cone", 2, 2, iters=5), Request(4, "cone", 2, 2), Request(5, "cone", 2, 2),
            Request(6, "cone", 4, 2)]
    be = BatchingBackend()
    srv = Server(be, hbm_budget=100_000, max_active=8, batch=True)
    res = srv.run(reqs)
    assert all(r.ok and r.x[0] == r.rid for r in res)
    batches = [e[1] for e in be.log if e[0] == "batch"]
    # 0 takes the next two of its key (up to max_batch = 3); 5 is left for
    # a batch of one; other keys and the fan request run alone
    assert batches == [(0, 2, 4)]
    assert srv.batched == 2
    rel = sorted(e[1] for e in be.log if e[0] == "release")
    assert rel == list(range(7)) and srv.in_use == 0
    # the budget caps a batch: with room for two, request 4 waits
    be = BatchingBackend()
    srv = Server(be, hbm_budget=4_000, max_active=8, batch=True)
    assert all(r.ok for r in srv.run(reqs[:5]))
    assert [e[1] for e in be.log if e[0] == "batch"][0] == (0, 2)
    assert srv.max_in_use <= 4_000
    # a failing batch fails all of its members and frees their slots
    reqs = [Request(8, "cone", 2, 2), Request(9, "cone", 2, 2), Request(10, "fan", 2, 2)]
    be = BatchingBackend()
    srv = Server(be, hbm_budget=100_000, max_active=2, batch=True)
    res = srv.run(reqs)
    assert [r.ok for r in res] == [False, False, True]
    assert "batch setup failed" in res[0].error and srv.in_use == 0


def test_poisson_trace_mix():
    tr = poisson_trace(4000, 200.0, seed=5)
    kinds = np.array([r.kind == "fan" for r in tr])
    assert abs(kinds.mean() - 0.6) < 0.03
    gaps = np.diff([r.arrival for r in tr])
    assert np.all(gaps > 0) and abs(gaps.mean() - 1 / 200) < 0.1 / 200
    fan = [r for r in tr if r.kind == "fan"]
    cone = [r for r in tr if r.kind == "cone"]
    assert {r.n for r in fan} == {128, 256, 512, 1024}
    assert {r.n for r in cone} == {64, 128, 256}
    assert min(r.views for r in fan) >= 180 and max(r.views for r in fan) <= 720
    assert min(r.views for r in cone) >= 120 and max(r.views for r in cone) <= 360
