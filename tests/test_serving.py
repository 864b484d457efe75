"""Host logic of the serving loop (admission, continuous batching, ordering),
driven by a simulated executor on a virtual clock -- no GPU needed."""

import numpy as np
import pytest

from glass_amd.serving import Scheduler, SimRequest, poisson_requests, run, summarize


class VClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class SimExecutor:
    """Launch time = per-shell cost x shells, lanes run concurrently, a lane's
    launches run one after another (stream order)."""

    def __init__(self, clock, cap=8, lane_bytes=100, shell_bytes=10, cost=0.01):
        self.clock, self.cap, self.lb, self.sb, self.cost = clock, cap, lane_bytes, shell_bytes, cost
        self.lane_free = {}
        self.open = set()
        self.log = []
        self.used_trace = []
        self.closed_lanes = []

    def lane_key(self, req):
        return (req.nside, req.lmax)

    def lane_capacity(self, key):
        return self.cap

    def lane_bytes(self, key):
        return self.lb

    def request_bytes(self, req):
        return req.nshell * self.sb

    def open_lane(self, key):
        assert key not in self.open
        self.open.add(key)

    def close_lane(self, key):
        self.open.remove(key)
        self.closed_lanes.append(key)

    def open_request(self, req):
        return {"req": req, "chunks": []}

    def close_request(self, st, keep):
        st["closed"] = True

    def launch(self, key, chunks):
        assert key in self.open
        now = self.clock.t
        start = max(now, self.lane_free.get(key, 0.0))
        n = sum(c[2] for c in chunks)
        assert 0 < n <= self.cap
        done = start + self.cost * n
        self.lane_free[key] = done
        for st, s0, k in chunks:
            st["chunks"].append((s0, k))
        self.log.append((key, [(st["req"].rid, s0, k) for st, s0, k in chunks], start, done))
        return done

    def poll(self, token):
        return self.clock.t >= token

    def finish(self, token):
        pass


def test_poisson_requests_match_the_workload():
    reqs = poisson_requests(rate=5.0, duration=600.0, seed=4)
    n = len(reqs)
    assert abs(n - 3000) < 5 * np.sqrt(3000)
    assert reqs == poisson_requests(rate=5.0, duration=600.0, seed=4)
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0) and arr[-1] < 600.0
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert abs(gaps.mean() - 0.2) < 0.02
    nsides = np.array([r.nside for r in reqs])
    for ns, w in zip((256, 512, 1024, 2048, 4096), (0.3, 0.3, 0.2, 0.15, 0.05)):
        frac = np.mean(nsides == ns)
        assert abs(frac - w) < 5 * np.sqrt(w * (1 - w) / n)
    shells = np.array([r.nshell for r in reqs])
    assert shells.min() >= 10 and shells.max() <= 80
    assert all(r.lmax == 2 * r.nside and r.ncorr == 3 for r in reqs)
    assert len({r.seed for r in reqs}) == n


def test_request_spectra_are_a_valid_chain():
    r = SimRequest(0, 0.0, 8, 5, 16, ncorr=3)
    gls = r.spectra()
    assert len(gls) == 5 * 6 // 2
    # GLASS order g00, g11, g10, g22, g21, g20, ...: neighbour correlations
    # 0.6 / 0.3 / 0.1, nothing beyond ncorr
    assert np.allclose(gls[2], 0.6 * np.sqrt(gls[0] * gls[1]))
    assert np.allclose(gls[5], 0.3 * np.sqrt(gls[0] * gls[3]))
    g1 = SimRequest(0, 0.0, 8, 5, 16, ncorr=1).spectra()
    assert len(g1[4]) == 17 and len(g1[5]) == 0


def _drive(reqs, ex, clock, budget, depth=2):
    sched = run(reqs, ex, budget, depth=depth, clock=clock, sleep=clock.sleep, idle_wait=1e-3)
    return sched


def test_all_requests_complete_in_chain_order_and_batched():
    clock = VClock()
    ex = SimExecutor(clock, cap=8)
    reqs = [SimRequest(i, 0.01 * i, 64 if i % 2 else 32, 5 + i, 128 if i % 2 else 64)
            for i in range(10)]
    sched = _drive(reqs, ex, clock, budget=1e9)
    assert sorted(c.rid for c in sched.completed) == list(range(10))
    for c in sched.completed:
        r = reqs[c.rid]
        assert c.latency >= ex.cost * r.nshell - 1e-12
    # every request's shells were launched once, in order, as contiguous chunks
    per = {}
    for key, chunks, start, done in ex.log:
        for rid, s0, k in chunks:
            per.setdefault(rid, []).append((s0, k))
    for rid, ch in per.items():
        pos = 0
        for s0, k in ch:
            assert s0 == pos
            pos += k
        assert pos == reqs[rid].nshell
    # continuous batching: some launch packed chunks of several requests
    assert sched.max_requests_per_launch >= 2
    # lanes of different geometries overlapped in (virtual) device time
    by_key = {}
    for key, _, start, done in ex.log:
        by_key.setdefault(key, []).append((start, done))
    (a, b) = list(by_key.values())[:2]
    assert any(s1 < d2 and s2 < d1 for s1, d1 in a for s2, d2 in b)


def test_admission_respects_budget_and_arrival_order():
    clock = VClock()
    ex = SimExecutor(clock, cap=4, lane_bytes=100, shell_bytes=10)
    # each request needs 10 shells x 10 = 100 bytes; budget: lane + 2 requests
    reqs = [SimRequest(i, 0.0, 16, 10, 32) for i in range(6)]
    sched = Scheduler(ex, budget_bytes=300)
    for r in reqs:
        sched.submit(r)
    admitted_order = []
    orig_open = ex.open_request

    def spy(req):
        admitted_order.append(req.rid)
        return orig_open(req)

    ex.open_request = spy
    while sched.busy():
        sched.step(clock.t)
        assert sched.used <= sched.budget
        clock.sleep(1e-3)
    assert admitted_order == list(range(6))
    assert sched.peak <= 300
    assert len(sched.completed) == 6


def test_request_larger_than_budget_is_rejected_others_served():
    clock = VClock()
    ex = SimExecutor(clock, cap=4, lane_bytes=100, shell_bytes=10)
    reqs = [SimRequest(0, 0.0, 16, 5, 32), SimRequest(1, 0.0, 16, 500, 32),
            SimRequest(2, 0.0, 16, 5, 32)]
    sched = _drive(reqs, ex, clock, budget=400)
    assert sched.rejected == [1]
    assert sorted(c.rid for c in sched.completed) == [0, 2]


def test_idle_lane_is_closed_under_memory_pressure():
    clock = VClock()
    ex = SimExecutor(clock, cap=4, lane_bytes=100, shell_bytes=10)
    reqs = [SimRequest(0, 0.0, 16, 5, 32), SimRequest(1, 1.0, 32, 5, 64)]
    # budget fits one lane + one request only
    sched = _drive(reqs, ex, clock, budget=160)
    assert sorted(c.rid for c in sched.completed) == [0, 1]
    assert ex.closed_lanes == [(16, 32)]


def test_summary_reports_latency_percentiles_and_throughput():
    clock = VClock()
    ex = SimExecutor(clock, cap=16, cost=0.002)
    reqs = poisson_requests(rate=20.0, duration=2.0, seed=1, nsides=(8, 16), weights=(0.5, 0.5),
                            shells=(2, 6))
    sched = _drive(reqs, ex, clock, budget=1e9)
    s = summarize(sched)
    assert s["requests"] == len(reqs) and s["rejected"] == 0
    assert s["maps"] == sum(r.nshell for r in reqs)
    assert 0 < s["latency_s"]["p50"] <= s["latency_s"]["p99"] <= s["latency_s"]["max"]
    assert s["maps_per_s"] > 0


@pytest.mark.parametrize("depth", [1, 2, 3])
def test_inflight_depth_is_bounded(depth):
    clock = VClock()
    ex = SimExecutor(clock, cap=2, cost=0.05)
    reqs = [SimRequest(i, 0.0, 16, 8, 32) for i in range(3)]
    sched = Scheduler(ex, 1e9, depth=depth)
    for r in reqs:
        sched.submit(r)
    while sched.busy():
        sched.step(clock.t)
        for ln in sched.lanes.values():
            assert len(ln.inflight) <= depth
        clock.sleep(1e-3)
    assert len(sched.completed) == 3


def test_executor_placement_failure_makes_the_request_wait():
    """open_request / open_lane raising MemoryError (no room in the executor's
    arena although the byte budget has room) delays the request until earlier
    ones complete; every request is still served, in arrival order."""
    clock = VClock()
    ex = SimExecutor(clock, cap=4, lane_bytes=100, shell_bytes=10)
    live = set()
    orig_open, orig_close = ex.open_request, ex.close_request
    refused = []

    def open_request(req):
        if len(live) >= 1:  # room for one request's buffers at a time
            refused.append(req.rid)
            raise MemoryError("no room")
        live.add(req.rid)
        return orig_open(req)

    def close_request(st, keep):
        live.discard(st["req"].rid)
        orig_close(st, keep)

    ex.open_request, ex.close_request = open_request, close_request
    reqs = [SimRequest(i, 0.0, 16, 6, 32) for i in range(4)]
    sched = _drive(reqs, ex, clock, budget=1e9)
    assert [c.rid for c in sched.completed] == [0, 1, 2, 3]
    assert refused and sched.rejected == []


def test_arena_first_fit_and_merge():
    """The device executor's arena bookkeeping (offsets only, no device)."""
    from glass_amd.serving import Arena

    a = Arena.__new__(Arena)
    a.size, a.free, a.taken = 40960, [(0, 40960)], {}
    lo = a.alloc(5000)
    hi = a.alloc(4096, high=True)
    assert lo == 0 and hi == 40960 - 4096 and a.taken[lo] == 8192
    mid = a.alloc(8192)
    assert mid == 8192
    a.release(lo)
    assert a.alloc(20000) == 16384  # first range large enough
    a.release(mid)
    a.release(16384)
    a.release(hi)
    assert a.free == [(0, 40960)] and not a.taken
    assert a.alloc(50000) is None
