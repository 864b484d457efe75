"""Serving loop for field-generation requests on one MI355X.

A request is one ``glass.generate`` call (glass/fields.py:839-894) of the
lognormal kind: a chain of ``nshell`` correlated shells at one (nside, lmax)
with ``ncorr`` neighbour correlations, to be turned into maps.  Requests arrive
over time; the server

  * admits a request (strict arrival order) only when its resident state --
    its covariance rows, iternorm state and its output maps -- fits the HBM
    budget, together with the geometry lane it needs;
  * keeps every admitted request's state on the device until it completes;
  * batches continuously: each geometry (nside, lmax) has a lane (a HIP
    stream, a plan and a batch buffer of ``capacity`` shells); whenever a lane
    has fewer than ``depth`` launches in flight, the next shells of all its
    admitted requests are packed into one launch (iternorm and correlated alm
    per request chunk, then ONE Legendre launch and ONE ring launch over the
    packed columns), so requests that arrive while others run join the next
    launch, and lanes of different geometries overlap on the device;
  * reports per-request latency (arrival to maps ready in HBM).

On the device, every lane buffer and every request's output maps are placed
in ONE arena of the budget's size, allocated once (``Arena``): lanes from its
top, requests from its bottom, freed ranges merged with their neighbours.
Byte accounting alone is not enough at this scale (a request of 80 maps at
nside 4096 holds 64 GB): with the device allocator's cache, ranges freed by
requests of other sizes fragment the memory until a large placement fails
even though the budget has room.  A placement that does not fit the arena
now makes the request wait, like one that does not fit the budget.

The chain order of a request is kept because all its chunks go through one
lane stream in order, and the iternorm state is carried between chunks.  The
device normals are keyed by (request seed, shell, alm index), so a request's
maps do not depend on what it was batched with (up to the FP32 rounding of the
ring stage, which transforms two shells in one complex FFT).

The scheduling logic is device-independent (``Scheduler``); ``DeviceExecutor``
runs the launches through the C ABI.
"""

from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np

from . import _lib
from .batch import shell_spectra


@dataclasses.dataclass
class SimRequest:
    """A simulation request: a chain of correlated lognormal shells."""

    rid: int
    arrival: float  # seconds from the start of the run
    nside: int
    nshell: int
    lmax: int
    ncorr: int = 3
    seed: int = 0
    tilt: float = -1.2
    amp: tuple = (1.0, 4.0)  # spectrum amplitude, linear from first to last shell
    rho: tuple = (0.6, 0.3, 0.1)
    scale: float = 1500.0
    shift: float = 1.0

    def spectra(self):
        a0, a1 = self.amp
        n = self.nshell
        return shell_spectra(n, self.lmax, amp=lambda i: a0 + (a1 - a0) * i / max(n - 1, 1),
                             rho=self.rho[: self.ncorr], scale=self.scale, tilt=self.tilt)


def poisson_requests(rate: float = 5.0, duration: float = 600.0, seed: int = 4,
                     nsides=(256, 512, 1024, 2048, 4096),
                     weights=(0.3, 0.3, 0.2, 0.15, 0.05), shells=(10, 80),
                     ncorr: int = 3, lmax_factor: int = 2) -> list[SimRequest]:
    """Poisson arrivals at ``rate`` per second over ``duration`` seconds; nside
    drawn with ``weights``, shell count uniform in ``shells`` (inclusive),
    lmax = lmax_factor * nside."""
    rng = np.random.default_rng(seed)
    out: list[SimRequest] = []
    t = 0.0
    p = np.asarray(weights, dtype=float)
    p = p / p.sum()
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        nside = int(nsides[int(rng.choice(len(nsides), p=p))])
        S = int(rng.integers(shells[0], shells[1] + 1))
        out.append(SimRequest(len(out), t, nside, S, lmax_factor * nside, ncorr=ncorr,
                              seed=int(rng.integers(1, 2**31))))
    return out


@dataclasses.dataclass
class _Active:
    req: SimRequest
    state: object
    nbytes: int
    next_shell: int = 0
    done_shells: int = 0
    t_admit: float = 0.0


@dataclasses.dataclass
class _Lane:
    key: tuple
    capacity: int
    nbytes: int
    active: list = dataclasses.field(default_factory=list)
    inflight: collections.deque = dataclasses.field(default_factory=collections.deque)


@dataclasses.dataclass
class Completion:
    rid: int
    arrival: float
    admitted: float
    done: float
    nshell: int
    nside: int

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class Scheduler:
    """Admission against an HBM budget plus continuous batching per geometry.

    ``executor`` provides: lane_key(req), lane_capacity(key), lane_bytes(key),
    request_bytes(req), open_lane(key), close_lane(key), open_request(req),
    close_request(state, keep), launch(key, chunks) -> token, poll(token) ->
    bool, finish(token) (called once when a launch has completed); and
    optionally bind(budget_bytes) (called once, before anything is opened).
    open_lane / open_request may raise MemoryError when the executor cannot
    place the buffers now: the request then waits (idle lanes are closed
    first), as when it does not fit the budget.
    """

    def __init__(self, executor, budget_bytes: float, depth: int = 2, keep_results: bool = False):
        self.ex = executor
        bind = getattr(executor, "bind", None)
        if bind is not None:
            bind(float(budget_bytes))
        self.budget = float(budget_bytes)
        self.depth = int(depth)
        self.keep = keep_results
        self.used = 0.0
        self.peak = 0.0
        self.waiting: collections.deque = collections.deque()
        self.lanes: dict = {}
        self.completed: list[Completion] = []
        self.rejected: list[int] = []
        self.launches = 0
        self.max_requests_per_launch = 0

    # -- arrivals and admission -------------------------------------------
    def submit(self, req: SimRequest) -> None:
        self.waiting.append(req)

    def _admit(self, now: float) -> bool:
        did = False
        while self.waiting:
            req = self.waiting[0]
            key = self.ex.lane_key(req)
            need = self.ex.request_bytes(req)
            lane_need = 0 if key in self.lanes else self.ex.lane_bytes(key)
            if need + lane_need > self.budget:
                self.waiting.popleft()  # can never fit
                self.rejected.append(req.rid)
                did = True
                continue
            if self.used + need + lane_need > self.budget:
                if not self._close_idle_lanes(keep=key):
                    break
                lane_need = 0 if key in self.lanes else self.ex.lane_bytes(key)
                if self.used + need + lane_need > self.budget:
                    break
            if key not in self.lanes:
                if not self._open_lane(key, lane_need):
                    break
            try:
                state = self.ex.open_request(req)
            except MemoryError:
                # no room in the executor's arena now: wait for completions
                # (an idle lane of another geometry is released first)
                if self._close_idle_lanes(keep=key):
                    continue
                break
            self.waiting.popleft()
            self.lanes[key].active.append(_Active(req, state, need, t_admit=now))
            self.used += need
            self.peak = max(self.peak, self.used)
            did = True
        return did

    def _open_lane(self, key, lane_need) -> bool:
        for attempt in range(2):
            try:
                self.ex.open_lane(key)
            except MemoryError:
                if attempt == 0 and self._close_idle_lanes(keep=key):
                    continue
                return False
            self.lanes[key] = _Lane(key, self.ex.lane_capacity(key), lane_need)
            self.used += lane_need
            return True
        return False

    def _close_idle_lanes(self, keep) -> bool:
        freed = False
        for key in [k for k, ln in self.lanes.items()
                    if k != keep and not ln.active and not ln.inflight]:
            ln = self.lanes.pop(key)
            self.ex.close_lane(key)
            self.used -= ln.nbytes
            freed = True
        return freed

    # -- launches -------------------------------------------------------------
    def _launch(self) -> bool:
        did = False
        for key, ln in self.lanes.items():
            while len(ln.inflight) < self.depth:
                chunks, room = [], ln.capacity
                for a in ln.active:
                    left = a.req.nshell - a.next_shell
                    if left <= 0 or room <= 0:
                        continue
                    n = min(left, room)
                    chunks.append((a, a.next_shell, n))
                    room -= n
                if not chunks:
                    break
                token = self.ex.launch(key, [(a.state, s0, n) for a, s0, n in chunks])
                for a, s0, n in chunks:
                    a.next_shell = s0 + n
                ln.inflight.append((token, chunks))
                self.launches += 1
                self.max_requests_per_launch = max(self.max_requests_per_launch, len(chunks))
                did = True
        return did

    # -- completions ------------------------------------------------------------
    def _reap(self, now: float) -> bool:
        did = False
        for ln in self.lanes.values():
            while ln.inflight and self.ex.poll(ln.inflight[0][0]):
                token, chunks = ln.inflight.popleft()
                self.ex.finish(token)
                did = True
                for a, s0, n in chunks:
                    a.done_shells += n
                    if a.done_shells == a.req.nshell:
                        ln.active.remove(a)
                        self.ex.close_request(a.state, self.keep)
                        self.used -= a.nbytes
                        self.completed.append(Completion(a.req.rid, a.req.arrival, a.t_admit, now,
                                                         a.req.nshell, a.req.nside))
        return did

    def step(self, now: float) -> bool:
        """One scheduling round: completions, admissions, launches."""
        d1 = self._reap(now)
        d2 = self._admit(now)
        d3 = self._launch()
        return d1 or d2 or d3

    def busy(self) -> bool:
        return bool(self.waiting) or any(ln.active or ln.inflight for ln in self.lanes.values())


def run(requests, executor, budget_bytes: float, *, depth: int = 2, keep_results: bool = False,
        clock=time.perf_counter, sleep=time.sleep, idle_wait: float = 2e-4,
        progress=None) -> Scheduler:
    """Real-time driver: requests are submitted at their arrival times (seconds
    from the start); returns the scheduler with its completions."""
    sched = Scheduler(executor, budget_bytes, depth=depth, keep_results=keep_results)
    reqs = sorted(requests, key=lambda r: r.arrival)
    i = 0
    t0 = clock()
    next_report = 10.0
    while i < len(reqs) or sched.busy():
        now = clock() - t0
        while i < len(reqs) and reqs[i].arrival <= now:
            sched.submit(reqs[i])
            i += 1
        if not sched.step(now):
            sleep(idle_wait)
        if progress is not None and now >= next_report:
            progress(now, len(sched.completed), len(reqs))
            next_report += 10.0
    sched.wall = clock() - t0
    close = getattr(executor, "close", None)
    if close is not None:
        close()
    return sched


def summarize(sched: Scheduler, t_end: float | None = None) -> dict:
    lat = np.asarray([c.latency for c in sched.completed], dtype=float)
    maps = sum(c.nshell for c in sched.completed)
    if t_end is None:
        t_end = max((c.done for c in sched.completed), default=0.0)
    t_begin = min((c.arrival for c in sched.completed), default=0.0)
    span = max(t_end - t_begin, 1e-9)
    return {
        "requests": len(sched.completed), "rejected": len(sched.rejected), "maps": int(maps),
        "maps_per_s": maps / span,
        "latency_s": {"p50": float(np.percentile(lat, 50)) if lat.size else None,
                      "p99": float(np.percentile(lat, 99)) if lat.size else None,
                      "max": float(lat.max()) if lat.size else None},
        "launches": sched.launches, "max_requests_per_launch": sched.max_requests_per_launch,
        "peak_hbm_bytes": sched.peak, "budget_bytes": sched.budget,
    }


# -- the device executor -------------------------------------------------------

def _nalm(lmax: int) -> int:
    return (lmax + 1) * (lmax + 2) // 2


DEFAULT_CAPACITY = {256: 128, 512: 128, 1024: 128, 2048: 64, 4096: 16}


class _ReqState:
    pass


class Arena:
    """Placement of buffers in one device allocation of ``nbytes``: first fit
    from the bottom (``high=False``) or the top (``high=True``), 4 KiB
    aligned; released ranges merge with free neighbours."""

    ALIGN = 4096

    def __init__(self, nbytes: int, device="cuda"):
        T = _lib.torch()
        self.size = int(nbytes) // self.ALIGN * self.ALIGN
        self.buf = T.empty(self.size, dtype=T.uint8, device=device)
        self.free: list = [(0, self.size)]  # (offset, length), sorted by offset
        self.taken: dict = {}

    def alloc(self, nbytes: int, high: bool = False):
        n = max(1, -(-int(nbytes) // self.ALIGN)) * self.ALIGN
        order = range(len(self.free) - 1, -1, -1) if high else range(len(self.free))
        for i in order:
            o, ln = self.free[i]
            if ln < n:
                continue
            off = o + ln - n if high else o
            rest = (o, ln - n) if high else (o + n, ln - n)
            if rest[1]:
                self.free[i] = rest
            else:
                del self.free[i]
            self.taken[off] = n
            return off
        return None

    def release(self, off: int) -> None:
        n = self.taken.pop(off)
        i = 0
        while i < len(self.free) and self.free[i][0] < off:
            i += 1
        self.free.insert(i, (off, n))
        # merge with the right, then the left neighbour
        if i + 1 < len(self.free) and off + n == self.free[i + 1][0]:
            self.free[i] = (off, n + self.free[i + 1][1])
            del self.free[i + 1]
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] == off:
            self.free[i - 1] = (self.free[i - 1][0], self.free[i - 1][1] + self.free[i][1])
            del self.free[i]

    def tensor(self, off: int, shape, dtype):
        T = _lib.torch()
        numel = int(np.prod(shape))
        nb = numel * T.empty(0, dtype=dtype).element_size()
        return self.buf[off: off + nb].view(dtype).view(tuple(shape))

    def largest_free(self) -> int:
        return max((ln for _, ln in self.free), default=0)


class DeviceExecutor:
    """Runs the scheduler's launches on the GPU through the C ABI."""

    def __init__(self, capacity: dict | None = None):
        _lib.require_gpu()
        self.T = _lib.torch()
        self.capacity = dict(DEFAULT_CAPACITY if capacity is None else capacity)
        self.lanes: dict = {}
        self.results: dict = {}
        self.arena = None
        nd = _lib.ctypes.c_int32()
        self._nd = {}
        for k in range(0, 9):
            _lib.call("ga_iternorm_state_doubles", k, _lib.ctypes.byref(nd))
            self._nd[k] = nd.value

    # sizes
    def lane_key(self, req: SimRequest):
        return (int(req.nside), int(req.lmax))

    def lane_capacity(self, key) -> int:
        nside, _ = key
        return int(self.capacity.get(nside, max(8, 128 * 1024 // max(nside, 1))))

    def lane_bytes(self, key) -> int:
        nside, lmax = key
        cap = self.lane_capacity(key)
        plan = _lib.get_plan(nside, lmax)
        return (plan.table_bytes() + plan.workspace_bytes(cap, 4) + _nalm(lmax) * cap * 8
                + cap * 12 * nside * nside * 4)

    def request_bytes(self, req: SimRequest) -> int:
        nl = req.lmax + 1
        k = req.ncorr
        return (req.nshell * 12 * req.nside**2 * 4 + 2 * req.nshell * nl * (k + 1) * 8
                + nl * self._nd[k] * 8)

    # the arena
    def bind(self, budget_bytes: float) -> None:
        """One arena of the budget's size for every lane and request buffer."""
        self.close()
        self.arena = Arena(int(budget_bytes))

    def close(self) -> None:
        for key in list(self.lanes):
            self.lanes[key].stream.synchronize()
        self.lanes.clear()
        self.arena = None
        self.T.cuda.empty_cache()

    def _place(self, parts, high: bool):
        """Offsets in the arena for [(shape, dtype)], or MemoryError (nothing
        kept) when they do not all fit now."""
        T = self.T
        offs = []
        for shape, dtype in parts:
            nb = int(np.prod(shape)) * T.empty(0, dtype=dtype).element_size()
            off = self.arena.alloc(nb, high=high)
            if off is None:
                for o in offs:
                    self.arena.release(o)
                raise MemoryError(f"arena: no free range of {nb} bytes "
                                  f"(largest {self.arena.largest_free()})")
            offs.append(off)
        return offs, [self.arena.tensor(o, shape, dtype) for o, (shape, dtype) in zip(offs, parts)]

    # lanes
    def open_lane(self, key) -> None:
        T = self.T
        if self.arena is None:
            raise RuntimeError("DeviceExecutor: bind(budget) before opening lanes")
        nside, lmax = key
        cap = self.lane_capacity(key)
        ln = _ReqState()
        ln.plan = _lib.get_plan(nside, lmax)
        ln.cap = cap
        ln.offs, (ln.alm, ln.ws, ln.maps) = self._place(
            [((_nalm(lmax) * cap,), T.complex64), ((ln.plan.workspace_bytes(cap, 4),), T.uint8),
             ((cap, 12 * nside * nside), T.float32)], high=True)
        ln.stream = T.cuda.Stream()
        self.lanes[key] = ln

    def close_lane(self, key) -> None:
        ln = self.lanes.pop(key)
        ln.stream.synchronize()
        for o in ln.offs:
            self.arena.release(o)

    # requests
    def open_request(self, req: SimRequest):
        from . import fields as _fields

        T = self.T
        ln = self.lanes[self.lane_key(req)]
        st = _ReqState()
        st.req = req
        st.key = self.lane_key(req)
        # the output maps first: if they do not fit, nothing else is made
        st.offs, (st.maps,) = self._place([((req.nshell, 12 * req.nside**2), T.float32)],
                                          high=False)
        nl = req.lmax + 1
        gls = req.spectra()
        rows = np.stack([np.array(c) for c in _fields.cls2cov(gls, nl, req.nshell, req.ncorr)])
        var = np.asarray([_fields._cltovar(_fields.getcl(gls, i, i)) for i in range(req.nshell)])
        st.var = var
        with T.cuda.stream(ln.stream):
            host = T.from_numpy(rows).pin_memory()
            st.cov = T.empty(host.shape, dtype=T.float64, device="cuda")
            st.cov.copy_(host, non_blocking=True)
            st._host_cov = host  # kept alive until the request closes
            st.w = T.empty_like(st.cov)
            st.state = T.empty(nl * self._nd[req.ncorr], dtype=T.float64, device="cuda")
            st.flag = T.zeros(1, dtype=T.int32, device="cuda")
            st.flag_host = T.zeros(1, dtype=T.int32).pin_memory()
        return st

    def close_request(self, st, keep: bool) -> None:
        # the flag was copied to pinned host memory by the request's last launch,
        # which has completed
        if int(st.flag_host[0]):
            raise ValueError(f"request {st.req.rid}: covariance matrix is not positive definite")
        if keep:
            self.results[st.req.rid] = st.maps.clone()  # out of the arena
        for o in st.offs:
            self.arena.release(o)
        st.maps = st.cov = st.w = st.state = st._host_cov = None

    # launches
    def launch(self, key, chunks):
        T = self.T
        ln = self.lanes[key]
        nside, lmax = key
        nl = lmax + 1
        ncol = sum(n for _, _, n in chunks)
        epi = (_lib.Epilogue * ncol)()
        col = 0
        for st, s0, n in chunks:
            for j in range(n):
                epi[col + j] = _lib.Epilogue(_lib.T_LOGNORMAL, 0, float(st.req.shift),
                                             float(st.var[s0 + j]))
            col += n
        sp = ln.stream.cuda_stream
        with T.cuda.stream(ln.stream):
            host_epi = T.frombuffer(bytearray(bytes(epi)), dtype=T.uint8).pin_memory()
            dev_epi = T.empty(host_epi.numel(), dtype=T.uint8, device="cuda")
            dev_epi.copy_(host_epi, non_blocking=True)
            alm = ln.alm[: _nalm(lmax) * ncol]
            col = 0
            for st, s0, n in chunks:
                k = st.req.ncorr
                _lib.call("ga_iternorm_batch", nl, k, k, k, n, 1 if s0 == 0 else 0,
                          _lib.ptr(st.cov[s0:]), _lib.ptr(st.state), _lib.ptr(st.w[s0:]),
                          _lib.ptr(st.flag), sp)
                _lib.call("ga_fields_alm", nl, k, s0, n, _lib.ptr(st.w[s0:]), None, 1, st.req.seed,
                          _lib.ptr(alm[col:]), ncol, None, sp)
                col += n
            _lib.call("ga_legendre_synth_batch_f32", ln.plan.h, ncol, _lib.ptr(alm), _lib.ptr(ln.ws),
                      ln.ws.numel(), sp)
            _lib.call("ga_ring_synth_batch_f32", ln.plan.h, ncol, _lib.ptr(ln.ws),
                      _lib.ptr(ln.maps), _lib.ptr(dev_epi), sp)
            col = 0
            for st, s0, n in chunks:
                st.maps[s0:s0 + n].copy_(ln.maps[col:col + n], non_blocking=True)
                st.flag_host.copy_(st.flag, non_blocking=True)
                col += n
            ev = T.cuda.Event()
            ev.record(ln.stream)
        return (ev, host_epi, dev_epi)

    def poll(self, token) -> bool:
        return token[0].query()

    def finish(self, token) -> None:
        pass
