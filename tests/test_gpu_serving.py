"""The serving loop on the GPU: a request's maps are the same alone and in a
mix of concurrent requests of other sizes, and equal the direct batch engine.

Tolerance: 1e-6 of the map RMS between alone / mixed / direct (the FP32 ring
stage transforms two shells in one complex FFT, so the rounding -- only --
depends on the pairing); 1e-5 against the FP64 path.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from glass_amd.batch import LognormalShells  # noqa: E402
from glass_amd.serving import DeviceExecutor, SimRequest, run, summarize  # noqa: E402


def rel(a, b):
    a = a.double()
    b = b.double()
    return float(torch.max(torch.abs(a - b)) / torch.sqrt(torch.mean(b * b)))


def _target():
    return SimRequest(0, 0.0, 64, 11, 128, ncorr=3, seed=77)


def test_request_is_the_same_alone_and_in_the_mix():
    cap = {32: 8, 64: 8}
    ex1 = DeviceExecutor(capacity=cap)
    s1 = run([_target()], ex1, 200e9, keep_results=True)
    assert [c.rid for c in s1.completed] == [0]
    alone = ex1.results[0].clone()

    others = [SimRequest(1, 0.0, 64, 7, 128, seed=5), SimRequest(2, 0.0, 32, 9, 64, seed=6),
              SimRequest(3, 0.0005, 64, 5, 128, seed=8), SimRequest(4, 0.001, 32, 20, 64, seed=9)]
    ex2 = DeviceExecutor(capacity=cap)
    s2 = run([_target(), *others], ex2, 200e9, keep_results=True)
    assert sorted(c.rid for c in s2.completed) == [0, 1, 2, 3, 4]
    # the target shared launches with other requests and was cut into chunks
    assert s2.max_requests_per_launch >= 2
    mixed = ex2.results[0]
    for s in range(11):
        assert rel(mixed[s], alone[s]) < 1e-6, s

    # the direct batch engine on the same spectra and seed
    r = _target()
    eng = LognormalShells(r.spectra(), r.nside, ncorr=r.ncorr, shift=r.shift, seed=r.seed,
                          lmax=r.lmax)
    eng.run()
    torch.cuda.synchronize()
    for s in range(11):
        assert rel(alone[s], eng.maps[s]) < 1e-6, s
    for s in (0, 5, 10):
        assert rel(alone[s], eng.fp64_reference_map(s)) < 1e-5, s
    summ = summarize(s2)
    assert summ["requests"] == 5 and summ["latency_s"]["p99"] > 0


def test_budget_admission_on_device():
    """A budget that holds one lane and one request at a time still serves
    every request, in arrival order, one after another."""
    ex = DeviceExecutor(capacity={32: 8})
    reqs = [SimRequest(i, 0.0, 32, 6, 64, seed=100 + i) for i in range(3)]
    need = ex.lane_bytes((32, 64)) + ex.request_bytes(reqs[0])
    sched = run(reqs, ex, need * 1.5, keep_results=True)
    assert [c.rid for c in sched.completed] == [0, 1, 2]
    assert sched.peak <= need * 1.5
    for i in range(3):
        assert torch.isfinite(ex.results[i]).all()
