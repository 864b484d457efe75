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


def test_c5_geometries_alone_mixed_and_fp64():
    """One request per serving geometry (nside 256 .. 4096, lmax = 2 nside,
    ncorr 3, a few shells each), all arriving together through the device
    executor: every shell equals the same request generated alone by the
    direct batch engine to 1e-6, and the FP64 path to 1e-5 (lmax 8192 at
    nside 4096 included)."""
    reqs = [SimRequest(i, 0.0, ns, n, 2 * ns, ncorr=3, seed=400 + i)
            for i, (ns, n) in enumerate(((256, 5), (512, 4), (1024, 4), (2048, 3), (4096, 3)))]
    ex = DeviceExecutor()
    sched = run(reqs, ex, 240e9, keep_results=True)
    assert sorted(c.rid for c in sched.completed) == list(range(len(reqs)))
    worst = {}
    for r in reqs:
        print(f"serving request nside={r.nside} lmax={r.lmax} shells={r.nshell}", flush=True)
        mixed = ex.results[r.rid]
        eng = LognormalShells(r.spectra(), r.nside, ncorr=r.ncorr, shift=r.shift, seed=r.seed,
                              lmax=r.lmax)
        eng.run()
        torch.cuda.synchronize()
        e_alone = max(rel(mixed[s], eng.maps[s]) for s in range(r.nshell))
        e_fp64 = max(rel(mixed[s], eng.fp64_reference_map(s)) for s in range(r.nshell))
        worst[r.nside] = (e_alone, e_fp64)
        print(f"  alone vs mixed {e_alone:.2e}, vs FP64 {e_fp64:.2e}", flush=True)
        del eng
        torch.cuda.empty_cache()
        assert e_alone < 1e-6 and e_fp64 < 1e-5, (r.nside, e_alone, e_fp64)
