"""The allocation-free synthesis entry point against the workspace one.

``ga_alm2map_batch_f32_noalloc`` runs the kernels of ``ga_alm2map_batch_f32``
on the library's fixed device scratch instead of a caller workspace, so on
the same inputs the two give the same maps bit for bit.  Two calls in a row,
the second smaller than the first (the scratch keeps the first call's
contents), plus a size that does not fit, which is refused on the host.
"""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from glass_amd import _lib  # noqa: E402


def _alm_batch(lmax, S, seed):
    rng = np.random.default_rng(seed)
    nalm = (lmax + 1) * (lmax + 2) // 2
    ell = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    amp = np.sqrt(1e-4 * (ell + 1.0) ** -1.2)[:, None]
    a = (rng.standard_normal((nalm, S)) + 1j * rng.standard_normal((nalm, S))) * amp
    a[: lmax + 1] = a[: lmax + 1].real  # m = 0 coefficients are real
    return torch.as_tensor(a.astype(np.complex64)).to("cuda")


def _epilogues(S):
    epi = (_lib.Epilogue * S)()
    for i in range(S):
        epi[i] = _lib.Epilogue(_lib.T_LOGNORMAL, 0, 1.0 + 0.1 * i, 1e-3 * (i + 1))
    return torch.frombuffer(bytearray(bytes(epi)), dtype=torch.uint8).to("cuda")


def _with_workspace(plan, S, alm, epi):
    ws = _lib.workspace(plan.workspace_bytes(S, 4))
    out = torch.full((S, 12 * plan.nside**2), float("nan"), dtype=torch.float32, device="cuda")
    rc = _lib.lib().ga_alm2map_batch_f32(plan.h, S, _lib.ptr(alm), _lib.ptr(out), _lib.ptr(epi),
                                         _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
    _lib.check(rc, "ga_alm2map_batch_f32")
    return out


def _no_alloc(plan, S, alm, epi):
    out = torch.full((S, 12 * plan.nside**2), float("nan"), dtype=torch.float32, device="cuda")
    rc = _lib.lib().ga_alm2map_batch_f32_noalloc(plan.h, S, _lib.ptr(alm), _lib.ptr(out),
                                                 _lib.ptr(epi), _lib.stream_ptr())
    _lib.check(rc, "ga_alm2map_batch_f32_noalloc")
    return out


def test_no_alloc_synthesis_matches_workspace_entry_point():
    cap = ctypes.c_size_t()
    _lib.check(_lib.lib().ga_noalloc_scratch_bytes(ctypes.byref(cap)))
    assert cap.value == 128 << 20
    # (nside, lmax, S): the second call is smaller in every dimension
    for nside, lmax, S, seed in ((128, 256, 40, 5), (32, 64, 5, 6)):
        print(f"no-alloc synthesis nside={nside} lmax={lmax} S={S}", flush=True)
        plan = _lib.get_plan(nside, lmax)
        assert plan.workspace_bytes(S, 4) <= cap.value
        alm = _alm_batch(lmax, S, seed)
        epi = _epilogues(S)
        # split the l-ranges into two chunks even at these sizes, so that the
        # parking of both forms (its own region here, the ring buffer in the
        # no-alloc call) is compared
        plan.set_tuning(_lib.TUNE_CHUNK_MIN, 2)
        try:
            ref = _with_workspace(plan, S, alm, epi)
            got = _no_alloc(plan, S, alm, epi)
            torch.cuda.synchronize()
        finally:
            plan.set_tuning(_lib.TUNE_CHUNK_MIN, 0)
        assert torch.isfinite(ref).all()
        assert torch.equal(got, ref)


def test_no_alloc_synthesis_rejects_oversize_before_launch():
    nside, lmax, S = 512, 1024, 64
    plan = _lib.get_plan(nside, lmax)
    cap = ctypes.c_size_t()
    _lib.check(_lib.lib().ga_noalloc_scratch_bytes(ctypes.byref(cap)))
    assert plan.workspace_bytes(S, 4) > cap.value
    alm = torch.zeros(((lmax + 1) * (lmax + 2) // 2, S), dtype=torch.complex64, device="cuda")
    out = torch.zeros((S, 12 * nside**2), dtype=torch.float32, device="cuda")
    rc = _lib.lib().ga_alm2map_batch_f32_noalloc(plan.h, S, _lib.ptr(alm), _lib.ptr(out), None,
                                                 _lib.stream_ptr())
    torch.cuda.synchronize()
    assert rc == _lib.GA_ERR_UNSUPPORTED
    assert b"scratch" in _lib.lib().ga_last_error()
