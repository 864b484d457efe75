"""Pixel quadrature weights on the device (ga_pixel_weights) and the weighted
analysis (map2alm(use_pixel_weights=True), glass/healpix.py:232-276 as called
at glass/lensing.py:306 and :408).

healpy reads these weights from data files that are not available here
(parity with them unpinned); the device computes the weights that the files'
purpose defines -- the pixel sum of every Y*_lm with l <= 2 nside equals its
integral -- pinned against the oracle's dense minimum-norm solve and by the
exact analysis of band-limited maps that the weights exist for.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

from glass_amd import _lib  # noqa: E402
from glass_amd import healpix as hp  # noqa: E402


def _rand_alm(lmax, seed, decay=0.0):
    rng = np.random.default_rng(seed)
    n = O.nalm(lmax)
    ell = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)]).astype(float)
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / (ell + 1.0) ** decay
    a[: lmax + 1] = a[: lmax + 1].real
    return a


@pytest.mark.parametrize("nside", [4, 8])
def test_weights_match_the_dense_solve(nside):
    print(f"pixel weights nside={nside}", flush=True)
    w = hp.pixel_weights(nside).cpu().numpy()
    ref = O.pixel_weights(nside, 2 * nside)
    assert np.max(np.abs(w - ref)) < 1e-12 * 4 * np.pi / w.size * 100


@pytest.mark.parametrize("nside", [16, 64])
def test_weights_match_the_oracle_cg(nside):
    print(f"pixel weights nside={nside}", flush=True)
    w = hp.pixel_weights(nside).cpu().numpy()
    ref = O.pixel_weights_cg(nside, 2 * nside)
    assert np.max(np.abs(w - ref)) < 1e-10 * 4 * np.pi / w.size
    assert abs(w.sum() - 4 * np.pi) < 1e-10


@pytest.mark.parametrize("nside", [32, 256])
def test_weighted_analysis_is_exact_for_band_limited_maps(nside):
    """A map band-limited to lmax = nside: the weighted analysis recovers its
    alm without any Jacobi pass (to rounding); uniform weights do not."""
    print(f"weighted analysis nside={nside}", flush=True)
    lmax = nside
    a = _rand_alm(lmax, nside)
    m = O.alm2map(a, nside, lmax)
    got = hp.map2alm(m, lmax=lmax, pol=False, use_pixel_weights=True, iter=0)
    uni = hp.map2alm(m, lmax=lmax, pol=False, use_pixel_weights=False, iter=0)
    e_w = np.max(np.abs(got - a)) / np.max(np.abs(a))
    e_u = np.max(np.abs(uni - a)) / np.max(np.abs(a))
    print(f"  weighted {e_w:.2e}, uniform {e_u:.2e}", flush=True)
    assert e_w < 1e-11 and e_u > 1e3 * e_w


@pytest.mark.parametrize("niter", [0, 3])
def test_weighted_analysis_matches_oracle(niter):
    nside, lmax = 16, 47
    rng = np.random.default_rng(niter)
    m = rng.standard_normal(12 * nside**2)
    got = hp.map2alm(m, lmax=lmax, pol=False, use_pixel_weights=True, iter=niter)
    ref = O.map2alm_weighted(m, O.pixel_weights_cg(nside, 2 * nside), lmax, niter=niter)
    assert np.max(np.abs(got - ref)) < 1e-11 * np.max(np.abs(ref))
    # the FP32 batch form with the same weights
    w = hp.pixel_weights(nside)
    plan = _lib.get_plan(nside, lmax)
    maps = torch.as_tensor(np.stack([m, -m]).astype(np.float32)).cuda()
    alm = torch.empty((O.nalm(lmax), 2), dtype=torch.complex64, device="cuda")
    tmp = torch.empty_like(maps)
    ws = _lib.workspace(plan.workspace_bytes(2, 4, analysis=True))
    _lib.call("ga_map2alm_weighted_batch_f32", plan.h, 2, _lib.ptr(maps), _lib.ptr(w),
              _lib.ptr(alm), niter, _lib.ptr(tmp), _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
    a32 = alm.cpu().numpy()
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(a32[:, 0] - ref)) < 2e-5 * scale
    assert np.max(np.abs(a32[:, 1] + ref)) < 2e-5 * scale


def test_errors():
    plan = _lib.get_plan(8, 40)  # lmax > 2 nside: not solved for
    n = _lib.ctypes.c_size_t()
    _lib.call("ga_pixel_weights_workspace_size", plan.h, _lib.ctypes.byref(n))
    ws = _lib.workspace(n.value)
    w = torch.empty(12 * 64, dtype=torch.float64, device="cuda")
    with pytest.raises(NotImplementedError, match="lmax <= 2 nside"):
        _lib.call("ga_pixel_weights", plan.h, 1e-12, 100, _lib.ptr(w), None, None,
                  _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
    m = np.zeros(12 * 64)
    with pytest.raises(ValueError, match="NULL weights"):
        d = torch.as_tensor(m).cuda()
        alm = torch.empty(O.nalm(16), dtype=torch.complex128, device="cuda")
        p2 = _lib.get_plan(8, 16)
        ws2 = _lib.workspace(p2.workspace_bytes(1, 8, analysis=True))
        _lib.call("ga_map2alm_weighted_f64", p2.h, _lib.ptr(d), None, _lib.ptr(alm), 0,
                  _lib.ptr(d), _lib.ptr(ws2), ws2.numel(), _lib.stream_ptr())
    torch.cuda.synchronize()
