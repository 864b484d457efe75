"""positions_from_delta on the device (glass/points.py:443-540).

* Counts: bit-exact.  With the device generator against the oracle's
  restatement of the Philox inversion; with a numpy generator against
  ``rng.poisson`` of the reference's expected counts (the reference's own
  count stream).
* Positions: the oracle's restatement to 1e-9 degrees, every point inside its
  pixel (ang2pix), and the reference test's cases
  (tests/core/test_points.py:235-386).
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

import glass_amd  # noqa: E402
from glass_amd.rng import DeviceRNG  # noqa: E402


def catpos(gen):
    lons, lats, counts = [], [], []
    for lo, la, co in gen:
        lons.append(np.asarray(lo))
        lats.append(np.asarray(la))
        counts.append(np.asarray(co))
    if not lons:
        return np.zeros(0), np.zeros(0), 0
    return np.concatenate(lons), np.concatenate(lats), sum(counts)


# expected counts per pixel: about 1700, 2400 (many Poisson parts) and 1.7
@pytest.mark.parametrize("nside,ngal,dtype", [(32, 0.5, np.float64), (16, 0.05, np.float32),
                                              (64, 2e-3, np.float64)])
def test_device_generator_counts_and_positions_match_oracle(nside, ngal, dtype):
    print(f"points nside={nside} ngal={ngal} {np.dtype(dtype).name}", flush=True)
    npix = 12 * nside**2
    rng = np.random.default_rng(nside)
    delta = (rng.lognormal(0.0, 0.6, npix) - 1.0).astype(dtype)
    vis = rng.uniform(0.5, 1.0, npix)
    for model, bias in (("linear", 1.3), ("loglinear", 0.7)):
        bm = glass_amd.linear_bias if model == "linear" else glass_amd.loglinear_bias
        lon, lat, count = catpos(glass_amd.positions_from_delta(
            ngal, delta, bias, vis, bias_model=bm, rng=DeviceRNG(9), batch=10_000))
        lam = O.expected_counts(delta, ngal, bias, vis, model=model)
        counts = O.device_poisson(lam, 9, 0)
        assert count == counts.sum() and lon.shape == (count,)
        rlon, rlat, ipix = O.device_positions(nside, counts, 9, 0)
        assert np.max(np.abs(lon - rlon)) < 1e-9 and np.max(np.abs(lat - rlat)) < 1e-9
        got = O.ang2pix_ring(nside, np.radians(90 - lat), np.radians(lon))
        assert np.array_equal(got, ipix)


def test_numpy_generator_counts_are_the_reference_stream():
    """rng = numpy Generator: counts are rng.poisson of the reference's expected
    counts (exactly the reference's draws), positions inside their pixels."""
    nside = 32
    npix = 12 * nside**2
    r = np.random.default_rng(5)
    delta = r.normal(0.0, 0.5, npix)
    vis = (r.uniform(size=npix) > 0.2).astype(np.float64)
    lon, lat, count = catpos(glass_amd.positions_from_delta(
        3.0, delta, 0.8, vis, rng=np.random.default_rng(123)))
    ref = np.random.default_rng(123).poisson(O.expected_counts(delta, 3.0, 0.8, vis))
    assert count == ref.sum()
    ipix = O.ang2pix_ring(nside, np.radians(90 - lat), np.radians(lon))
    assert np.array_equal(np.bincount(ipix, minlength=npix), ref)
    # rng = None: the reference's default generator (seed 42)
    _, _, c0 = catpos(glass_amd.positions_from_delta(3.0, delta, 0.8, vis))
    assert c0 == np.random.default_rng(42).poisson(O.expected_counts(delta, 3.0, 0.8, vis)).sum()


def test_remove_monopole_and_torch_inputs():
    nside = 16
    npix = 12 * nside**2
    delta = np.random.default_rng(2).lognormal(0, 0.4, npix) - 1
    lon, lat, count = catpos(glass_amd.positions_from_delta(
        0.2, delta, 1.0, remove_monopole=True, rng=DeviceRNG(3)))
    counts = O.device_poisson(O.expected_counts(delta, 0.2, 1.0, remove_monopole=True), 3, 0)
    # the device mean is summed in another order: counts agree except where a
    # uniform falls within rounding of a CDF step (none expected here)
    assert count == counts.sum()
    t = torch.as_tensor(delta).cuda()
    outs = list(glass_amd.positions_from_delta(0.2, t, 1.0, rng=DeviceRNG(3)))
    assert all(isinstance(lo, torch.Tensor) and lo.is_cuda for lo, _, _ in outs)


def test_reference_cases():
    """tests/core/test_points.py:235-386 on the drop-in."""
    nside = 128
    npix = 12 * nside**2
    ngal, bias = 1e-3, 0.8
    delta, vis = np.zeros(npix), np.ones(npix)
    for kw in ({}, {"rng": np.random.default_rng(0)}, {"rng": DeviceRNG(1)}):
        lon, lat, count = catpos(glass_amd.positions_from_delta(ngal, delta, bias, vis, **kw))
        assert int(count) == count and lon.shape == lat.shape == (count,)
    lon, lat, count = catpos(glass_amd.positions_from_delta(ngal, delta, None, vis,
                                                            bias_model=lambda x: x))
    assert lon.shape == lat.shape == (count,)
    lon, lat, count = catpos(glass_amd.positions_from_delta(ngal, delta, bias, None))
    assert lon.shape == (count,)
    lon, lat, count = catpos(glass_amd.positions_from_delta(ngal, delta, bias, vis,
                                                            remove_monopole=True))
    assert lon.shape == (count,)
    lon, lat, count = catpos(glass_amd.positions_from_delta(ngal, np.linspace(-1, -1, npix),
                                                            None, vis))
    assert count == 0 and lon.size == 0 and lat.size == 0
    big = np.random.default_rng(1).normal(100, 1, size=(npix,))
    lon, lat, count = catpos(glass_amd.positions_from_delta(ngal, big, bias, vis))
    assert lon.shape == (count,) and count > 0
    # several populations
    lon, lat, count = catpos(glass_amd.positions_from_delta(np.asarray([1e-3, 2e-3]),
                                                            np.zeros(12), 0.8, np.ones(12)))
    assert count.shape == (2,) and lon.shape == (count.sum(),)
    lon, lat, count = catpos(glass_amd.positions_from_delta(1e-3, np.zeros((3, 2, 12)), 0.8,
                                                            np.ones(12)))
    assert count.shape == (3, 2) and lat.shape == (count.sum(),)
    vis12 = np.ones(12)
    lon, lat, count = catpos(glass_amd.positions_from_delta(np.asarray([1e-3, 2e-3]),
                                                            np.zeros((3, 1, 12)), 0.8, vis12))
    assert count.shape == (3, 2) and lon.shape == (count.sum(),)
    vis12[:6] = 0.0
    lon, lat, count = catpos(glass_amd.positions_from_delta(np.asarray([1e-3, 2e-3]),
                                                            np.zeros((3, 1, 12)), 0.8, vis12))
    assert count.shape == (3, 2) and lon.shape == (count.sum(),)
    with pytest.raises(TypeError, match="bias_model must be callable"):
        next(glass_amd.positions_from_delta(1e-3, np.zeros(12), 0.8, np.ones(12), bias_model=0))


def test_batches_respect_the_batch_size():
    nside = 16
    npix = 12 * nside**2
    batches = list(glass_amd.positions_from_delta(5.0, np.zeros(npix), rng=DeviceRNG(4),
                                                  batch=1000))
    assert all(lo.size <= 1000 for lo, _, _ in batches)
    assert sum(c for _, _, c in batches) == sum(lo.size for lo, _, _ in batches)
