"""GPU parity of the field-side kernels and of glass.generate against the
oracle, with the reference's own known-answer vectors and invariants."""

import json
import pathlib
import warnings

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

import glass_amd  # noqa: E402
from glass_amd import fields as F  # noqa: E402

GOLD = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


@pytest.mark.parametrize("nd", [False, True])
@pytest.mark.parametrize("k", [0, 1, 2])
def test_iternorm_device_matches_oracle(k, nd):
    cov = np.asarray(GOLD["iternorm_cov_nd" if nd else "iternorm_cov_0d"])
    rows = [cov[..., i, i::-1][..., : min(i, k) + 1] for i in range(cov.shape[-1])]
    got = list(glass_amd.iternorm(rows))
    ref = list(O.iternorm(rows))
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        np.testing.assert_allclose(g, r, rtol=1e-14, atol=1e-15)


def test_iternorm_device_errors():
    with pytest.raises(ValueError, match="empty covariance"):
        list(glass_amd.iternorm([np.ones(0)]))
    with pytest.raises(ValueError, match="shape mismatch"):
        list(glass_amd.iternorm([np.ones(1), np.ones((5, 2))]))
    with pytest.raises(ValueError, match="not positive definite"):
        list(glass_amd.iternorm([np.asarray([1.0]), np.asarray([0.1, 1.0])]))


def test_iternorm_long_chain():
    rng = np.random.default_rng(0)
    n, k, nl = 12, 3, 50
    A = rng.standard_normal((nl, n, n))
    cov = A @ np.swapaxes(A, -1, -2) + n * np.eye(n)
    rows = [cov[:, i, i::-1][:, : min(i, k) + 1] for i in range(n)]
    rows = [np.pad(r, ((0, 0), (0, k + 1 - r.shape[1]))) for r in rows]
    for g, r in zip(glass_amd.iternorm(rows), O.iternorm(rows)):
        np.testing.assert_allclose(g, r, rtol=1e-12, atol=1e-14)


def test_multalm_and_reorder_vectors():
    for case in GOLD["multalm"]:
        out = glass_amd.harmonics.multalm(np.asarray(case["alm"]), np.asarray(case["bl"]))
        np.testing.assert_array_equal(out, case["out"])
    g = GOLD["glass_to_healpix_alm"]
    out = F._glass_to_healpix_alm(np.asarray(g["in"], dtype=np.complex128))
    np.testing.assert_array_equal(out, np.asarray(g["out"], dtype=np.complex128))
    rng = np.random.default_rng(4)
    x = rng.standard_normal(O.nalm(40)) + 1j * rng.standard_normal(O.nalm(40))
    np.testing.assert_array_equal(F._glass_to_healpix_alm(x), O.glass_to_healpix_alm(x))


def test_cls2cov_matches_reference_vectors():
    g = GOLD["cls2cov_multi"]
    gen = glass_amd.cls2cov([np.asarray(c) for c in g["cls"]], g["nl"], g["nf"], g["nc"])
    c1 = next(gen)
    first = c1.copy()
    next(gen)
    c3 = next(gen)
    assert c1 is c3
    np.testing.assert_array_equal(first[:, 0], g["first_col0"])
    np.testing.assert_array_equal(c3[:, 0], g["final_col0"])


def _gls(nfield, lmax, ncorr, rho=(0.5,), amp=lambda i: 1.0):
    ell = np.arange(lmax + 1)
    auto = [amp(i) * 1e-4 * (ell + 1.0) ** -1.2 * np.exp(-((ell / 200.0) ** 2))
            for i in range(nfield)]
    gls = []
    for i in range(nfield):
        for j in range(i, -1, -1):
            d = i - j
            if d == 0:
                gls.append(auto[i])
            elif d <= len(rho):
                gls.append(rho[d - 1] * np.sqrt(auto[i] * auto[j]))
            else:
                gls.append(np.zeros(0))
    return gls


@pytest.mark.parametrize("ncorr", [None, 1])
def test_generate_fp64_equals_oracle(ncorr):
    gls = _gls(4, 40, 1)
    nside = 16
    fields = [glass_amd.grf.Lognormal(1.0), glass_amd.grf.Normal(),
              glass_amd.grf.Lognormal(0.7), glass_amd.grf.SquaredNormal(0.8, 1.3)]
    ref_fields = [lambda x, v: O.lognormal(x, v, 1.0), lambda x, v: x,
                  lambda x, v: O.lognormal(x, v, 0.7), lambda x, v: O.squared_normal(x, 0.8, 1.3)]
    got = list(glass_amd.generate(fields, gls, nside, ncorr=ncorr,
                                  rng=np.random.default_rng(0)))
    ref = list(O.generate(ref_fields, gls, nside, ncorr=ncorr, rng=np.random.default_rng(0)))
    assert len(got) == 4
    for g, r in zip(got, ref):
        assert g.shape == (12 * nside**2,) and g.dtype == np.float64
        assert np.max(np.abs(g - r)) / np.sqrt(np.mean(r**2)) < 1e-12


def test_generate_fp32_close_to_oracle():
    gls = _gls(6, 64, 3, rho=(0.6, 0.3, 0.1), amp=lambda i: 1 + i)
    nside = 32
    fields = [glass_amd.grf.Lognormal(1.0) for _ in range(6)]
    got = list(glass_amd.generate(fields, gls, nside, ncorr=3, rng=np.random.default_rng(5),
                                  precision="fp32", batch=4))
    ref = list(O.generate([lambda x, v: O.lognormal(x, v) for _ in range(6)], gls, nside,
                          ncorr=3, rng=np.random.default_rng(5)))
    for g, r in zip(got, ref):
        assert g.dtype == np.float32
        assert np.max(np.abs(g - r)) / np.sqrt(np.mean(r**2)) < 1e-5


def test_generate_reference_invariants():
    # tests/core/test_fields.py:362-438
    gls = [np.asarray([1.0, 0.5, 0.1])]
    a = list(F._generate_grf(gls, 4, rng=np.random.default_rng(42)))
    b = list(F._generate_grf(gls, 4, ncorr=1, rng=np.random.default_rng(42)))
    np.testing.assert_array_equal(a[0], b[0])
    assert a[0].shape == (192,)
    with pytest.raises(ValueError, match="all gls are empty"):
        list(F._generate_grf([np.asarray([])], 4))
    fields = [lambda x, var: x, lambda x, var: x]
    with pytest.raises(ValueError, match="mismatch between number of fields and gls"):
        list(glass_amd.generate(fields, [np.ones(10), np.ones(10)], nside=16))
    gls = [np.ones(10), np.ones(10), np.ones(10)]
    res = list(glass_amd.generate(fields, gls, nside=16))
    assert len(res) == 2 and all(r.shape == (12 * 256,) for r in res)
    fields = [lambda x, var: x, lambda x, var: x**2]
    res = list(glass_amd.generate(fields, gls, nside=16))
    np.testing.assert_allclose(res[1], res[0] ** 2, atol=1e-5)
    with pytest.deprecated_call():
        r = glass_amd.generate_lognormal([np.asarray([1.0, 0.5, 0.1])], 4)
    next(r)
    with pytest.deprecated_call():
        r = glass_amd.generate_gaussian([np.asarray([1.0, 0.5, 0.1])], 4)
    next(r)


def test_generate_default_seed_and_errors():
    gls = [np.asarray([1.0, 0.5, 0.1])]
    a = next(glass_amd.generate([glass_amd.grf.Normal()], gls, 4))
    b = next(O.generate([lambda x, v: x], gls, 4, rng=np.random.default_rng(42)))
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-13)
    with pytest.raises(ValueError, match="negative values in cl"):
        list(glass_amd.generate([glass_amd.grf.Normal()], [np.asarray([-1.0, 0.5])], 4))
    bad = [np.asarray([1.0]), np.asarray([1.0]), np.asarray([2.0])]
    with pytest.raises(ValueError, match="not positive definite"):
        list(glass_amd.generate([glass_amd.grf.Normal()] * 2, bad, 4))


def test_device_rng_statistics_and_batch_independence():
    gls = _gls(3, 32, 1)
    nside = 16
    rng = glass_amd.DeviceRNG(123)
    f = [glass_amd.grf.Normal()] * 3
    a = list(glass_amd.generate(f, gls, nside, ncorr=1, rng=rng, precision="fp32", batch=3))
    b = list(glass_amd.generate(f, gls, nside, ncorr=1, rng=rng, precision="fp32", batch=1))
    # same normals whatever the batching; the FP32 ring stage transforms two
    # shells in one complex FFT, so rounding (only) depends on the pairing
    for x, y in zip(a, b):
        assert np.max(np.abs(x - y)) / np.sqrt(np.mean(y.astype(float) ** 2)) < 1e-6
    # variance of a Gaussian field with spectrum g_l is sum (2l+1) g_l / 4 pi
    var = O.cltovar(gls[0])
    samples = [np.var(next(glass_amd.generate([glass_amd.grf.Normal()], [gls[0]], 32,
                                              rng=glass_amd.DeviceRNG(s), precision="fp32")))
               for s in range(12)]
    assert abs(np.mean(samples) / var - 1) < 0.15


def test_ten_lognormal_shells_nside128_parity():
    """The reference's CPU-sized case: 10 shells, nside 128, lmax 256, ncorr 1,
    seed 0, Lognormal(1); FP64 path vs the oracle (and the FP32 path within 1e-5)."""
    lmax, nside, n = 256, 128, 10
    gls = _gls(n, lmax, 1, rho=(0.5,))
    fields = [glass_amd.grf.Lognormal(1.0) for _ in range(n)]
    got = list(glass_amd.generate(fields, gls, nside, ncorr=1, rng=np.random.default_rng(0)))
    got32 = list(glass_amd.generate(fields, gls, nside, ncorr=1, rng=np.random.default_rng(0),
                                    precision="fp32", batch=5))
    ref = list(O.generate([lambda x, v: O.lognormal(x, v)] * n, gls, nside, ncorr=1,
                          rng=np.random.default_rng(0)))
    e64, e32 = [], []
    for g, g32, r in zip(got, got32, ref):
        rms = np.sqrt(np.mean(r**2))
        e64.append(np.max(np.abs(g - r)) / rms)
        e32.append(np.max(np.abs(g32 - r)) / rms)
    print(f"10 shells, nside 128: worst of {n} shells vs oracle: fp64 {max(e64):.2e}, fp32 {max(e32):.2e}",
          flush=True)
    assert max(e64) < 1e-12 and max(e32) < 1e-5


def test_from_convergence_chain():
    nside, lmax = 16, 40
    rng = np.random.default_rng(8)
    alm = rng.standard_normal(O.nalm(lmax)) + 1j * rng.standard_normal(O.nalm(lmax))
    kappa = O.alm2map(alm, nside)
    assert glass_amd.from_convergence(kappa) == ()
    res = glass_amd.from_convergence(kappa, lmax, potential=True, deflection=True, shear=True,
                                     discretized=False)
    assert len(res) == 3
    # oracle chain (glass/lensing.py:306-367): the pixel-weighted analysis with
    # healpy's iter=3 passes
    a = O.map2alm_weighted(kappa, O.pixel_weights_cg(nside, 2 * nside), lmax, niter=3)
    ell = np.arange(lmax + 1, dtype=float)
    fl = np.zeros(lmax + 1)
    fl[1:] = -2 / (ell[1:] * (ell[1:] + 1))
    a = O.almxfl(a, fl)
    psi = O.alm2map(a, nside)
    a = O.almxfl(a, np.sqrt(ell * (ell + 1)))
    q, u = O.alm2map_spin([a, np.zeros_like(a)], nside, 1, lmax)
    alpha = q + 1j * u
    fl = np.zeros(lmax + 1)
    fl[1:] = np.sqrt((ell[1:] - 1) * (ell[1:] + 2)) / 2
    a = O.almxfl(a, fl)
    q, u = O.alm2map_spin([a, np.zeros_like(a)], nside, 2, lmax)
    gamma = q + 1j * u
    for got, ref in zip(res, (psi, alpha, gamma)):
        assert np.max(np.abs(got - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2)) < 1e-10
    g1, g2 = glass_amd.shear_from_convergence(kappa, lmax, discretized=False)
    assert np.max(np.abs(g1 + 1j * g2 - gamma)) / np.sqrt(np.mean(np.abs(gamma) ** 2)) < 1e-10


@pytest.mark.parametrize("shell,seed", [(0, 7), (5, 2**40 + 3)])
def test_device_normals_match_oracle(shell, seed):
    """The device generator's normals (k = 0, w = 1: alm = healpy-ordered z
    with the m = 0 fix) against the oracle's Philox4x32-10 + Box-Muller; the
    device evaluates the transcendental functions in FP32."""
    from glass_amd import _lib

    nl = 40
    z = O.device_normals(seed, shell, np.arange(nl * (nl + 1) // 2))
    ref = O.glass_to_healpix_alm(z)
    ref[:nl] = ref[:nl].real + ref[:nl].imag
    w = torch.ones((1, nl, 1), dtype=torch.float64, device="cuda")
    out = torch.empty(nl * (nl + 1) // 2, dtype=torch.complex128, device="cuda")
    _lib.call("ga_fields_alm", nl, 0, shell, 1, _lib.ptr(w), None, 1, seed, None, 0,
              _lib.ptr(out), _lib.stream_ptr())
    got = out.cpu().numpy()
    assert np.max(np.abs(got - ref)) < 2e-5
    # the batched FP32 form (one thread walks the shells; a Philox call serves
    # an even shell and the odd one after it): three shells from `shell`
    ns = 3
    w3 = torch.ones((ns, nl, 1), dtype=torch.float64, device="cuda")
    o32 = torch.empty((nl * (nl + 1) // 2, ns), dtype=torch.complex64, device="cuda")
    _lib.call("ga_fields_alm", nl, 0, shell, ns, _lib.ptr(w3), None, 1, seed, _lib.ptr(o32), 0,
              None, _lib.stream_ptr())
    got32 = o32.cpu().numpy()
    for j in range(ns):
        zj = O.glass_to_healpix_alm(O.device_normals(seed, shell + j, np.arange(nl * (nl + 1) // 2)))
        zj[:nl] = zj[:nl].real + zj[:nl].imag
        assert np.max(np.abs(got32[:, j] - zj)) < 2e-5, (shell, j)
