"""Pixel windows on the GPU (``ga_pixwin``) and the reference's default
lensing path (``discretized=True``) that uses them.

Tolerances: the device bins the pixel pairs by separation and replaces each
bin by a three-node rule matching its first three moments, so it agrees with
the oracle's exact per-pair sums to <= 1e-8 (absolute, on windows of order 1).
Parity with healpy's tabulated windows is unpinned (the tables are not
available); the definition itself is pinned by tests/test_oracle.py.
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

import glass_amd  # noqa: E402
from glass_amd import _lib  # noqa: E402
from glass_amd import healpix as hp  # noqa: E402


def device_pixwin(nside, lmax, spin):
    n = ctypes.c_size_t()
    _lib.call("ga_pixwin_workspace_size", nside, lmax, ctypes.byref(n))
    ws = _lib.workspace(n.value)
    pw0 = torch.empty(lmax + 1, dtype=torch.float64, device="cuda")
    pws = torch.empty(lmax + 1, dtype=torch.float64, device="cuda")
    _lib.call("ga_pixwin", nside, lmax, spin, _lib.ptr(pw0), _lib.ptr(pws) if spin else None,
              _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
    return pw0.cpu().numpy(), pws.cpu().numpy()


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16, 32])
def test_pixwin_matches_oracle_pairs(nside):
    lmax = 4 * nside
    print(f"pixwin nside={nside} lmax={lmax}", flush=True)
    for spin in (0, 2, 1):
        pw0, pws = device_pixwin(nside, lmax, spin)
        ref0 = O.pixwin_pairs(nside, lmax, 0)
        assert np.max(np.abs(pw0 - ref0)) < 1e-8, (spin, np.max(np.abs(pw0 - ref0)))
        if spin:
            refs = O.pixwin_pairs(nside, lmax, spin)
            assert np.max(np.abs(pws - refs)) < 1e-8, (spin, np.max(np.abs(pws - refs)))


def test_pixwin_is_deterministic_and_bounded_at_large_nside():
    """Fixed-point binning and fixed-order reductions: two runs agree bit for
    bit.  w_0 = 1, windows in (0, 1] and decreasing up to 4 nside."""
    nside = 1024
    print(f"pixwin nside={nside}", flush=True)
    a0, a2 = device_pixwin(nside, 4 * nside, 2)
    b0, b2 = device_pixwin(nside, 4 * nside, 2)
    assert np.array_equal(a0, b0) and np.array_equal(a2, b2)
    assert abs(a0[0] - 1) < 1e-12
    assert np.all(np.diff(a0) < 0) and a0[-1] > 0
    assert np.all(a2[:2] == 0) and np.all(np.diff(a2[2:]) < 0) and a2[-1] > 0
    # pixel windows at fixed l / nside converge as the pixels flatten
    c0, _ = device_pixwin(512, 4 * 512, 0)
    assert np.max(np.abs(a0[::2] - c0)) < 2e-3


def test_pixwin_reference_interface():
    """glass/healpix.py:313-356 as the reference's tests use it
    (tests/core/test_healpix.py:323-350): shapes at nside 4, lmax 11, the pair
    for pol=True, and the datapath error."""
    pw = hp.pixwin(4, lmax=11)
    assert isinstance(pw, np.ndarray) and pw.shape == (12,) and pw.dtype == np.float64
    pw0, pw2 = hp.pixwin(4, lmax=11, pol=True)
    assert np.stack([pw0, pw2]).shape == (2, 12)
    assert np.allclose(pw0, O.pixwin_pairs(4, 11, 0), atol=1e-8)
    assert np.allclose(pw2, O.pixwin_pairs(4, 11, 2), atol=1e-8)
    t0 = hp.pixwin(4, lmax=11, xp=torch)
    assert isinstance(t0, torch.Tensor) and t0.dtype == torch.float64
    assert len(hp.pixwin(16)) == 48  # default lmax = 3 nside - 1
    with pytest.raises(ValueError, match="lmax"):
        hp.pixwin(4, lmax=17)


def test_pixwin_wrong_datapath(monkeypatch, tmp_path):
    bad = str(tmp_path / "missing")
    monkeypatch.setenv("HEALPY_DATAPATH", bad)
    with pytest.raises(ValueError, match=f"Healpy datapath not found at '{bad}"):
        hp.pixwin(4, lmax=11, pol=True)


def test_from_convergence_reference_defaults():
    """The reference's test of from_convergence (tests/core/test_lensing.py:21-56)
    with its default arguments (discretized=True): result counts for every
    combination, then the values of the default shear against the oracle
    chain with the oracle's pixel-window ratio."""
    n_side = 32
    rng = np.random.default_rng(42)
    kappa = rng.random(hp.nside2npix(n_side)) * 10.0
    assert glass_amd.from_convergence(kappa) == ()
    assert len(glass_amd.from_convergence(kappa, potential=True)) == 1
    assert len(glass_amd.from_convergence(kappa, deflection=True)) == 1
    assert len(glass_amd.from_convergence(kappa, shear=True)) == 1
    assert len(glass_amd.from_convergence(kappa, potential=True, deflection=True)) == 2
    assert len(glass_amd.from_convergence(kappa, potential=True, shear=True)) == 2
    assert len(glass_amd.from_convergence(kappa, deflection=True, shear=True)) == 2
    res = glass_amd.from_convergence(kappa, potential=True, deflection=True, shear=True)
    assert len(res) == 3

    # default shear (lmax = 3 nside - 1) against the oracle chain of
    # glass/lensing.py:306-367 with pw2 / pw0 from the oracle's pair form
    (gamma,) = glass_amd.from_convergence(kappa, shear=True)
    lmax = 3 * n_side - 1
    ratio = O.pixwin_pairs(n_side, lmax, 2) / O.pixwin_pairs(n_side, lmax, 0)
    a = O.map2alm_weighted(kappa, O.pixel_weights_cg(n_side, 2 * n_side), lmax, niter=3)
    a = O.almxfl(a, O.kappa_to_shear_factor(lmax, ratio))
    q, u = O.alm2map_spin([a, np.zeros_like(a)], n_side, 2, lmax)
    ref = q + 1j * u
    err = np.max(np.abs(gamma - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2))
    assert err < 1e-9, err
    g1, g2 = glass_amd.shear_from_convergence(kappa)
    err = np.max(np.abs(g1 + 1j * g2 - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2))
    assert err < 1e-9, err


def test_alm2map_pixwin_smooths_with_the_windows():
    """alm2map(pixwin=True) = alm2map of almxfl(alm, pw0) with healpy's
    default-length window (l < 3 nside; higher l dropped)."""
    nside, lmax = 8, 30
    rng = np.random.default_rng(5)
    alm = rng.standard_normal(O.nalm(lmax)) + 1j * rng.standard_normal(O.nalm(lmax))
    got = hp.alm2map(alm, nside, lmax=lmax, pixwin=True, pol=False)
    pw0 = O.pixwin_pairs(nside, 3 * nside - 1, 0)
    ref = O.alm2map(O.almxfl(alm, pw0), nside)
    assert np.max(np.abs(got - ref)) / np.sqrt(np.mean(ref**2)) < 1e-8
