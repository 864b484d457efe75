"""GPU parity of the batched device engines (lognormal shells, shear).

Tolerances (max |error| / RMS of the reference output, per map):
  FP32 batched shear vs the FP64 oracle chain   <= 1e-5
  FP32 batched maps / shear vs the FP64 GPU path <= 1e-5
"""

import contextlib

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

from glass_amd import healpix as hp  # noqa: E402
from glass_amd import lensing  # noqa: E402
from glass_amd.batch import LognormalShells, ShearShells, shear_factor, shell_spectra  # noqa: E402


@contextlib.contextmanager
def plan_knob(nside, lmax, key, value):
    """Set one tuning knob of the (nside, lmax) plan for the block, then restore it."""
    from glass_amd import _lib

    plan = _lib.get_plan(nside, lmax)
    old = plan.get_tuning(key)
    plan.set_tuning(key, value)
    try:
        yield
    finally:
        plan.set_tuning(key, old)


def relerr(got, ref):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.max(np.abs(got - ref)) / np.sqrt(np.mean(ref**2)))


def test_shear_factor_is_the_reference_product():
    # glass/lensing.py:316-363: (-2 / (l(l+1))) * sqrt(l(l+1)) * sqrt((l-1)(l+2)) / 2
    lmax = 50
    ell = np.arange(lmax + 1, dtype=float)
    prod = np.zeros(lmax + 1)
    prod[1:] = (-2 / (ell[1:] * (ell[1:] + 1))) * np.sqrt(ell[1:] * (ell[1:] + 1)) \
        * np.sqrt((ell[1:] - 1) * (ell[1:] + 2)) / 2
    assert np.allclose(shear_factor(lmax), prod, rtol=1e-15, atol=0)
    assert np.allclose(shear_factor(lmax), O.kappa_to_shear_factor(lmax), rtol=1e-15, atol=0)


@pytest.mark.parametrize("nside,lmax,S", [(16, 31, 3), (32, 63, 6)])
def test_shear_batch_matches_oracle(nside, lmax, S):
    print(f"shear batch nside={nside} lmax={lmax} S={S}", flush=True)
    rng = np.random.default_rng(nside + S)
    kappas = []
    for _ in range(S):
        n = O.nalm(lmax)
        a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        ell = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
        a *= np.sqrt(1e-3 / (ell + 1.0) ** 1.5)
        kappas.append(O.alm2map(a, nside).astype(np.float32))
    eng = ShearShells(nside, lmax, S)
    gamma = eng.run(torch.as_tensor(np.stack(kappas)).cuda()).cpu().numpy()
    fl = O.kappa_to_shear_factor(lmax)
    w = O.pixel_weights_cg(nside, 2 * nside)
    for s in range(S):
        a = O.almxfl(O.map2alm_weighted(kappas[s].astype(np.float64), w, lmax, niter=3), fl)
        q, u = O.alm2map_spin([a, np.zeros_like(a)], nside, 2, lmax)
        assert relerr(gamma[s, 0], q) < 1e-5
        assert relerr(gamma[s, 1], u) < 1e-5


def test_shear_batch_rejects_bad_shape():
    eng = ShearShells(8, 15, 2)
    with pytest.raises(ValueError, match="contiguous"):
        eng.run(torch.zeros((3, 12 * 64), dtype=torch.float32, device="cuda"))


def _trelerr(got, ref):
    """max |got - ref| / RMS(ref) and RMS(got - ref) / RMS(ref), on the device."""
    got, ref = got.double(), ref.double()
    rms = torch.sqrt(torch.mean(ref * ref))
    d = got - ref
    return float(torch.max(torch.abs(d)) / rms), float(torch.sqrt(torch.mean(d * d)) / rms)


def _check_shells(nside, lmax, S, shells):
    gls = shell_spectra(S, lmax, amp=lambda i: 1.0 + 3.0 * i / max(S - 1, 1),
                        rho=(0.6, 0.3, 0.1), scale=1500.0)
    eng = LognormalShells(gls, nside, ncorr=3, seed=7, lmax=lmax, shear=True)
    eng.run()
    torch.cuda.synchronize()
    eng.check()
    fl = shear_factor(lmax)
    out = []
    for s in shells:
        ref = eng.fp64_reference_map(s)
        e_map, r_map = _trelerr(eng.maps[s], ref)
        # FP64 shear of the FP32 map (same input, so only the shear's own error counts)
        kappa = eng.maps[s].double()
        alm = hp.map2alm(kappa, lmax=lmax, pol=False, precision="fp64", use_pixel_weights=True)
        alm = hp.almxfl(alm, fl)
        q, u = hp.alm2map_spin([alm, torch.zeros_like(alm)], nside, 2, lmax)
        e_q, _ = _trelerr(eng.gamma[s, 0], q)
        e_u, _ = _trelerr(eng.gamma[s, 1], u)
        print(f"shell {s}: map max {e_map:.2e} rms {r_map:.2e}  gamma1 {e_q:.2e} "
              f"gamma2 {e_u:.2e}", flush=True)
        out.append((e_map, e_q, e_u))
    return out


@pytest.mark.parametrize("nside,lmax,S,k", [(128, 300, 20, 3), (64, 150, 5, 0), (128, 255, 37, 7)])
def test_fused_column_maxima_give_identical_maps(nside, lmax, S, k):
    """The field kernel's per-block column maxima (ga_fields_alm_colmax) reduce
    to the column maxima of the alm batch, and the operand preparation from
    them leaves the same maps as the preparation that reads the alm, bit for
    bit (37 shells: a partial last chunk of 5; lmax 255: a partial last block)."""
    gls = shell_spectra(S, lmax, amp=lambda i: 1.0 + i / S,
                        rho=(0.6, 0.3, 0.1, 0.05, 0.02, 0.01, 0.005))
    eng = LognormalShells(gls, nside, ncorr=k, seed=3, lmax=lmax)
    assert eng.colmax_fused
    eng.run()
    fused = eng.maps.clone()
    a = torch.view_as_real(eng.alm).reshape(eng.nalm, 2 * S).abs().amax(dim=0)
    assert eng.colpart.shape == (-(-eng.nalm // 256), 2 * S)
    assert torch.equal(eng.colpart.amax(dim=0), a)
    eng.colmax_fused = False
    eng.run()
    assert torch.equal(eng.maps, fused)


def test_lognormal_shells_with_shear_small():
    for e_map, e_q, e_u in _check_shells(64, 127, 8, [0, 3, 7]):
        assert e_map < 1e-5 and e_q < 1e-5 and e_u < 1e-5


def test_lognormal_shells_with_shear_64_shells_nside1024():
    """64 correlated lognormal shells at nside 1024, lmax 2047, ncorr 3 (spectra
    amplitude x1 -> x4, neighbour correlation 0.6/0.3/0.1), shear for all 64;
    every shell's map and both shear components within 1e-5 of the FP64 path
    (max |error| / RMS, each printed)."""
    errs = _check_shells(1024, 2047, 64, range(64))
    worst = np.max(np.array(errs), axis=0)
    print(f"worst of 64 shells: map {worst[0]:.2e} gamma1 {worst[1]:.2e} gamma2 {worst[2]:.2e}")
    assert np.all(worst < 1e-5), worst


def test_from_convergence_fp32_batch_agrees_with_module_function():
    nside, lmax = 16, 31
    rng = np.random.default_rng(2)
    a = rng.standard_normal(O.nalm(lmax)) + 1j * rng.standard_normal(O.nalm(lmax))
    kappa = O.alm2map(a, nside)
    g1, g2 = lensing.shear_from_convergence(kappa, lmax, discretized=False)
    eng = ShearShells(nside, lmax, 1)
    gamma = eng.run(torch.as_tensor(kappa[None].astype(np.float32)).cuda()).cpu().numpy()
    assert relerr(gamma[0, 0], g1) < 1e-5
    assert relerr(gamma[0, 1], g2) < 1e-5


@pytest.mark.parametrize("nside,lmax", [(16, 40), (64, 128), (128, 255), (256, 512)])
def test_ring_paths_agree(nside, lmax):
    """The FP32 ring-synthesis forms (radix-4 LDS stages / fused radix-2^K
    passes / the default: fused passes with the register-split kernels of the
    wide and Bluestein rings; power-of-two and Bluestein rings) give the same
    maps, and all are within 1e-5 of the FP64 path."""
    from glass_amd import _lib

    S = 6
    gls = shell_spectra(S, lmax, rho=(0.6, 0.3, 0.1))
    eng = LognormalShells(gls, nside, ncorr=3, seed=5, lmax=lmax)
    out = {}
    for path in (0, 1, 2):
        with plan_knob(nside, lmax, _lib.TUNE_RING_PATH, path):
            eng.run()
            torch.cuda.synchronize()
            out[path] = eng.maps.clone()
    for s in range(S):
        for path in (0, 2):
            assert relerr(out[path][s].cpu().numpy(), out[1][s].cpu().numpy()) < 2e-6, (path, s)
    for s in (0, S - 1):
        ref = eng.fp64_reference_map(s).cpu().numpy()
        for path in (0, 2):
            assert relerr(out[path][s].cpu().numpy(), ref) < 1e-5, (path, s)


@pytest.mark.parametrize("nside,lmax,S", [(256, 511, 6), (256, 700, 5), (512, 1023, 4)])
def test_wide_ring_kernel_matches_fused_passes(nside, lmax, S):
    """The register-fold kernel of the wide power-of-two rings (4 r > lmax)
    agrees with the four-array fused-pass kernel (same terms, order and
    passes; the compiler may contract products into FMAs differently, so to
    FP32 rounding: max |difference| / RMS < 2e-6), including bins with both a
    direct and a conjugate term (lmax > 2 r) and a lone last shell (odd S);
    and stays within 1e-5 of FP64."""
    from glass_amd import _lib

    print(f"wide rings nside={nside} lmax={lmax} S={S}", flush=True)
    gls = shell_spectra(S, lmax, rho=(0.6, 0.3, 0.1))
    eng = LognormalShells(gls, nside, ncorr=3, seed=11, lmax=lmax)
    out = {}
    for path in (0, 2):
        with plan_knob(nside, lmax, _lib.TUNE_RING_PATH, path):
            eng.run()
            torch.cuda.synchronize()
            out[path] = eng.maps.clone()
    for s in range(S):
        assert relerr(out[0][s].cpu().numpy(), out[2][s].cpu().numpy()) < 2e-6, s
    ref = eng.fp64_reference_map(S - 1).cpu().numpy()
    assert relerr(out[0][S - 1].cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("spin", [1, 2])
@pytest.mark.parametrize("nside,lmax,S", [(8, 20, 3), (64, 150, 5), (256, 511, 4)])
def test_spin_mfma_matches_valu_and_fp64(spin, nside, lmax, S):
    """FP32 spin synthesis: the split-FP16 MFMA path (E only, and E + B in two
    passes) against the VALU path and the FP64 path, max |error| / RMS < 1e-5;
    B = NULL gives the same maps as an explicit zero B."""
    from glass_amd import _lib
    from glass_amd.healpix import spin_batch_f32

    print(f"spin mfma spin={spin} nside={nside} lmax={lmax} S={S}", flush=True)
    rng = np.random.default_rng(spin * 7 + nside)
    n = O.nalm(lmax)
    ell = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)]).astype(float)
    amp = np.sqrt(1e-3 / (ell + 1.0) ** 1.5)
    E = (rng.standard_normal((n, S)) + 1j * rng.standard_normal((n, S))) * amp[:, None]
    B = (rng.standard_normal((n, S)) + 1j * rng.standard_normal((n, S))) * amp[:, None]
    Ed = torch.as_tensor(E).to(torch.complex64).cuda().contiguous()
    Bd = torch.as_tensor(B).to(torch.complex64).cuda().contiguous()
    out = {}
    for path in (1, 2):
        with plan_knob(nside, lmax, _lib.TUNE_SPIN_PATH, path):
            out[path, "eb"] = spin_batch_f32(Ed, Bd, nside, spin, lmax).cpu().numpy()
            out[path, "e0"] = spin_batch_f32(Ed, torch.zeros_like(Ed), nside, spin, lmax).cpu().numpy()
    with plan_knob(nside, lmax, _lib.TUNE_SPIN_PATH, 2):
        out[2, "none"] = spin_batch_f32(Ed, None, nside, spin, lmax).cpu().numpy()
    assert np.array_equal(out[2, "none"], out[2, "e0"]) or \
        relerr(out[2, "none"], out[2, "e0"]) < 1e-6
    for kind in ("eb", "e0"):
        for s in range(S):
            for c in range(2):
                assert relerr(out[2, kind][s, c], out[1, kind][s, c]) < 1e-5, (kind, s, c)
    for s in (0, S - 1):
        q, u = hp.alm2map_spin([E[:, s], B[:, s]], nside, spin, lmax)
        assert relerr(out[2, "eb"][s, 0], q) < 1e-5
        assert relerr(out[2, "eb"][s, 1], u) < 1e-5
        q, u = hp.alm2map_spin([E[:, s], np.zeros_like(E[:, s])], nside, spin, lmax)
        assert relerr(out[2, "none"][s, 0], q) < 1e-5
        assert relerr(out[2, "none"][s, 1], u) < 1e-5


@pytest.mark.parametrize("nside,lmax,S,niter", [(8, 20, 3, 0), (64, 150, 7, 0), (256, 600, 5, 1),
                                                (512, 1023, 17, 0), (128, 300, 40, 1)])
def test_analysis_mfma_matches_valu_and_fp64(nside, lmax, S, niter):
    """FP32 batched analysis: the split-FP16 MFMA contraction against the VALU
    one and against the FP64 path (max |error| / RMS of the alm < 1e-5)."""
    from glass_amd import _lib

    print(f"analysis mfma nside={nside} lmax={lmax} S={S} niter={niter}", flush=True)
    rng = np.random.default_rng(nside + S)
    maps = torch.as_tensor(rng.standard_normal((S, 12 * nside * nside))).cuda()
    m32 = maps.float().contiguous()
    out = {}
    for path in (1, 2):
        with plan_knob(nside, lmax, _lib.TUNE_ANAL_PATH, path):
            out[path] = hp.map2alm_batch_f32(m32, lmax, niter).T.cpu().numpy()
    for s in range(S):
        a1, a2 = out[1][s], out[2][s]
        scale = np.sqrt(np.mean(np.abs(a1) ** 2))
        assert np.max(np.abs(a2 - a1)) / scale < 1e-5, s
    for s in (0, S - 1):
        ref = hp.map2alm(maps[s], lmax=lmax, iter=niter, pol=False).cpu().numpy()
        scale = np.sqrt(np.mean(np.abs(ref) ** 2))
        assert np.max(np.abs(out[2][s] - ref)) / scale < 1e-5, s


@pytest.mark.parametrize("nside,lmax,S,niter", [(16, 40, 3, 0), (64, 150, 4, 1), (256, 600, 5, 0),
                                                (512, 1023, 6, 0)])
def test_ring_analysis_forms_agree(nside, lmax, S, niter):
    """FP32 batched analysis with the register-form ring stage (default)
    against the first form (radix-4 LDS stages, GA_TUNE_RING_PATH = 1):
    max |difference| / RMS of the alm < 2e-6 (power-of-two and Bluestein
    rings, rings with n <= lmax and > lmax, odd S), and against FP64 < 1e-5."""
    from glass_amd import _lib

    print(f"ring analysis nside={nside} lmax={lmax} S={S} niter={niter}", flush=True)
    rng = np.random.default_rng(3 * nside + S)
    maps = torch.as_tensor(rng.standard_normal((S, 12 * nside * nside))).cuda()
    m32 = maps.float().contiguous()
    out = {}
    for path in (0, 1):
        with plan_knob(nside, lmax, _lib.TUNE_RING_PATH, path):
            out[path] = hp.map2alm_batch_f32(m32, lmax, niter).T.cpu().numpy()
    for s in range(S):
        scale = np.sqrt(np.mean(np.abs(out[1][s]) ** 2))
        assert np.max(np.abs(out[0][s] - out[1][s])) / scale < 2e-6, s
    for s in (0, S - 1):
        ref = hp.map2alm(maps[s], lmax=lmax, iter=niter, pol=False).cpu().numpy()
        scale = np.sqrt(np.mean(np.abs(ref) ** 2))
        assert np.max(np.abs(out[0][s] - ref)) / scale < 1e-5, s


@pytest.mark.parametrize("nside,lmax,S", [(32, 63, 40), (128, 255, 64), (64, 150, 100)])
def test_wave_pair_legendre_matches_single_wave(nside, lmax, S):
    """The producer / consumer wave-pair form of the split-FP16 Legendre stage
    (tiles = 4: one wave runs the recurrence, its partner the MFMAs) issues
    the same MFMAs in the same order as the single-wave form with two tiles,
    so the maps agree bit for bit."""
    from glass_amd import _lib, autotune

    print(f"wave pairs nside={nside} lmax={lmax} S={S}", flush=True)
    gls = shell_spectra(S, lmax, rho=(0.6, 0.3, 0.1))
    eng = LognormalShells(gls, nside, ncorr=3, seed=11, lmax=lmax)
    saved = autotune.current(eng.plan)
    saved_flush = eng.plan.get_tuning(_lib.TUNE_F16_FLUSH)
    out = {}
    try:
        # the wave-pair form accumulates each l-sum in one chain: compare it
        # with the single-wave form in the same accumulation form
        eng.plan.set_tuning(_lib.TUNE_F16_FLUSH, -1)
        for nt in (2, 4):
            autotune.apply(eng.plan, {"path": 2, "nt": nt})
            eng.run()
            torch.cuda.synchronize()
            out[nt] = eng.maps.clone()
    finally:
        autotune.apply(eng.plan, saved)
        eng.plan.set_tuning(_lib.TUNE_F16_FLUSH, saved_flush)
    assert torch.equal(out[2], out[4])
    ref = eng.fp64_reference_map(S - 1).cpu().numpy()
    assert relerr(out[4][S - 1].cpu().numpy(), ref) < 1e-5
