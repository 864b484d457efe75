"""Pins the FP64 oracle before anything is compared against it.

Each check is chosen so that a plausible mistake in the oracle (dropped term,
wrong sign or index, transposed operand, broken exponent scaling) fails it:
literal vectors from the reference's own tests, closed forms, independent
library routines (scipy, numpy.fft, sympy), brute force on tiny inputs, and
invariants (orthonormality, adjointness, rotation).
"""

from __future__ import annotations

import json
import math
import pathlib

import numpy as np
import pytest
from scipy.special import roots_legendre, sph_harm_y

from oracle import oracle as O

GOLD = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def _hp(lmax, l, m):
    return m * (2 * lmax + 1 - m) // 2 + l


def _brute_synth(alm, lmax, theta, phi):
    out = np.zeros_like(theta)
    for m in range(lmax + 1):
        for l in range(m, lmax + 1):
            a = alm[_hp(lmax, l, m)]
            y = sph_harm_y(l, m, theta, phi)
            out += (a.real * y).real if m == 0 else 2 * (a * y).real
    return out


# --------------------------------------------------------------------------
# reference-owned arithmetic against the reference tests' literal vectors


@pytest.mark.parametrize("nd", [False, True])
@pytest.mark.parametrize("k", [0, 1, 2])
def test_iternorm_against_pinv_cholesky(k, nd):
    # tests/core/test_fields.py:31-104
    cov = np.asarray(GOLD["iternorm_cov_nd" if nd else "iternorm_cov_0d"])
    rows = (cov[..., i, i::-1][..., : min(i, k) + 1] for i in range(cov.shape[-1]))
    for n, w in enumerate(O.iternorm(rows)):
        Sn = cov[..., :n, :n].copy()
        cn = cov[..., :n, n].copy()
        vn = cov[..., n, n]
        if n > k:
            band = (np.abs(np.arange(n)[:, None] - np.arange(n)) <= k).astype(float)
            Sn *= band
            cn *= (np.arange(n) >= n - k).astype(float)
        Sinv = np.linalg.pinv(Sn)
        An = np.swapaxes(np.linalg.cholesky(Sinv + 1e-100 * np.eye(n)), -1, -2)
        an = (An @ cn[..., None])[..., 0]
        sn = np.sqrt(vn - np.einsum("...i,...i", an, an))
        assert w.shape[-1] <= min(n, k) + 1
        a, s = w[..., :-1], w[..., -1]
        np.testing.assert_allclose(np.einsum("...i,...i", a, a),
                                   np.einsum("...i,...i", an, an), rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(s, sn, rtol=1e-12)


def test_iternorm_errors():
    # tests/core/test_fields.py:107-119
    with pytest.raises(ValueError, match="empty covariance"):
        list(O.iternorm([np.ones(0)]))
    with pytest.raises(ValueError, match="shape mismatch"):
        list(O.iternorm([np.ones(1), np.ones((5, 2))]))
    with pytest.raises(ValueError, match="not positive definite"):
        list(O.iternorm([np.asarray([1.0]), np.asarray([0.1, 1.0])]))


def test_cls2cov_vectors():
    # tests/core/test_fields.py:166-252
    g = GOLD["cls2cov_single"]
    cls = [None if c is None else np.asarray(c) for c in g["cls"]]
    cov = next(O.cls2cov(cls, g["nl"], g["nf"], g["nc"]))
    assert cov.shape == (3, 3)
    np.testing.assert_array_equal(cov[:, 0], g["col0"])
    np.testing.assert_array_equal(cov[:, 1:], 0.0)

    g = GOLD["cls2cov_negative"]
    with pytest.raises(ValueError, match="negative values in cl"):
        next(O.cls2cov([np.asarray(c) for c in g["cls"]], g["nl"], g["nf"], g["nc"]))

    g = GOLD["cls2cov_multi"]
    gen = O.cls2cov([np.asarray(c) for c in g["cls"]], g["nl"], g["nf"], g["nc"])
    c1 = next(gen)
    c1_copy = c1.copy()
    c2 = next(gen)
    c3 = next(gen)
    # same buffer reused for every yield
    assert c1 is c2 is c3
    np.testing.assert_array_equal(c3[:, 0], g["final_col0"])
    np.testing.assert_array_equal(c1_copy[:, 0], g["first_col0"])


def test_multalm_vectors():
    # tests/core/test_harmonics.py:15-52
    for case in GOLD["multalm"]:
        out = O.multalm(np.asarray(case["alm"]), np.asarray(case["bl"]))
        np.testing.assert_array_equal(out, case["out"])


def test_glass_to_healpix_alm():
    g = GOLD["glass_to_healpix_alm"]
    out = O.glass_to_healpix_alm(np.asarray(g["in"], dtype=np.complex128))
    np.testing.assert_array_equal(out, np.asarray(g["out"], dtype=np.complex128))
    # closed form: GLASS l(l+1)/2 + m lands at healpy m(2 lmax + 1 - m)/2 + l
    lmax = 9
    idx = np.asarray([l * (l + 1) // 2 + m for l in range(lmax + 1) for m in range(l + 1)])
    x = np.arange(idx.size).astype(np.complex128)
    y = O.glass_to_healpix_alm(x)
    for l in range(lmax + 1):
        for m in range(l + 1):
            assert y[_hp(lmax, l, m)] == l * (l + 1) // 2 + m


def test_rng_stream_pinned():
    z = np.random.default_rng(42).standard_normal((6, 2))
    np.testing.assert_array_equal(z, np.asarray(GOLD["pcg64_seed42_normals_6x2"]))


def test_transform_closed_forms():
    rng = np.random.default_rng(7)
    x = rng.standard_normal(10)
    for lam in (1.0, 0.37):
        var = 0.6
        np.testing.assert_array_equal(O.lognormal(x, var, lam), lam * np.expm1(x - var / 2))
        a = math.sqrt(1 - var)
        np.testing.assert_array_equal(O.squared_normal(x, a, lam), lam * ((x - a) ** 2 - 1))


def test_getcl_and_nfields():
    cls = [np.asarray([i, j], dtype=float) for i in range(6) for j in range(i, -1, -1)]
    for i in range(6):
        for j in range(6):
            np.testing.assert_array_equal(O.getcl(cls, i, j), sorted([i, j], reverse=True))
    assert O.nfields_from_nspectra(10) == 4
    with pytest.raises(ValueError, match="invalid number of spectra"):
        O.nfields_from_nspectra(7)


# --------------------------------------------------------------------------
# geometry and transforms


def test_ring_geometry():
    for nside in (1, 2, 4, 16):
        npix = 12 * nside * nside
        total = 0
        prev_z = 2.0
        for r in range(1, 4 * nside):
            s, n, z, st, p0 = O.ring_info(nside, r)
            assert s == total
            total += n
            assert z < prev_z
            prev_z = z
            assert abs(st * st + z * z - 1) < 1e-15
        assert total == npix
    # nside 1: 4 pixels per ring, first pixel of ring 2 at phi = 0, ring 1 at pi/4
    assert O.ring_info(1, 1)[4] == pytest.approx(math.pi / 4)
    assert O.ring_info(1, 2)[4] == 0.0
    assert O.ring_info(1, 2)[2] == 0.0
    # equal-area: mean z over pixels is 0 and pixels are uniform in z per cap
    th, _ = O.pixel_angles(8)
    assert abs(np.mean(np.cos(th))) < 1e-14


def test_closed_forms_low_l():
    nside, lmax = 4, 2
    th, ph = O.pixel_angles(nside)
    # Y_00
    a = np.zeros(O.nalm(lmax), complex)
    a[_hp(lmax, 0, 0)] = 1.0
    np.testing.assert_allclose(O.alm2map(a, nside), 1 / math.sqrt(4 * math.pi), rtol=1e-14)
    # Y_10
    a[:] = 0
    a[_hp(lmax, 1, 0)] = 1.0
    np.testing.assert_allclose(O.alm2map(a, nside), math.sqrt(3 / (4 * math.pi)) * np.cos(th),
                               atol=1e-15)
    # a_11 = 1: 2 Re(Y_11) = -2 sqrt(3/8pi) sin th cos phi (Condon-Shortley)
    a[:] = 0
    a[_hp(lmax, 1, 1)] = 1.0
    np.testing.assert_allclose(O.alm2map(a, nside),
                               -2 * math.sqrt(3 / (8 * math.pi)) * np.sin(th) * np.cos(ph),
                               atol=1e-15)


@pytest.mark.parametrize("nside,lmax", [(1, 3), (2, 5), (4, 11), (8, 23)])
def test_alm2map_brute_force(nside, lmax):
    rng = np.random.default_rng(nside)
    alm = rng.standard_normal(O.nalm(lmax)) + 1j * rng.standard_normal(O.nalm(lmax))
    th, ph = O.pixel_angles(nside)
    ref = _brute_synth(alm, lmax, th, ph)
    got = O.alm2map(alm, nside, nthreads=2)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("nside,lmax", [(1, 3), (4, 11), (8, 23), (64, 300)])
def test_symmetric_synthesis_matches(nside, lmax):
    """The CPU-baseline synthesis (equatorial symmetry, early scale folding)
    against scipy brute force at small sizes and against the plain oracle at
    lmax > 4 nside (polar rings whose high-m values start far below range)."""
    rng = np.random.default_rng(nside + lmax)
    alm = rng.standard_normal(O.nalm(lmax)) + 1j * rng.standard_normal(O.nalm(lmax))
    got = O.alm2map_sym(alm, nside, nthreads=2)
    if nside <= 8:
        th, ph = O.pixel_angles(nside)
        ref = _brute_synth(alm, lmax, th, ph)
    else:
        ref = O.alm2map(alm, nside, nthreads=2)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))


def test_ring_dft_against_numpy():
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 4, 7, 12, 20, 28, 44, 64, 97, 100, 8156):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        np.testing.assert_allclose(O.dft(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * n)
        np.testing.assert_allclose(O.dft(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-12 * n)


def test_legendre_orthonormal_high_l():
    """Gauss-Legendre quadrature is exact for lambda_lm lambda_l'm when l + l' < 2N,
    so the Gram matrix must be the identity / 2 pi.  Covers l up to 2500 with
    seeds that underflow FP64 by hundreds of orders of magnitude near the poles."""
    lmax = 2500
    N = lmax + 1
    x, w = roots_legendre(N)
    th = np.arccos(x)
    for m in (0, 1, 700, 2400):
        lam = np.stack([O.legendre(m, lmax, t) for t in th])  # (N, lmax-m+1)
        gram = 2 * np.pi * (lam * w[:, None]).T @ lam
        err = np.abs(gram - np.eye(gram.shape[0]))
        # the median entry is at round-off (~1e-14); the worst entries at m <= 1
        # are limited by the accuracy of the polar quadrature nodes themselves
        assert np.median(err) < 1e-13, (m, np.median(err))
        assert err.max() < 1e-9, (m, err.max())


def test_legendre_scaling_against_arbitrary_exponent():
    """mpmath floats have an unbounded exponent: run the recurrence there without
    any rescaling and compare, at seeds far below the FP64 range."""
    import mpmath as mp

    mp.mp.dps = 40
    # seeds sin^m(theta) of 1e-430, 1e-700, 1e-1050, 1e-346; the functions turn
    # on (l ~ m / sin theta) before lmax so the compared tail is O(1)
    for m, lmax, theta in [(4000, 6000, 0.9), (1000, 6000, 0.2), (1500, 8000, 0.2),
                           (300, 5000, 0.07)]:
        st, x = mp.sin(mp.mpf(theta)), mp.cos(mp.mpf(theta))
        v = 1 / mp.sqrt(4 * mp.pi)
        for k in range(1, m + 1):
            v *= -mp.sqrt(mp.mpf(2 * k + 1) / (2 * k)) * st
        p0, p1 = mp.mpf(0), v
        vals = [p1]
        for l in range(m + 1, lmax + 1):
            a = mp.sqrt(mp.mpf(4 * l * l - 1) / (l * l - m * m))
            b = mp.sqrt(mp.mpf((l - 1) ** 2 - m * m) / (4 * (l - 1) ** 2 - 1))
            p0, p1 = p1, a * (x * p1 - b * p0)
            vals.append(p1)
        got = O.legendre(m, lmax, theta)
        ref = np.asarray([float(v) for v in vals])
        scale = np.max(np.abs(ref))
        assert scale > 1e-3
        assert np.max(np.abs(got - ref)) < 1e-12 * scale


def test_phi_rotation_invariant():
    """Rotating alm by exp(-i m alpha) with alpha = 2 pi / n on the equator
    shifts that ring by one pixel: f_rot(phi) = f(phi - alpha)."""
    nside, lmax = 8, 20
    rng = np.random.default_rng(5)
    alm = rng.standard_normal(O.nalm(lmax)) + 1j * rng.standard_normal(O.nalm(lmax))
    for l in range(lmax + 1):
        alm[_hp(lmax, l, 0)] = alm[_hp(lmax, l, 0)].real
    alpha = 2 * np.pi / (4 * nside)
    rot = alm.copy()
    for m in range(lmax + 1):
        for l in range(m, lmax + 1):
            rot[_hp(lmax, l, m)] *= np.exp(-1j * m * alpha)
    a = O.alm2map(alm, nside)
    b = O.alm2map(rot, nside)
    s, n, *_ = O.ring_info(nside, 2 * nside)
    np.testing.assert_allclose(b[s:s + n], np.roll(a[s:s + n], 1), atol=1e-12)


def test_map2alm_adjoint_and_round_trip():
    nside, lmax = 8, 12
    npix = 12 * nside * nside
    rng = np.random.default_rng(11)
    alm = rng.standard_normal(O.nalm(lmax)) + 1j * rng.standard_normal(O.nalm(lmax))
    for l in range(lmax + 1):
        alm[_hp(lmax, l, 0)] = alm[_hp(lmax, l, 0)].real
    f = rng.standard_normal(npix)
    Sa = O.alm2map(alm, nside)
    Af = O.map2alm(f, lmax, niter=0)
    # <S a, f> = (npix/4pi) Re sum_m (2 - delta_m0) conj(a) (A f)
    lhs = np.dot(Sa, f)
    w = np.asarray([1.0 if m == 0 else 2.0 for m in range(lmax + 1)
                    for _ in range(m, lmax + 1)])
    rhs = npix / (4 * np.pi) * np.sum(w * (np.conj(alm) * Af).real)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # band-limited round trip converges with Jacobi iterations
    e0, e3, e10 = (np.max(np.abs(O.map2alm(Sa, lmax, niter=k) - alm)) for k in (0, 3, 10))
    assert e3 < e0 / 100
    assert e10 < 1e-10


def test_almxfl_placement():
    lmax = 12
    rng = np.random.default_rng(9)
    alm = rng.standard_normal(O.nalm(lmax)) + 1j * rng.standard_normal(O.nalm(lmax))
    np.testing.assert_array_equal(O.almxfl(alm, np.ones(lmax + 1)), alm)
    for L0 in (0, 1, 5, lmax):
        f = np.zeros(lmax + 1)
        f[L0] = 2.0
        out = O.almxfl(alm, f)
        for m in range(lmax + 1):
            for l in range(m, lmax + 1):
                i = _hp(lmax, l, m)
                assert out[i] == (2.0 * alm[i] if l == L0 else 0.0)
    # zero beyond len(fl)
    out = O.almxfl(alm, np.ones(4))
    assert all(out[_hp(lmax, l, m)] == 0 for m in range(lmax + 1) for l in range(max(m, 4), lmax + 1))


def test_cltovar_against_map_variance():
    nside, lmax = 16, 30
    cl = 1.0 / (np.arange(lmax + 1) + 1.0) ** 2
    rng = np.random.default_rng(2)
    vals = []
    for _ in range(40):
        alm = np.zeros(O.nalm(lmax), complex)
        for m in range(lmax + 1):
            for l in range(m, lmax + 1):
                if m == 0:
                    alm[_hp(lmax, l, 0)] = rng.standard_normal() * math.sqrt(cl[l])
                else:
                    alm[_hp(lmax, l, m)] = (rng.standard_normal() + 1j * rng.standard_normal()) \
                        * math.sqrt(cl[l] / 2)
        vals.append(np.mean(O.alm2map(alm, nside) ** 2))
    assert abs(np.mean(vals) / O.cltovar(cl) - 1) < 0.1


def test_spin_against_eth_operator():
    """sY_lm from the spin-raising operator applied symbolically to Y_lm, and
    the oracle's [Q, U] synthesis against -sum_{l, |m| <= l} (E + iB) sY_lm
    with E_{l,-m} = (-1)^m conj(E_lm) evaluated by brute force."""
    import sympy as sp

    th, ph = sp.symbols("theta phi", real=True)

    def eth(f, s):
        g = sp.sin(th) ** (-s) * f
        return -sp.sin(th) ** s * (sp.diff(g, th) + sp.I / sp.sin(th) * sp.diff(g, ph))

    def sY(l, m, s):
        f = sp.Ynm(l, m, th, ph).expand(func=True)
        for k in range(s):
            f = eth(f, k)
        return sp.lambdify((th, ph), f * sp.sqrt(sp.factorial(l - s) / sp.factorial(l + s)),
                           "numpy")

    lmax, nside = 4, 2
    for s in (1, 2):
        rng = np.random.default_rng(s)
        E = rng.standard_normal(O.nalm(lmax)) + 1j * rng.standard_normal(O.nalm(lmax))
        B = rng.standard_normal(O.nalm(lmax)) + 1j * rng.standard_normal(O.nalm(lmax))
        for l in range(lmax + 1):
            E[_hp(lmax, l, 0)] = E[_hp(lmax, l, 0)].real
            B[_hp(lmax, l, 0)] = B[_hp(lmax, l, 0)].real
        theta, phi = O.pixel_angles(nside)
        P = np.zeros(theta.size, complex)
        for l in range(s, lmax + 1):
            for m in range(-l, l + 1):
                am = abs(m)
                e, b = E[_hp(lmax, l, am)], B[_hp(lmax, l, am)]
                if m < 0:
                    e, b = (-1) ** am * np.conj(e), (-1) ** am * np.conj(b)
                P += -(e + 1j * b) * sY(l, m, s)(theta, phi)
        q, u = O.alm2map_spin([E, B], nside, s, lmax)
        np.testing.assert_allclose(q, P.real, atol=1e-12)
        np.testing.assert_allclose(u, P.imag, atol=1e-12)


def test_generate_invariants():
    """The reference's own invariants for generate (test_fields.py:362-438)."""
    gls = [np.asarray([1.0, 0.5, 0.1])]
    a = list(O.generate_grf_alm(gls, rng=np.random.default_rng(42)))
    b = list(O.generate_grf_alm(gls, ncorr=1, rng=np.random.default_rng(42)))
    np.testing.assert_array_equal(a[0], b[0])
    with pytest.raises(ValueError, match="all gls are empty"):
        list(O.generate_grf_alm([np.asarray([])]))
    gls = [np.ones(10), np.ones(10), np.ones(10)]
    fields = [lambda x, var: x, lambda x, var: x**2]
    res = list(O.generate(fields, gls, nside=4, rng=np.random.default_rng(1)))
    np.testing.assert_allclose(res[1], res[0] ** 2, atol=1e-5)
    with pytest.raises(ValueError, match="mismatch between number of fields and gls"):
        list(O.generate(fields, [np.ones(10), np.ones(10)], nside=4))


def test_philox_known_answers():
    """Philox4x32-10 against the generator's published known-answer vectors
    (Random123 distribution, kat_vectors: counter, key -> output)."""
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF),
            (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    ctr = np.array([c for c, _, _ in kat], dtype=np.uint32)
    key = np.array([k for _, k, _ in kat], dtype=np.uint32)
    out = O.philox4x32_10(ctr, key)
    assert out.tolist() == [list(e) for _, _, e in kat]
    # Box-Muller on top: standard complex normals
    for shell in (2, 3):  # words 0, 1 and words 2, 3 of the same calls
        z = O.device_normals(11, shell, np.arange(200000))
        assert abs(z.real.mean()) < 0.02 and abs(z.imag.mean()) < 0.02
        assert abs(z.real.var() - 1) < 0.02 and abs(z.imag.var() - 1) < 0.02
        assert abs(np.mean(z.real * z.imag)) < 0.02
    # the two shells of one call are uncorrelated, and shell 3's value at
    # g = 5 comes from words 2 and 3 of the call with counter (5, 0, 1, 0)
    z2, z3 = (O.device_normals(11, s_, np.arange(200000)) for s_ in (2, 3))
    assert abs(np.mean(z2.real * z3.real)) < 0.02 and abs(np.mean(z2.imag * z3.imag)) < 0.02
    w = O.philox4x32_10(np.array([[5, 0, 1, 0]], dtype=np.uint32), np.array([[11, 0]], dtype=np.uint32))
    u1 = (np.float64(np.float32(w[0, 2])) + 0.5) * 2.0**-32
    u2 = float(w[0, 3] >> 8) * 2.0**-24
    r = np.sqrt(-2.0 * np.log(u1))
    assert abs(O.device_normals(11, 3, np.array([5]))[0] - r * np.exp(2j * np.pi * u2)) < 1e-12


class _UnitDrawRng:
    """Stand-in generator for the window test: the ``hot``-th call of
    standard_normal returns real parts 1 and imaginary parts 0, every other
    call zeros, so each yielded alm is the coefficient of that one draw."""

    def __init__(self, hot: int):
        self.hot, self.calls = hot, 0

    def standard_normal(self, shape):
        out = np.zeros(shape)
        if self.calls == self.hot:
            out[:, 0] = 1.0
        self.calls += 1
        return out


@pytest.mark.parametrize("nfields,ncorr", [(5, 2), (6, 3), (4, 1), (5, 0)])
def test_generate_window_implies_the_banded_covariance(nfields, ncorr):
    """The multi-shell window of glass/fields.py:401-425 (deque of the last
    k + 1 normal vectors, offset mis, alm = sum_i w[:, i + mis] z_i): with the
    draws as unit vectors, the yielded alm give the coefficient matrix L_l of
    every shell on every draw; 2 L L^T (E|z|^2 = 2 for complex normals) must be
    the target covariance C_l(t, u) inside the band |t - u| <= ncorr and zero
    outside it.  A wrong window offset, deque order or trimming breaks this
    for shells past ncorr + 1."""
    nl = 7
    ell = np.arange(nl, dtype=float)
    auto = [(1.0 + 0.3 * i) * (ell + 1.0) ** -1.1 for i in range(nfields)]
    rho = (0.55, 0.3, 0.12)
    gls = []
    for i in range(nfields):
        for j in range(i, -1, -1):
            d = i - j
            if d == 0:
                gls.append(auto[i])
            elif d <= len(rho):
                gls.append(rho[d - 1] * np.sqrt(auto[i] * auto[j]))
            else:
                gls.append(np.zeros(0))
    # L[l, t, s]: coefficient of draw s in shell t, read at the (l, m = 0) slot
    L = np.zeros((nl, nfields, nfields))
    for s in range(nfields):
        alms = list(O.generate_grf_alm(gls, ncorr=ncorr, rng=_UnitDrawRng(s)))
        assert len(alms) == nfields
        for t, a in enumerate(alms):
            L[:, t, s] = a[:nl].real  # healpy order: (l, 0) at index l
            # every m of a given l carries the same coefficient
            for m in range(1, nl):
                idx = [O.nalm(nl - 1) - O.nalm(nl - 1 - m) + (l - m) for l in range(m, nl)]
                assert np.allclose(a[idx].real, L[m:, t, s], rtol=0, atol=1e-15)
    K = 2.0 * np.einsum("lts,lus->ltu", L, L)
    for t in range(nfields):
        for u in range(nfields):
            want = O.getcl(gls, t, u, lmax=nl - 1) if abs(t - u) <= ncorr else np.zeros(nl)
            assert np.allclose(K[:, t, u], want, rtol=1e-12, atol=1e-15), (t, u)
    # causal and banded: shell t uses only draws t - ncorr .. t
    for t in range(nfields):
        for s in range(nfields):
            if s > t or s < t - ncorr:
                assert np.all(L[:, t, s] == 0.0), (t, s)


# ---- pixel windows -------------------------------------------------------


@pytest.mark.parametrize("nside", [1, 2, 4])
def test_face_geometry_gives_the_pixel_centres(nside):
    """Face coordinates at pixel centres ((ix + 1/2) / nside, (iy + 1/2) / nside)
    land on the HEALPix pixel centres of the ring geometry, one for one."""
    th, ph = O.pixel_angles(nside)
    ref = np.stack([np.cos(th), np.mod(ph, 2 * np.pi)], -1)
    got = []
    for f in range(12):
        for ix in range(nside):
            for iy in range(nside):
                z, p = O.face_zphi(f, (ix + 0.5) / nside, (iy + 0.5) / nside)
                got.append((float(z), float(np.mod(p, 2 * np.pi))))
    got = np.array(sorted(got))
    ref = ref[np.lexsort((ref[:, 1], ref[:, 0]))]
    assert got.shape == ref.shape
    assert np.allclose(got, ref, atol=1e-12)


def test_pixel_quadrature_is_area_weighted():
    """The nodes of a pixel integrate z and z^2 like the exact area average of
    its ring range, for a belt pixel (z linear in x + y) and a polar one."""
    nside = 4
    for f, ix, iy in ((4, 1, 2), (0, 3, 3), (0, 2, 3)):
        z, _, w = O.pixel_nodes(f, ix, iy, nside, G=6)
        assert abs(w.sum() - 1) < 1e-14
        # brute force: a fine uniform grid in the face coordinates
        g = (np.arange(400) + 0.5) / 400
        X, Y = np.meshgrid((ix + g) / nside, (iy + g) / nside)
        zb, _ = O.face_zphi(f, X, Y)
        assert abs(np.sum(w * z) - zb.mean()) < 1e-6
        assert abs(np.sum(w * z * z) - (zb**2).mean()) < 1e-6


def test_wigner_d_matches_the_harmonics():
    """sqrt((2l+1)/4pi) |d^l_{m,-s}| equals |sY_lm| of the oracle's Legendre
    (s = 0, pinned to scipy) and spin-weighted (s = 2, pinned by sympy) tables."""
    theta = 0.7
    for m in range(0, 5):
        lam = O.legendre(m, 8, theta)
        for l in range(m, 9):
            d = O.wigner_d(l, m, 0, np.array([theta]))[0]
            assert abs(abs(d) * math.sqrt((2 * l + 1) / (4 * math.pi)) - abs(lam[l - m])) < 1e-13
        sp, _ = O.spin_legendre(2, m, 8, theta)
        l0 = max(m, 2)
        for l in range(l0, 9):
            d = O.wigner_d(l, m, -2, np.array([theta]))[0]
            assert abs(abs(d) * math.sqrt((2 * l + 1) / (4 * math.pi)) - abs(sp[l - l0])) < 1e-13


@pytest.mark.parametrize("nside,spin", [(1, 0), (1, 1), (1, 2), (2, 0), (2, 1), (2, 2)])
def test_pixwin_pair_form_matches_definition(nside, spin):
    """The pair form (what the device computes) against the definition evaluated
    directly from the harmonics, every pixel: pins the addition theorem and the
    frame-rotation angle of the spin-2 window; both quadratures converged to
    ~1e-7 at these (largest) pixel sizes."""
    lmax = 4 * nside
    d = O.pixwin_definition(nside, lmax, spin, G=10)
    p = O.pixwin_pairs(nside, lmax, spin)
    assert np.max(np.abs(d - p)) < 3e-7
    if spin:
        assert np.all(p[:spin] == 0)


@pytest.mark.parametrize("nside", [2, 4])
def test_pixwin_symmetry_reduction(nside):
    """The reduced pixel set with its multiplicities equals the sum over every pixel."""
    for spin in (0, 2):
        full = O.pixwin_pairs(nside, 4 * nside, spin, reduced=False)
        red = O.pixwin_pairs(nside, 4 * nside, spin)
        assert np.max(np.abs(full - red)) < 1e-13


def test_pixwin_limits():
    """w_0 = 1; the windows decrease with l up to 4 nside and stay in (0, 1]."""
    for nside in (4, 8):
        pw0 = O.pixwin_pairs(nside, 4 * nside, 0)
        pw2 = O.pixwin_pairs(nside, 4 * nside, 2)
        assert abs(pw0[0] - 1) < 1e-14
        assert np.all(np.diff(pw0) < 0) and pw0[-1] > 0
        assert np.all(np.diff(pw2[2:]) < 0) and pw2[-1] > 0 and pw2[2] < 1


# ---- multi-plane convergence ----------------------------------------------


def test_multi_plane_matrix_identity(cosmo, shells):
    """tests/core/test_lensing.py:64-86: the matrix is lower triangular and
    maps random mass planes to the convergence sequence."""
    mat = O.multi_plane_matrix(shells, cosmo)
    assert np.array_equal(mat, np.tril(mat))
    rng = np.random.default_rng(1)
    deltas = rng.random((len(shells), 10))
    conv = O.MultiPlaneConvergence(cosmo)
    kappas = []
    for i, w in enumerate(shells):
        conv.add_window(deltas[i], w)
        kappas.append(conv.kappa.copy())
    assert np.allclose(mat @ deltas, np.stack(kappas), rtol=1e-12, atol=1e-15)
    # the first source plane sees no mass in front of it
    assert np.all(mat[0] == 0)


def test_multi_plane_weights_identity(cosmo, shells):
    """tests/core/test_lensing.py:89-120."""
    w_out = O.multi_plane_weights(np.eye(len(shells)), shells, cosmo)
    assert np.array_equal(w_out, np.triu(w_out, k=1))
    rng = np.random.default_rng(2)
    deltas = rng.random((len(shells), 10))
    weights = rng.random((len(shells), 3))
    conv = O.MultiPlaneConvergence(cosmo)
    kappa = 0
    for i, w in enumerate(shells):
        conv.add_window(deltas[i], w)
        kappa = kappa + weights[i][..., None] * conv.kappa
    kappa /= np.sum(weights, axis=0)[..., None]
    wmat = O.multi_plane_weights(weights, shells, cosmo)
    assert np.allclose(np.einsum("ij,ik", wmat, deltas), kappa, rtol=1e-12)
    with pytest.raises(ValueError, match="shape mismatch"):
        O.multi_plane_weights(np.ones(3), shells, cosmo)


def test_multi_plane_source_order(cosmo):
    conv = O.MultiPlaneConvergence(cosmo)
    conv.add_plane(np.zeros(4), 1.0)
    with pytest.raises(ValueError, match="source redshift must be increasing"):
        conv.add_plane(np.zeros(4), 1.0)


# ---- points -----------------------------------------------------------------


@pytest.mark.parametrize("nside", [1, 2, 3, 4, 8, 16])
def test_ring_geometry_of_points(nside):
    """ang2pix_ring inverts the pixel centres of the ring geometry, and the
    face coordinates of ring_to_xyf at (ix + 1/2, iy + 1/2) land on them."""
    th, ph = O.pixel_angles(nside)
    assert np.array_equal(O.ang2pix_ring(nside, th, ph), np.arange(12 * nside**2))
    f, ix, iy = O.ring_to_xyf(nside, np.arange(12 * nside**2))
    assert np.all((ix >= 0) & (ix < nside) & (iy >= 0) & (iy < nside))
    lon, lat = O.face_lonlat(f, (ix + 0.5) / nside, (iy + 0.5) / nside)
    assert np.allclose(np.radians(90 - lat), th, atol=1e-12)
    d = np.mod(np.radians(lon) - ph + np.pi, 2 * np.pi) - np.pi
    assert np.allclose(d, 0, atol=1e-12)


def test_device_poisson_moments_and_limits():
    lam = np.array([0.0, 0.5, 3.0, 31.9, 32.0, 33.0, 100.5])
    c = O.device_poisson(np.repeat(lam, 4000), 11, 3).reshape(len(lam), 4000)
    assert np.all(c[0] == 0) and c.dtype == np.int64
    se = np.sqrt(lam[1:] / 4000)
    assert np.all(np.abs(c[1:].mean(1) - lam[1:]) < 5 * se)
    assert np.all(np.abs(c[1:].var(1) / lam[1:] - 1) < 0.15)
    # deterministic in (seed, population, pixel)
    assert np.array_equal(O.device_poisson(lam, 11, 3), O.device_poisson(lam, 11, 3))
    assert not np.array_equal(O.device_poisson(np.full(64, 5.0), 11, 3),
                              O.device_poisson(np.full(64, 5.0), 11, 4))


def test_device_positions_in_their_pixels_and_uniform():
    """Every point lies in its pixel; inside one nside-1 pixel the points fall
    evenly on its 64 nside-8 sub-pixels (chi-square)."""
    nside = 4
    counts = np.random.default_rng(1).integers(0, 5, 12 * nside**2)
    lon, lat, ipix = O.device_positions(nside, counts, 42, 0)
    assert lon.size == counts.sum() and np.all((lon >= 0) & (lon < 360))
    assert np.array_equal(O.ang2pix_ring(nside, np.radians(90 - lat), np.radians(lon)), ipix)
    counts = np.zeros(12, np.int64)
    counts[5] = 64000
    lon, lat, _ = O.device_positions(1, counts, 42, 0)
    sub = O.ang2pix_ring(8, np.radians(90 - lat), np.radians(lon))
    h = np.bincount(sub, minlength=768)
    occ = h[h > 0]
    assert occ.size == 64
    chi2 = np.sum((occ - 1000.0) ** 2 / 1000.0)
    assert chi2 < 120  # 63 degrees of freedom


def test_expected_counts_reference_cases():
    """glass/points.py:214-348 on the reference test's cases
    (tests/core/test_points.py:235-386): delta = -1 gives no points, a zero
    contrast gives ngal per arcmin^2, vis scales, negative counts are clipped."""
    npix = 12 * 16**2
    n = O.expected_counts(np.full(npix, -1.0), 1e-3, None)
    assert np.all(n == 0)
    n = O.expected_counts(np.zeros(npix), 2e-3, 0.8, np.ones(npix))
    assert np.allclose(n.sum(), O.ARCMIN2_SPHERE * 2e-3, rtol=1e-12)
    vis = np.ones(npix)
    vis[: npix // 2] = 0
    assert np.all(O.expected_counts(np.zeros(npix), 1e-3, 0.8, vis)[: npix // 2] == 0)
    n = O.expected_counts(np.full(npix, -3.0), 1e-3, 1.0)
    assert np.all(n == 0)
    d = np.random.default_rng(0).lognormal(0, 0.3, npix) - 1
    n = O.expected_counts(d, 1e-3, 0.8, remove_monopole=True)
    assert np.allclose(n.sum(), O.ARCMIN2_SPHERE * 1e-3, rtol=1e-10)
    ll = O.expected_counts(d, 1e-3, 1.0, model="loglinear")
    assert np.allclose(ll, O.expected_counts(d, 1e-3, None), rtol=1e-12)


# ---- spectra preparation ------------------------------------------------------


def test_gauss_legendre_nodes():
    """Nodes and weights: numpy's leggauss for small n; at n = 301, roots of
    P_n and the weights 2 / ((1 - x^2) P_n'(x)^2) in 40-digit arithmetic
    (mpmath) at the ends, the middle and a node between; sum of weights 2."""
    for n in (1, 2, 3, 8, 33):
        x, w = O.gl_nodes(n)
        rx, rw = np.polynomial.legendre.leggauss(n)
        assert np.allclose(x, rx, rtol=0, atol=2e-16) and np.allclose(w, rw, rtol=1e-14)
    import mpmath

    n = 301
    x, w = O.gl_nodes(n)
    assert abs(w.sum() - 2.0) < 1e-14
    with mpmath.workdps(40):
        for k in (0, 1, 77, 150, 300):
            r = mpmath.mpf(x[k])
            for _ in range(4):  # Newton from the float64 node
                d = n * (r * mpmath.legendre(n, r) - mpmath.legendre(n - 1, r)) / (r * r - 1)
                r -= mpmath.legendre(n, r) / d
            d = n * (r * mpmath.legendre(n, r) - mpmath.legendre(n - 1, r)) / (r * r - 1)
            wr = 2 / ((1 - r * r) * d * d)
            assert abs(float(r) - x[k]) <= 1.2e-16
            assert abs(float(wr) - w[k]) <= 1e-14 * float(wr)


def test_gauss_legendre_spectrum_transforms():
    """C_l <-> C(theta) on n Gauss-Legendre nodes: exact round trip for l < n,
    a single multipole gives (2L+1)/(4 pi) P_L, a constant correlation gives
    only the monopole 4 pi c (closed forms)."""
    rng = np.random.default_rng(0)
    cl = rng.random(40)
    assert np.allclose(O.gl_corrtocl(O.gl_cltocorr(cl)), cl, rtol=0, atol=1e-12)
    L, n = 7, 20
    d = np.zeros(n)
    d[L] = 1.0
    x, _ = O.gl_nodes(n)
    from scipy.special import eval_legendre

    assert np.allclose(O.gl_cltocorr(d), (2 * L + 1) / (4 * np.pi) * eval_legendre(L, x),
                       atol=1e-14)
    c = O.gl_corrtocl(np.full(n, 0.25))
    assert abs(c[0] - np.pi) < 1e-13 and np.all(np.abs(c[1:]) < 1e-13)


def test_cls_file_round_trip(tmp_path):
    """glass/user.py:42-86: save_cls / load_cls keep every spectrum, empty ones
    included (host-only, no device)."""
    from glass_amd.user import load_cls, save_cls

    cls = [np.arange(5.0), np.zeros(0), np.linspace(0, 1, 9), np.ones(3)]
    save_cls(tmp_path / "cls.npz", cls)
    back = load_cls(tmp_path / "cls.npz")
    assert len(back) == len(cls)
    for a, b in zip(back, cls):
        assert np.array_equal(a, b)


def test_spectra_order_helpers():
    """glass/fields.py:563-605 docstring examples."""
    from glass_amd.fields import enumerate_spectra, spectra_indices

    assert list(enumerate_spectra([[1, 2, 3], [4, 5, 6], [7, 8, 9]])) == [
        (0, 0, [1, 2, 3]), (1, 1, [4, 5, 6]), (1, 0, [7, 8, 9])]
    assert spectra_indices(3).tolist() == [[0, 0], [1, 1], [1, 0], [2, 2], [2, 1], [2, 0]]


# ---- pixel quadrature weights ------------------------------------------------


@pytest.mark.parametrize("nside", [1, 2, 4])
def test_pixel_weights_oracle(nside):
    """The dense solve: the quadrature conditions hold, the correction is the
    minimum-norm one (it lies in the span of the conditions' rows), the weights
    sum to 4 pi and carry the HEALPix symmetries (rotation by pi/2 within a
    ring, phi -> -phi, the equatorial mirror), and the weighted analysis of a
    map band-limited to nside is exact without iterations."""
    L = 2 * nside
    w = O.pixel_weights(nside, L)
    Y = O.ylm_matrix(nside, L)
    A = O._constraint_rows(Y, L)
    b = np.zeros(A.shape[0])
    b[0] = np.sqrt(4 * np.pi)
    assert np.max(np.abs(A @ w - b)) < 1e-12
    assert abs(w.sum() - 4 * np.pi) < 1e-12
    dw = w - 4 * np.pi / w.size
    coef = np.linalg.lstsq(A.T, dw, rcond=None)[0]
    assert np.max(np.abs(A.T @ coef - dw)) < 1e-12
    # symmetries, ring by ring
    for r in range(1, 4 * nside):
        start, n, z, _, phi0 = O.ring_info(nside, r)
        wr = w[start:start + n]
        assert np.allclose(wr, np.roll(wr, n // 4), rtol=0, atol=1e-13)
        # phi -> -phi: pixel j -> n - 1 - j (phi0 = pi / n) or -j (phi0 = 0)
        mirror = wr[::-1] if phi0 > 0 else np.roll(wr[::-1], 1)
        assert np.allclose(wr, mirror, rtol=0, atol=1e-13)
        s2, n2, _, _, _ = O.ring_info(nside, 4 * nside - r)
        assert np.allclose(np.sort(wr), np.sort(w[s2:s2 + n2]), rtol=0, atol=1e-13)
    # exact analysis of a band-limited map (lmax = nside) with no iteration
    rng = np.random.default_rng(nside)
    lm = nside
    alm = rng.standard_normal(O.nalm(lm)) + 1j * rng.standard_normal(O.nalm(lm))
    alm[: lm + 1] = alm[: lm + 1].real
    m = O.alm2map(alm, nside, lm)
    got = O.map2alm_weighted(m, w, lm, niter=0)
    assert np.max(np.abs(got - alm)) < 1e-11


@pytest.mark.parametrize("nside", [2, 4, 8])
def test_pixel_weights_cg_equals_dense(nside):
    """The oracle's conjugate-gradient form (for larger nside) against the
    dense minimum-norm solve, and the weighted analysis through the uniform
    one against the dense harmonic matrix."""
    L = 2 * nside
    wd = O.pixel_weights(nside, L)
    wc = O.pixel_weights_cg(nside, L)
    assert np.max(np.abs(wc - wd)) < 1e-12 * np.max(np.abs(wd)) * 10
    rng = np.random.default_rng(nside)
    m = rng.standard_normal(wd.size)
    lm = 3 * nside - 1
    Y = O.ylm_matrix(nside, lm)
    ref = Y.conj().T @ (wd * m)
    got = O.map2alm_weighted(m, wd, lm, niter=0)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref)) * 10
