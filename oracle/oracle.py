"""FP64 CPU oracle for the GLASS field-generation path.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, by ``__graft_entry__.smoke``
and by the CPU-baseline leg of ``bench.py`` -- always to check results, never as
the thing measured or shipped.  The product package ``glass_amd`` never imports
this module.

Two halves:

* numpy restatements of the parts of the path the reference owns
  (``glass/fields.py``, ``glass/harmonics.py``, ``glass/grf/_transformations.py``,
  ``glass/lensing.py``), each citing the reference lines it follows;
* ctypes bindings of ``sht_oracle.c`` (FP64 HEALPix transforms, restating the
  algorithm healpy implements behind ``glass/healpix.py``).

Pinning (what checks each function, all in ``tests/test_oracle.py``):

* iternorm       -- explicit pinv/Cholesky restatement on the reference test's
                    literal covariances (tests/core/test_fields.py:31-104) and
                    its error cases (:107-119).
* cls2cov        -- literal vectors incl. buffer reuse (test_fields.py:166-252).
* multalm        -- literal vectors (tests/core/test_harmonics.py:15-52).
* glass_to_healpix_alm -- literal permutation (test_fields.py:649-655) and the
                    closed-form healpy index m(2 lmax + 1 - m)/2 + l.
* lognormal etc. -- closed forms (tests/core/grf/test_transformations.py).
* cltovar        -- closed form sum (2l+1) C_l / 4 pi checked against the
                    variance of a synthesised map (Parseval / addition theorem).
* alm2map        -- brute force with scipy.special.sph_harm_y at every pixel
                    centre (nside <= 8), closed forms Y_00 / Y_10 / Y_11,
                    Gauss-Legendre orthonormality of lambda_lm up to l = 2500,
                    the recurrence in mpmath's unbounded exponent range for
                    seeds 10^-346 .. 10^-1050 (checks the exponent scaling),
                    numpy.fft for the ring DFT, phi-rotation invariant.
* almxfl         -- a delta function in l keeps exactly the entries of that l
                    (placement in healpy order), ones keep everything.
* map2alm        -- adjointness <S a, f> = (npix / 4 pi) <a, A f> and round trip
                    of band-limited maps (lmax <= nside) with Jacobi iterations.
* alm2map_spin   -- spin-weighted harmonics from the spin-raising operator
                    applied symbolically (sympy) to Y_lm at small l.
* generate       -- the reference's own invariants (test_fields.py:362-438):
                    same seed -> same first field for ncorr None vs 1; two
                    identical gls -> field 1 equals field 0.
* philox4x32_10  -- the published known-answer vectors of Philox4x32-10
                    (Salmon et al., SC'11, Random123 distribution).  The glass_amd
                    device generator (not the reference's numpy stream)
                    is then checked against device_normals built on it.

* MultiPlaneConvergence -- the reference's own identities
                    (tests/core/test_lensing.py:64-120): multi_plane_matrix is lower
                    triangular and reproduces the convergence sequence of random
                    mass planes; multi_plane_weights reproduces weighted sums.
* pixwin         -- pixwin_definition evaluates the definition directly (pixel
                    averages of the spin-weighted harmonics from Wigner's explicit
                    d formula, every pixel, high-order quadrature); pixwin_pairs
                    (the pair form the device computes, exact per pair, over the
                    symmetry-reduced pixel set) is checked against it, against
                    itself over all pixels, and against w_0 = 1.

* pixel_weights  -- dense minimum-norm least squares on the harmonic matrix of
                    the pixel centres (scipy sph_harm_y); pinned by the
                    quadrature conditions it solves (sum_p w_p Y*_lm =
                    sqrt(4 pi) delta), by the correction lying in the row
                    space of the conditions (minimum norm), by Gauss-Legendre-
                    like exactness for band-limited products, and by the
                    HEALPix symmetries of the result.

Parity unpinned: healpy's pixel-weight files (use_pixel_weights=True) and the
tabulated pixel windows that healpy ships (pixwin here is computed from its
definition; the tables themselves are not available to compare with).
"""

from __future__ import annotations

import ctypes
import math
import os
import pathlib
import subprocess
from collections import deque

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "libsht_oracle.so"
_lib = None


def build() -> None:
    """Compile ``sht_oracle.c`` with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        dp = ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.c_int64
        L.oracle_alm2map.argtypes = [i64, i64, dp, dp, ctypes.c_int]
        L.oracle_alm2map_rings.argtypes = [i64, i64, dp, dp, i64, i64, ctypes.c_int]
        L.oracle_alm2map_sym.argtypes = [i64, i64, dp, dp, ctypes.c_int]
        L.oracle_map2alm.argtypes = [i64, i64, dp, dp, ctypes.c_int, ctypes.c_int]
        L.oracle_alm2map_spin.argtypes = [i64, i64, i64, dp, dp, dp, dp, ctypes.c_int]
        L.oracle_legendre.argtypes = [i64, i64, ctypes.c_double, ctypes.c_double, dp]
        L.oracle_spin_legendre.argtypes = [i64, i64, i64, ctypes.c_double, ctypes.c_double, dp, dp]
        L.oracle_dft.argtypes = [dp, dp, i64, ctypes.c_int]
        L.oracle_ring_info.argtypes = [i64, i64] + [ctypes.POINTER(i64)] * 2 + [dp] * 3
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _threads(nthreads: int | None) -> int:
    if nthreads is None:
        return len(os.sched_getaffinity(0))
    return int(nthreads)


# ---------------------------------------------------------------------------
# HEALPix geometry and transforms (C oracle)


def nside2npix(nside: int) -> int:
    return 12 * nside * nside


def npix2nside(npix: int) -> int:
    nside = math.isqrt(npix // 12)
    if 12 * nside * nside != npix:
        raise ValueError(f"{npix} is not a valid number of HEALPix pixels")
    return nside


def nalm(lmax: int) -> int:
    return (lmax + 1) * (lmax + 2) // 2


def lmax_from_nalm(n: int) -> int:
    lmax = (math.isqrt(8 * n + 1) - 3) // 2
    if nalm(lmax) != n:
        raise ValueError(f"{n} is not a valid number of alm")
    return lmax


def ring_info(nside: int, ring: int) -> tuple[int, int, float, float, float]:
    """(start pixel, pixel count, z, sin theta, phi0) of 1-based ring."""
    s, n = ctypes.c_int64(), ctypes.c_int64()
    z, st, p = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().oracle_ring_info(nside, ring, ctypes.byref(s), ctypes.byref(n),
                           ctypes.byref(z), ctypes.byref(st), ctypes.byref(p))
    return s.value, n.value, z.value, st.value, p.value


def pixel_angles(nside: int) -> tuple[np.ndarray, np.ndarray]:
    """theta, phi of every pixel centre in RING order."""
    npix = nside2npix(nside)
    theta = np.empty(npix)
    phi = np.empty(npix)
    for r in range(1, 4 * nside):
        s, n, z, st, p0 = ring_info(nside, r)
        theta[s:s + n] = math.atan2(st, z)
        phi[s:s + n] = p0 + 2 * math.pi * np.arange(n) / n
    return theta, phi


def legendre(m: int, lmax: int, theta: float) -> np.ndarray:
    """Normalised lambda_lm(theta), l = m..lmax (Condon-Shortley phase)."""
    out = np.zeros(lmax - m + 1)
    lib().oracle_legendre(m, lmax, math.cos(theta), math.sin(theta), _dp(out))
    return out


def spin_legendre(s: int, m: int, lmax: int, theta: float) -> tuple[np.ndarray, np.ndarray]:
    """(sY, -sY) theta parts for l = max(m, s)..lmax."""
    l0 = max(m, s)
    op = np.zeros(lmax - l0 + 1)
    om = np.zeros(lmax - l0 + 1)
    lib().oracle_spin_legendre(s, m, lmax, math.cos(theta), math.sin(theta), _dp(op), _dp(om))
    return op, om


def dft(x: np.ndarray, sign: int) -> np.ndarray:
    """x_j = sum_k a_k exp(sign 2 pi i j k / n)."""
    re = np.ascontiguousarray(x.real, dtype=np.float64).copy()
    im = np.ascontiguousarray(x.imag, dtype=np.float64).copy()
    lib().oracle_dft(_dp(re), _dp(im), len(re), sign)
    return re + 1j * im


def alm2map(alm: np.ndarray, nside: int, lmax: int | None = None,
            nthreads: int | None = None) -> np.ndarray:
    """FP64 scalar synthesis, healpy order in, RING map out."""
    alm = np.ascontiguousarray(alm, dtype=np.complex128)
    if lmax is None:
        lmax = lmax_from_nalm(alm.size)
    buf = alm.view(np.float64).copy()
    out = np.zeros(nside2npix(nside))
    lib().oracle_alm2map(nside, lmax, _dp(buf), _dp(out), _threads(nthreads))
    return out


def alm2map_rings(alm: np.ndarray, nside: int, r_first: int, r_step: int,
                  nthreads: int | None = None) -> np.ndarray:
    """Synthesis of rings r_first, r_first + r_step, ... only (1-based); the
    other pixels stay zero.  Used to time a bounded sample of a large map."""
    alm = np.ascontiguousarray(alm, dtype=np.complex128)
    lmax = lmax_from_nalm(alm.size)
    buf = alm.view(np.float64).copy()
    out = np.zeros(nside2npix(nside))
    lib().oracle_alm2map_rings(nside, lmax, _dp(buf), _dp(out), int(r_first), int(r_step),
                               _threads(nthreads))
    return out


def alm2map_sym(alm: np.ndarray, nside: int, nthreads: int | None = None) -> np.ndarray:
    """The full synthesis with equatorial symmetry (one recurrence per north
    ring and m serves the mirror ring): the tuned form timed as the CPU
    baseline, checked against ``alm2map``."""
    alm = np.ascontiguousarray(alm, dtype=np.complex128)
    lmax = lmax_from_nalm(alm.size)
    buf = alm.view(np.float64).copy()
    out = np.zeros(nside2npix(nside))
    lib().oracle_alm2map_sym(nside, lmax, _dp(buf), _dp(out), _threads(nthreads))
    return out


def map2alm(m: np.ndarray, lmax: int, niter: int = 3,
            nthreads: int | None = None) -> np.ndarray:
    """FP64 scalar analysis with healpy's Jacobi refinement."""
    m = np.ascontiguousarray(m, dtype=np.float64)
    nside = npix2nside(m.size)
    out = np.zeros(2 * nalm(lmax))
    lib().oracle_map2alm(nside, lmax, _dp(m), _dp(out), int(niter), _threads(nthreads))
    return out.view(np.complex128)


def alm2map_spin(alms, nside: int, spin: int, lmax: int,
                 nthreads: int | None = None) -> list[np.ndarray]:
    """FP64 spin-s synthesis of [E, B] into [Q, U] (HEALPix convention)."""
    e = np.ascontiguousarray(alms[0], dtype=np.complex128).view(np.float64).copy()
    b = np.ascontiguousarray(alms[1], dtype=np.complex128).view(np.float64).copy()
    q = np.zeros(nside2npix(nside))
    u = np.zeros(nside2npix(nside))
    lib().oracle_alm2map_spin(nside, lmax, spin, _dp(e), _dp(b), _dp(q), _dp(u),
                              _threads(nthreads))
    return [q, u]


def almxfl(alm: np.ndarray, fl: np.ndarray) -> np.ndarray:
    """alm_lm * fl[l], zero beyond len(fl) (healpy almxfl semantics)."""
    lmax = lmax_from_nalm(alm.size)
    out = np.array(alm, dtype=np.complex128, copy=True)
    f = np.zeros(lmax + 1)
    k = min(len(fl), lmax + 1)
    f[:k] = np.asarray(fl, dtype=np.float64)[:k]
    for m in range(lmax + 1):
        i0 = m * (2 * lmax + 1 - m) // 2 + m  # index of (l = m, m)
        out[i0:i0 + lmax - m + 1] *= f[m:]
    return out


# ---------------------------------------------------------------------------
# Reference-owned arithmetic, restated in numpy


def inv_triangle_number(t: int) -> int:
    """glass/fields.py:70-81."""
    n = math.floor(math.sqrt(2 * t))
    if n * (n + 1) // 2 != t:
        raise ValueError(f"not a triangle number: {t}")
    return n


def nfields_from_nspectra(nspectra: int) -> int:
    """glass/fields.py:84-98."""
    try:
        return inv_triangle_number(nspectra)
    except ValueError:
        raise ValueError(f"invalid number of spectra: {nspectra}") from None


def getcl(cls, i: int, j: int, lmax: int | None = None) -> np.ndarray:
    """glass/fields.py:525-560 (Christmas-tree index i(i+1)/2 + i - j)."""
    if j > i:
        i, j = j, i
    cl = np.asarray(cls[i * (i + 1) // 2 + i - j])
    if lmax is not None:
        if cl.shape[0] > lmax + 1:
            cl = cl[: lmax + 1]
        else:
            cl = np.pad(cl, (0, lmax + 1 - cl.shape[0]))
    return cl


def cls2cov(cls, nl: int, nf: int, nc: int):
    """glass/fields.py:191-236: one reused (nl, nc+1) buffer per field, halved."""
    cov = np.zeros((nl, nc + 1))
    end = 0
    for j in range(nf):
        begin, end = end, end + j + 1
        for i, cl in enumerate(cls[begin:end][: nc + 1]):
            if cl is None:
                cl = np.zeros(0)
            cl = np.asarray(cl, dtype=np.float64)
            if i == 0 and np.any(cl < 0):
                raise ValueError("negative values in cl")
            n = cl.shape[0]
            cov[:n, i] = cl
            cov[n:, i] = 0.0
        cov /= 2
        yield cov


def iternorm(cov):
    """glass/fields.py:101-188, restated as the conditional-Gaussian update.

    State per mode: the k x k matrix M mapping the last k unit normals to the
    last k field values' whitened form, the previous coefficient vector a and
    the previous conditional standard deviation s (arXiv:2302.01942 App. A).
    """
    M = a = s = None
    shape = None
    for i, row in enumerate(cov):
        row = np.asarray(row, dtype=np.float64)
        k = row.shape[-1] - 1
        if k < 0:
            raise ValueError("empty covariance matrix")
        if i == 0:
            shape = row.shape[:-1]
            M = np.zeros((*shape, k, k))
            a = np.zeros((*shape, k))
            s = np.ones(shape)
        else:
            if row.shape[:-1] != shape:
                raise ValueError("shape mismatch in covariance")
            # new bottom row of M: [-a^T M, 1] / s  (1 -> 0 where s == 0)
            aM = np.einsum("...i,...ij->...j", a, M)
            grown = np.zeros((*shape, M.shape[-2] + 1, M.shape[-1] + 1))
            grown[..., :-1, :-1] = M
            pos = s > 0
            safe_s = np.where(pos, s, 1.0)
            grown[..., -1, :-1] = -aM / safe_s[..., None]
            grown[..., -1, -1] = np.where(pos, 1.0, 0.0) / safe_s
            M = grown
        M = M[..., M.shape[-2] - k:, M.shape[-1] - k:]
        c = row[..., :0:-1]  # correlations, oldest first
        a = np.einsum("...ij,...j->...i", M, c)
        var = row[..., 0] - np.einsum("...i,...i->...", a, a)
        if np.any(var < 0):
            raise ValueError("covariance matrix is not positive definite")
        s = np.sqrt(var)
        yield np.concatenate([a, s[..., None]], axis=-1)


def multalm(alm: np.ndarray, bl: np.ndarray) -> np.ndarray:
    """glass/harmonics.py:17-47: GLASS (l-major) order, alm[l(l+1)/2+m] *= bl[l]."""
    bl = np.asarray(bl)
    return np.asarray(alm) * np.repeat(bl, np.arange(bl.size) + 1)


def glass_to_healpix_alm(alm: np.ndarray) -> np.ndarray:
    """glass/fields.py:943-962: l(l+1)/2 + m  ->  m(2 lmax + 1 - m)/2 + l."""
    n = inv_triangle_number(alm.size)
    ell = np.arange(n)
    return np.concatenate([alm[ell[m:] * (ell[m:] + 1) // 2 + m] for m in range(n)])


def cltovar(cl: np.ndarray) -> float:
    """transformcl.cltovar as called at glass/fields.py:890: sum (2l+1) C_l / 4 pi."""
    cl = np.asarray(cl, dtype=np.float64)
    ell = np.arange(cl.shape[-1])
    return float(np.sum((2 * ell + 1) * cl) / (4 * np.pi))


def lognormal(x, var: float, lamda: float = 1.0):
    """glass/grf/_transformations.py:83-89."""
    y = np.expm1(x - var / 2)
    if lamda != 1.0:
        y = lamda * y
    return y


def squared_normal(x, a: float, lamda: float = 1.0):
    """glass/grf/_transformations.py:170-175."""
    y = (x - a) ** 2 - 1
    if lamda != 1.0:
        y = lamda * y
    return y


def standard_normals(rng: np.random.Generator, z_size: int) -> np.ndarray:
    """glass/fields.py:407: standard_normal((z_size, 2)) @ [1, 1j]."""
    return rng.standard_normal((z_size, 2)) @ np.asarray([1, 1j])


def generate_grf_alm(gls, ncorr: int | None = None, rng=None):
    """glass/fields.py:334-426: yields the healpy-ordered alm of each shell
    (after the m = 0 fix) exactly as _generate_grf hands them to alm2map."""
    if rng is None:
        rng = np.random.default_rng(42)  # glass/rng.py:18, :94-102
    ngrf = nfields_from_nspectra(len(gls))
    if ncorr is None:
        ncorr = ngrf - 1
    n = max((np.asarray(gl).shape[0] for gl in gls if gl is not None), default=0)
    if n == 0:
        raise ValueError("all gls are empty")
    z_size = n * (n + 1) // 2
    y = deque()
    for w in iternorm(cls2cov(gls, n, ngrf, ncorr)):
        y.append(standard_normals(rng, z_size))
        while len(y) > w.shape[-1]:
            y.popleft()
        mis = w.shape[-1] - len(y)
        alm = sum(multalm(z, w[..., i + mis]) for i, z in enumerate(y))
        alm = glass_to_healpix_alm(alm)
        alm[:n] = alm[:n].real + alm[:n].imag + 0j
        yield alm


def generate(fields, gls, nside: int, ncorr: int | None = None, rng=None,
             nthreads: int | None = None):
    """glass/fields.py:839-894 with the FP64 oracle synthesis in place of healpy."""
    n = len(fields)
    if len(gls) != n * (n + 1) // 2:
        raise ValueError("mismatch between number of fields and gls")
    variances = (cltovar(getcl(gls, i, i)) for i in range(n))
    for t, alm, var in zip(fields, generate_grf_alm(gls, ncorr=ncorr, rng=rng),
                           variances, strict=True):
        x = alm2map(alm, nside, nthreads=nthreads)
        yield t(x, var)


def kappa_to_shear_factor(lmax: int, pw_ratio: np.ndarray | None = None) -> np.ndarray:
    """Net kappa -> gamma E-mode factor of glass/lensing.py:316-363 (and the
    identical single factor of :414-417): -sqrt((l-1)(l+2) / (l(l+1))), 0 at l < 2."""
    ell = np.arange(lmax + 1, dtype=np.float64)
    f = np.zeros(lmax + 1)
    pos = ell >= 2
    f[pos] = -np.sqrt((ell[pos] - 1) * (ell[pos] + 2) / (ell[pos] * (ell[pos] + 1)))
    if pw_ratio is not None:
        f *= np.asarray(pw_ratio)[: lmax + 1]
    return f


# --------------------------------------------------------------------------
# Counter-based generator of the device RNG (no counterpart in the reference,
# which draws from numpy's PCG64; see rng.py of the package).

_PHILOX_M0, _PHILOX_M1 = 0xD2511F53, 0xCD9E8D57
_PHILOX_W0, _PHILOX_W1 = 0x9E3779B9, 0xBB67AE85


def philox4x32_10(ctr: np.ndarray, key: np.ndarray) -> np.ndarray:
    """Philox4x32 with 10 rounds (Salmon, Moraes, Dror, Shaw, SC'11):
    ctr (..., 4) and key (..., 2) uint32 -> (..., 4) uint32."""
    c = [np.asarray(ctr[..., i], dtype=np.uint64) for i in range(4)]
    k0 = np.asarray(key[..., 0], dtype=np.uint64)
    k1 = np.asarray(key[..., 1], dtype=np.uint64)
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(_PHILOX_M0) * c[0]
        p1 = np.uint64(_PHILOX_M1) * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        c = [(hi1 ^ c[1] ^ k0) & mask, lo1, (hi0 ^ c[3] ^ k1) & mask, lo0]
        k0 = (k0 + np.uint64(_PHILOX_W0)) & mask
        k1 = (k1 + np.uint64(_PHILOX_W1)) & mask
    return np.stack([x.astype(np.uint32) for x in c], axis=-1)


def device_normals(seed: int, shell: int, g: np.ndarray) -> np.ndarray:
    """Complex standard normals of the device generator for GLASS indices g of
    one shell: one Philox call per (g, shell pair), counter (g lo, g hi,
    u lo, u hi) with u = shell // 2, key (seed lo, seed hi); words (0, 1) for an
    even shell, (2, 3) for an odd one; Box-Muller on u1 = (w_a + 1/2) 2^-32 and
    u2 = (w_b >> 8) 2^-24 (FP64 here; the device evaluates log / sqrt / sin /
    cos in FP32)."""
    g = np.asarray(g, dtype=np.uint64)
    u = shell >> 1
    ctr = np.stack([g & np.uint64(0xFFFFFFFF), g >> np.uint64(32),
                    np.full_like(g, u & 0xFFFFFFFF), np.full_like(g, u >> 32)], axis=-1)
    key = np.broadcast_to(np.array([seed & 0xFFFFFFFF, seed >> 32], dtype=np.uint64), ctr.shape[:-1] + (2,))
    w = philox4x32_10(ctr.astype(np.uint32), key.astype(np.uint32))
    a = 2 * (shell & 1)
    u1 = (np.float32(w[..., a]).astype(np.float64) + 0.5) * 2.0**-32
    u2 = (w[..., a + 1] >> 8).astype(np.float64) * 2.0**-24
    rad = np.sqrt(-2.0 * np.log(u1))
    return rad * np.cos(2 * np.pi * u2) + 1j * rad * np.sin(2 * np.pi * u2)


# --------------------------------------------------------------------------
# HEALPix pixel windows from their definition (restates what healpy.pixwin
# tabulates, glass/healpix.py:313-356; the tables are not available here).
#
# w^s_l = sqrt(mean_p (4 pi / (2l+1)) sum_m |<sY_lm>_p|^2), <.>_p the average
# over pixel p.  The pair form: (4 pi/(2l+1)) sum_m sY_lm(a) sY_lm(b)^* =
# d^l_ss(beta_ab) e^{i s psi_ab}, so the power is the pair average of
# d^l_ss(beta) cos(s psi).

# base-face offsets of the HEALPix grid (Gorski et al. 2005): jr = jrll - x - y
# in units of nside rings from the north pole, phi offset jpll in units of pi/4
_JRLL = (2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4)
_JPLL = (1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7)


def face_zphi(face: int, x, y):
    """(z, phi) of continuous face coordinates (x, y) in [0, 1]^2 (equal-area)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    jr = _JRLL[face] - x - y
    z = np.where(jr < 1, 1 - jr**2 / 3, np.where(jr > 3, (4 - jr) ** 2 / 3 - 1, (2 - jr) * 2 / 3))
    nr = np.where(jr < 1, jr, np.where(jr > 3, 4 - jr, 1.0))
    phi = (np.pi / 4) * (_JPLL[face] * nr + x - y) / np.where(nr > 0, nr, 1.0)
    return z, phi


def pixel_nodes(face: int, ix: int, iy: int, nside: int, G: int = 6):
    """Quadrature (z, phi, w) of one pixel, weights summing to 1: the pixel is
    split along its centre line x + y = const into two triangles, each
    integrated in collapsed coordinates toward its apex (G x G Gauss-Legendre)."""
    t, wt = np.polynomial.legendre.leggauss(G)
    u, wu = (t + 1) / 2, wt / 2
    h0 = 1.0 / nside
    sc, dc = (ix + iy + 1) * h0, (ix - iy) * h0
    zs, ps, ws = [], [], []
    for sgn in (1.0, -1.0):
        s = np.repeat(sc + sgn * h0 * u, G)
        h = np.repeat(h0 * (1 - u), G)
        d = dc + h * np.tile(t, G)
        w = np.repeat(wu * (1 - u), G) * np.tile(wt, G) / 2
        z, phi = face_zphi(face, (s + d) / 2, (s - d) / 2)
        zs.append(z)
        ps.append(phi)
        ws.append(w)
    return np.concatenate(zs), np.concatenate(ps), np.concatenate(ws)


def wigner_d(l: int, mp: int, m: int, beta) -> np.ndarray:
    """d^l_{mp, m}(beta) by Wigner's explicit sum (small l only)."""
    f = math.factorial
    c, s = np.cos(np.asarray(beta) / 2), np.sin(np.asarray(beta) / 2)
    out = np.zeros_like(c)
    pre = math.sqrt(f(l + mp) * f(l - mp) * f(l + m) * f(l - m))
    for k in range(max(0, m - mp), min(l + m, l - mp) + 1):
        den = f(l + m - k) * f(k) * f(mp - m + k) * f(l - mp - k)
        out = out + (-1) ** (mp - m + k) * pre / den * c ** (2 * l + m - mp - 2 * k) * s ** (mp - m + 2 * k)
    return out


def pixwin_definition(nside: int, lmax: int, spin: int = 0, G: int = 10) -> np.ndarray:
    """w^s_l from the definition, every pixel (small nside and lmax only)."""
    W = np.zeros(lmax + 1)
    for f in range(12):
        for ix in range(nside):
            for iy in range(nside):
                z, phi, w = pixel_nodes(f, ix, iy, nside, G)
                th = np.arccos(np.clip(z, -1, 1))
                for l in range(abs(spin), lmax + 1):
                    W[l] += sum(abs(np.sum(w * wigner_d(l, m, -spin, th) * np.exp(1j * m * phi))) ** 2
                                for m in range(-l, l + 1))
    return np.sqrt(W / (12 * nside * nside))


def _pair_power(z, phi, w, lmax: int, spin: int):
    st = np.sqrt(np.maximum(0.0, (1 - z) * (1 + z)))
    n = np.stack([st * np.cos(phi), st * np.sin(phi), z], -1)
    eth = np.stack([z * np.cos(phi), z * np.sin(phi), -st], -1)
    eph = np.stack([-np.sin(phi), np.cos(phi), np.zeros_like(phi)], -1)
    D = n[None, :, :] - n[:, None, :]  # [a, b] = n_b - n_a
    q = 0.5 * np.sum(D * D, axis=-1)   # 1 - cos beta
    beta = 2 * np.arcsin(np.minimum(1.0, np.sqrt(2 * q) / 2))
    x = np.cos(beta)
    t1 = D + q[..., None] * n[:, None, :]   # at a, toward b
    t2 = D - q[..., None] * n[None, :, :]   # at b, away from a
    c1 = np.einsum("abk,ak->ab", t1, eth)
    s1 = np.einsum("abk,ak->ab", t1, eph)
    c2 = np.einsum("abk,bk->ab", t2, eth)
    s2 = np.einsum("abk,bk->ab", t2, eph)
    re, im = c2 * c1 + s2 * s1, s2 * c1 - c2 * s1
    nn = re * re + im * im
    safe = np.where(nn > 0, nn, 1.0)
    if spin == 2:
        cs = np.where(nn > 0, (re * re - im * im) / safe, 1.0)
    elif spin == 1:
        cs = np.where(nn > 0, re / np.sqrt(safe), 1.0)
    else:
        cs = np.ones_like(x)
    wp = np.outer(w, w) * cs
    out = np.zeros(lmax + 1)
    s2_ = spin * spin
    if spin == 0:
        p0, p1 = np.ones_like(x), x
        out[0] = np.sum(wp)
        if lmax >= 1:
            out[1] = np.sum(wp * x)
        for l in range(2, lmax + 1):
            p0, p1 = p1, ((2 * l - 1) * x * p1 - (l - 1) * p0) / l
            out[l] = np.sum(wp * p1)
        return out
    d0, d1 = np.zeros_like(x), ((1 + x) / 2) ** spin
    if lmax >= spin:
        out[spin] = np.sum(wp * d1)
    for l in range(spin + 1, lmax + 1):
        lm = l - 1
        d0, d1 = d1, ((2 * l - 1) * (lm * (lm + 1) * x - s2_) * d1 - l * (lm * lm - s2_) * d0) / (lm * (l * l - s2_))
        out[l] = np.sum(wp * d1)
    return out


def pixwin_pairs(nside: int, lmax: int, spin: int = 0, G: int = 6, reduced: bool = True) -> np.ndarray:
    """w^s_l in the pair form, exact per pair.  reduced: only the pixels the
    device computes (face 0 with ix <= iy, ix + iy >= nside - 1, weight 8 or 16,
    and one pixel per belt ring nside + 1 .. 2 nside, weight 8 nside or 4 nside
    on the equator); otherwise every pixel."""
    W = np.zeros(lmax + 1)
    if reduced:
        items = [(0, ix, iy, 8.0 if ix == iy else 16.0)
                 for iy in range(nside) for ix in range(iy + 1) if ix + iy >= nside - 1]
        items += [(4, nside - 1, 2 * nside - r, 4.0 * nside if r == 2 * nside else 8.0 * nside)
                  for r in range(nside + 1, 2 * nside + 1)]
    else:
        items = [(f, ix, iy, 1.0) for f in range(12) for ix in range(nside) for iy in range(nside)]
    for f, ix, iy, mult in items:
        W += mult * _pair_power(*pixel_nodes(f, ix, iy, nside, G), lmax, spin)
    return np.sqrt(np.maximum(W / (12 * nside * nside), 0.0))


# --------------------------------------------------------------------------
# Multi-plane convergence (glass/lensing.py:430-673), restated with numpy
# arrays exactly as the reference computes it.


class MultiPlaneConvergence:
    """glass/lensing.py:430-586."""

    def __init__(self, cosmo) -> None:
        self.cosmo = cosmo
        self.z2 = 0.0
        self.z3 = 0.0
        self.x3 = 0.0
        self.w3 = 0.0
        self.r23 = 1.0
        self.delta3 = None
        self.kappa2 = None
        self.kappa3 = None

    def add_window(self, delta, w) -> None:
        """glass/lensing.py:488-508."""
        if self.delta3 is None:
            self.delta3 = np.asarray(0.0)
        zsrc = w.zeff
        lens_weight = float(np.trapezoid(w.wa, w.za) / np.interp(zsrc, w.za, w.wa))
        self.add_plane(delta, zsrc, lens_weight)

    def add_plane(self, delta, zsrc, wlens: float = 1.0) -> None:
        """glass/lensing.py:510-586."""
        if self.delta3 is None:
            self.delta3 = np.asarray(0.0)
        if zsrc <= self.z3:
            raise ValueError("source redshift must be increasing")
        delta2, self.delta3 = self.delta3, delta
        z1, self.z2, self.z3 = self.z2, self.z3, zsrc
        w2, self.w3 = self.w3, wlens
        x2, self.x3 = (self.x3,
                       self.cosmo.transverse_comoving_distance(self.z3) / self.cosmo.hubble_distance)
        r12 = self.r23
        r13, self.r23 = self.cosmo.transverse_comoving_distance(
            [z1, self.z2], self.z3) / (self.cosmo.hubble_distance * self.x3)
        t = r13 / r12
        f = 3 * self.cosmo.Omega_m0 / 2
        f *= x2 * self.r23
        f *= (1 + self.z2) / self.cosmo.H_over_H0(self.z2)
        f *= w2
        if self.kappa2 is None:
            self.kappa2 = np.zeros_like(delta)
            self.kappa3 = np.zeros_like(delta)
        self.kappa2, self.kappa3 = self.kappa3, self.kappa2
        t = np.asarray(t)
        self.kappa3 *= 1 - t
        self.kappa3 += t * self.kappa2
        self.kappa3 += np.asarray(f) * delta2

    @property
    def kappa(self):
        return self.kappa3


def multi_plane_matrix(shells, cosmo) -> np.ndarray:
    """glass/lensing.py:589-616."""
    mpc = MultiPlaneConvergence(cosmo)
    wmat = np.eye(len(shells))
    for i, w in enumerate(shells):
        mpc.add_window(wmat[i, :].copy(), w)
        wmat[i, :] = mpc.kappa
    return wmat


def multi_plane_weights(weights, shells, cosmo) -> np.ndarray:
    """glass/lensing.py:636-673."""
    shape = weights.shape
    if not shape or shape[0] != len(shells):
        raise ValueError("shape mismatch between weights and shells")
    weights = weights / np.sum(weights, axis=0)
    return np.matmul(multi_plane_matrix(shells, cosmo).T, weights)


# --------------------------------------------------------------------------
# Points from a density contrast (glass/points.py:115-540), and the device
# generator's counts and positions restated (no counterpart in the reference,
# whose positions come from the healpix package's randang).

ARCMIN2_SPHERE = 60**6 // 100 / math.pi


def expected_counts(delta, ngal, bias=None, vis=None, model="linear", remove_monopole=False):
    """glass/points.py:214-348 up to the Poisson draw: the clipped expected
    number of points per pixel, with the reference's array operations."""
    if bias is None:
        n = np.array(delta, dtype=np.float64, copy=True)
    elif model == "linear":
        n = bias * np.asarray(delta, dtype=np.float64)
    else:
        n = np.expm1(np.log1p(np.asarray(delta, dtype=np.float64)) * bias)
    if remove_monopole:
        n = n - np.mean(n, keepdims=True)
    n = n + 1
    n *= ARCMIN2_SPHERE / n.size * ngal
    if vis is not None:
        n *= vis
    return np.clip(n, 0.0, None)


def uniform53(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return ((a >> np.uint32(5)).astype(np.float64) * 67108864.0
            + (b >> np.uint32(6)).astype(np.float64)) * 2.0**-53


def _philox_words(lo, hi, w2, w3, seed):
    ctr = np.stack([lo, hi, w2, w3], axis=-1).astype(np.uint32)
    key = np.broadcast_to(np.array([seed & 0xFFFFFFFF, seed >> 32], dtype=np.uint32),
                          ctr.shape[:-1] + (2,))
    return philox4x32_10(ctr, key)


def device_poisson(lam: np.ndarray, seed: int, population: int) -> np.ndarray:
    """Counts of the device generator: lambda cut into parts of at most 32,
    each inverted on the CDF with one 53-bit uniform from Philox4x32-10 at
    counter (pixel lo, pixel hi, population, 0x40000000 + part)."""
    lam = np.asarray(lam, dtype=np.float64).copy()
    pix = np.arange(lam.size, dtype=np.uint64)
    out = np.zeros(lam.size, dtype=np.int64)
    part = 0
    while np.any(lam > 0):
        act = np.nonzero(lam > 0)[0]
        lpart = np.minimum(lam[act], 32.0)
        lam[act] = lam[act] - lpart
        w = _philox_words(pix[act] & np.uint64(0xFFFFFFFF), pix[act] >> np.uint64(32),
                          np.full(act.size, population, np.uint64),
                          np.full(act.size, 0x40000000 + part, np.uint64), seed)
        part += 1
        u = uniform53(w[:, 0], w[:, 1])
        p = np.exp(-lpart)
        cdf = p.copy()
        k = np.zeros(act.size, dtype=np.int64)
        run = u > cdf
        while np.any(run):
            k[run] += 1
            p[run] = p[run] * (lpart[run] / k[run])
            cdf[run] = cdf[run] + p[run]
            run = run & (u > cdf) & (k < 2048)
        out[act] += k
    return out


_JRLL_A = np.array(_JRLL)
_JPLL_A = np.array(_JPLL)


def ring_to_xyf(nside: int, pix) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """RING pixel -> (face, ix, iy) of the HEALPix face grid (Gorski et al. 2005)."""
    pix = np.asarray(pix, dtype=np.int64)
    N = int(nside)
    ncap, npix, nl2 = 2 * N * (N - 1), 12 * N * N, 2 * N
    face = np.zeros(pix.shape, np.int64)
    iring = np.zeros(pix.shape, np.int64)
    iphi = np.zeros(pix.shape, np.int64)
    kshift = np.zeros(pix.shape, np.int64)
    nr = np.zeros(pix.shape, np.int64)
    n = pix < ncap
    if np.any(n):
        p = pix[n]
        ir = (1 + np.floor(np.sqrt(1.0 + 2 * p)).astype(np.int64)) >> 1
        ir = np.where(2 * ir * (ir - 1) > p, ir - 1, ir)
        ir = np.where(2 * (ir + 1) * ir <= p, ir + 1, ir)
        ip = (p + 1) - 2 * ir * (ir - 1)
        iring[n], iphi[n], nr[n] = ir, ip, ir
        face[n] = (ip - 1) // ir
    e = (pix >= ncap) & (pix < npix - ncap)
    if np.any(e):
        ip0 = pix[e] - ncap
        tmp = ip0 // (4 * N)
        ir = tmp + N
        ip = ip0 - tmp * 4 * N + 1
        ks = (ir + N) & 1
        ire = ir - N + 1
        irm = nl2 + 2 - ire
        ifm = (ip - ire // 2 + N - 1) // N
        ifp = (ip - irm // 2 + N - 1) // N
        f = np.where(ifp == ifm, ifp | 4, np.where(ifp < ifm, ifp, ifm + 8))
        iring[e], iphi[e], kshift[e], nr[e], face[e] = ir, ip, ks, N, f
    s = pix >= npix - ncap
    if np.any(s):
        ip0 = npix - pix[s]
        ir = (1 + np.floor(np.sqrt(2.0 * ip0 - 1)).astype(np.int64)) >> 1
        ir = np.where(2 * ir * (ir - 1) >= ip0, ir - 1, ir)
        ir = np.where(2 * (ir + 1) * ir < ip0, ir + 1, ir)
        ip = 4 * ir + 1 - (ip0 - 2 * ir * (ir - 1))
        iring[s], iphi[s], nr[s] = 2 * nl2 - ir, ip, ir
        face[s] = 8 + (ip - 1) // ir
    irt = iring - _JRLL_A[face] * N + 1
    ipt = 2 * iphi - _JPLL_A[face] * nr - kshift - 1
    ipt = np.where(ipt >= nl2, ipt - 8 * N, ipt)
    return face, (ipt - irt) >> 1, (-ipt - irt) >> 1


def face_lonlat(face, x, y) -> tuple[np.ndarray, np.ndarray]:
    """(lon in [0, 360), lat) in degrees of face coordinates (x, y)."""
    face = np.asarray(face)
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    jr = _JRLL_A[face] - x - y
    polar_n, polar_s = jr < 1, jr > 3
    nr = np.where(polar_n, jr, np.where(polar_s, 4 - jr, 1.0))
    t = nr * nr / 3
    z = np.where(polar_n, 1 - t, np.where(polar_s, t - 1, (2 - jr) * (2 / 3)))
    st = np.where(polar_n | polar_s, np.sqrt(t * (2 - t)), np.sqrt((1 - z) * (1 + z)))
    q = np.where(nr > 0, (_JPLL_A[face] * nr + x - y) / np.where(nr > 0, nr, 1.0), 0.0)
    q = q - 8.0 * np.floor(q / 8.0)
    return 45.0 * q, np.degrees(np.arctan2(z, st))


def device_positions(nside: int, counts: np.ndarray, seed: int, population: int):
    """Positions of the device generator: point q of pixel p at face
    coordinates ((ix + u) / nside, (iy + v) / nside), u, v 53-bit uniforms from
    Philox4x32-10 at counter (q lo, q hi, population, 0x80000000)."""
    counts = np.asarray(counts, dtype=np.int64)
    ipix = np.repeat(np.arange(counts.size), counts)
    q = np.arange(ipix.size, dtype=np.uint64)
    w = _philox_words(q & np.uint64(0xFFFFFFFF), q >> np.uint64(32),
                      np.full(q.size, population, np.uint64),
                      np.full(q.size, 0x80000000, np.uint64), seed)
    u, v = uniform53(w[:, 0], w[:, 1]), uniform53(w[:, 2], w[:, 3])
    f, ix, iy = ring_to_xyf(nside, ipix)
    lon, lat = face_lonlat(f, (ix + u) / nside, (iy + v) / nside)
    return lon, lat, ipix


def ang2pix_ring(nside: int, theta, phi) -> np.ndarray:
    """RING pixel containing (theta, phi) (Gorski et al. 2005, the ring scheme)."""
    N = int(nside)
    z = np.cos(np.asarray(theta, dtype=np.float64))
    za = np.abs(z)
    tt = np.mod(np.asarray(phi, dtype=np.float64), 2 * np.pi) / (np.pi / 2)  # in [0, 4)
    out = np.zeros(z.shape, np.int64)
    eq = za <= 2 / 3
    if np.any(eq):
        t1 = N * (0.5 + tt[eq])
        t2 = N * z[eq] * 0.75
        jp = np.floor(t1 - t2).astype(np.int64)
        jm = np.floor(t1 + t2).astype(np.int64)
        ir = N + 1 + jp - jm
        ks = 1 - (ir & 1)
        ip = (jp + jm - N + ks + 1) // 2
        ip = np.mod(ip, 4 * N)
        out[eq] = 2 * N * (N - 1) + (ir - 1) * 4 * N + ip
    pc = ~eq
    if np.any(pc):
        ttp = tt[pc]
        tp = ttp - np.floor(ttp)
        tmp = N * np.sqrt(3 * (1 - za[pc]))
        jp = np.floor(tp * tmp).astype(np.int64)
        jm = np.floor((1 - tp) * tmp).astype(np.int64)
        ir = jp + jm + 1
        ip = np.floor(ttp * ir).astype(np.int64)
        ip = np.mod(ip, 4 * ir)
        out[pc] = np.where(z[pc] > 0, 2 * ir * (ir - 1) + ip, 12 * N * N - 2 * ir * (ir + 1) + ip)
    return out


# --------------------------------------------------------------------------
# C_l <-> C(theta) on Gauss-Legendre nodes (the transforms the device runs for
# the Gaussian-spectra solver, glass/grf/_solver.py:27-148, in place of the
# transformcl package): exact for l < n.


def _legendre_table(x: np.ndarray, n: int) -> np.ndarray:
    P = np.zeros((n, x.size))
    P[0] = 1.0
    if n > 1:
        P[1] = x
    for l in range(2, n):
        P[l] = ((2 * l - 1) * x * P[l - 1] - (l - 1) * P[l - 2]) / l
    return P


def gl_nodes(n: int) -> tuple[np.ndarray, np.ndarray]:
    """Gauss-Legendre nodes (ascending) and weights by Newton's method on P_n
    in extended precision (numpy's ``longdouble``), rounded to float64.
    numpy's ``leggauss`` evaluates the derivative at the unpolished root, so
    its weights next to +-1 carry relative errors ~1e-10 at n ~ 300; these
    carry the rounding of the last step only."""
    ld = np.longdouble
    i = np.arange((n + 1) // 2)
    z = np.cos(np.pi * (i + ld(0.75)) / (n + ld(0.5))).astype(ld)

    def pn(z):
        p0, p1 = np.ones_like(z), z.copy()
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
        return p1, n * (z * p1 - p0) / (z * z - 1)

    for _ in range(100):
        p, dp = pn(z)
        dz = p / dp
        z = z - dz
        if np.max(np.abs(dz)) < 1e-19:
            break
    _, dp = pn(z)
    w = 2 / ((1 - z * z) * dp * dp)
    x = np.concatenate([-z, z[::-1][n % 2:]]).astype(np.float64)
    wt = np.concatenate([w, w[::-1][n % 2:]]).astype(np.float64)
    if n % 2:
        x[n // 2] = 0.0
    return x, wt


def gl_cltocorr(cl) -> np.ndarray:
    cl = np.asarray(cl, dtype=np.float64)
    n = cl.size
    x, _ = gl_nodes(n)
    ell = np.arange(n)
    return ((2 * ell + 1) / (4 * np.pi) * cl) @ _legendre_table(x, n)


def gl_corrtocl(corr) -> np.ndarray:
    corr = np.asarray(corr, dtype=np.float64)
    n = corr.size
    x, w = gl_nodes(n)
    return 2 * np.pi * (_legendre_table(x, n) @ (w * corr))


# --------------------------------------------------------------------------
# Pixel quadrature weights (computed in place of healpy's weight files that
# map2alm(use_pixel_weights=True) reads, glass/healpix.py:232-276): the
# smallest correction of the uniform weights 4 pi / npix for which the pixel
# sum of every Y*_lm with l <= lmax is the integral sqrt(4 pi) delta_l0.


def ylm_matrix(nside: int, lmax: int) -> np.ndarray:
    """Y_lm (Condon-Shortley) at every pixel centre, [npix, nalm] in healpy order."""
    from scipy.special import sph_harm_y

    th, ph = pixel_angles(nside)
    Y = np.empty((th.size, nalm(lmax)), dtype=np.complex128)
    i = 0
    for m in range(lmax + 1):
        for l in range(m, lmax + 1):
            Y[:, i] = sph_harm_y(l, m, th, ph)
            i += 1
    return Y


def _constraint_rows(Y: np.ndarray, lmax: int) -> np.ndarray:
    """Real rows of the conditions: Y_l0 (real), then Re and Im of Y_lm, m > 0."""
    return np.concatenate([Y[:, : lmax + 1].real.T, Y[:, lmax + 1:].real.T,
                           Y[:, lmax + 1:].imag.T])


def _alm_dot(y: np.ndarray, z: np.ndarray, lmax: int) -> float:
    f = np.full(y.size, 2.0)
    f[: lmax + 1] = 1.0
    return float(np.sum(f * (y.real * z.real + y.imag * z.imag)))


def pixel_weights_cg(nside: int, lmax: int, tol: float = 1e-13, maxiter: int = 2000):
    """The same weights by conjugate gradients on C S lambda = b - C(w0) with
    this module's C transforms (C = (npix / 4 pi) map2alm without iterations,
    S = alm2map), for sizes where the dense solve is too large."""
    npix = nside2npix(nside)
    kappa = npix / (4 * np.pi)

    def C(f):
        c = kappa * map2alm(f, lmax, niter=0)
        c[: lmax + 1] = c[: lmax + 1].real
        return c

    b = np.zeros(nalm(lmax), dtype=np.complex128)
    b[0] = np.sqrt(4 * np.pi)
    r = b - C(np.full(npix, 4 * np.pi / npix))
    lam = np.zeros_like(r)
    p = r.copy()
    rr = rr0 = _alm_dot(r, r, lmax)
    for _ in range(maxiter):
        if rr <= tol * tol * rr0:
            break
        q = C(alm2map(p, nside, lmax))
        alpha = rr / _alm_dot(p, q, lmax)
        lam += alpha * p
        r -= alpha * q
        rn = _alm_dot(r, r, lmax)
        p = r + (rn / rr) * p
        rr = rn
    return 4 * np.pi / npix + alm2map(lam, nside, lmax)


def pixel_weights(nside: int, lmax: int) -> np.ndarray:
    Y = ylm_matrix(nside, lmax)
    A = _constraint_rows(Y, lmax)
    b = np.zeros(A.shape[0])
    b[0] = np.sqrt(4 * np.pi)
    w0 = np.full(A.shape[1], 4 * np.pi / A.shape[1])
    dw = np.linalg.lstsq(A, b - A @ w0, rcond=None)[0]
    return w0 + dw


def map2alm_weighted(m: np.ndarray, w: np.ndarray, lmax: int, niter: int = 3) -> np.ndarray:
    """a_lm = sum_p w_p f_p Y*_lm(p), then healpy's Jacobi passes
    a += A_w(f - S a); A_w(f) is the uniform-weight analysis of
    f w npix / (4 pi)."""
    nside = npix2nside(m.size)
    c = w * (m.size / (4 * np.pi))
    a = map2alm(m * c, lmax, niter=0)
    for _ in range(niter):
        a = a + map2alm((m - alm2map(a, nside, lmax)) * c, lmax, niter=0)
    return a
