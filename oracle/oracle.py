"""FP64 CPU oracle for the GLASS field-generation path.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, by ``__graft_entry__.smoke``
and by the CPU-baseline leg of ``bench.py`` -- always as the checker, never as
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
                    (Salmon et al., SC'11, Random123 distribution).  The device
                    generator of this build (not the reference's numpy stream)
                    is then checked against device_normals built on it.

Parity unpinned: healpy's pixel-weight files (use_pixel_weights=True) and pixwin
tables are not available, so those inputs are taken from the caller.
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
    one shell: counter (g lo, g hi, shell lo, shell hi), key (seed lo, seed hi),
    Box-Muller on u1 = (w0 + 1/2) 2^-32 and u2 = (w1 >> 8) 2^-24 (FP64 here; the
    device evaluates log / sqrt / sin / cos in FP32)."""
    g = np.asarray(g, dtype=np.uint64)
    ctr = np.stack([g & np.uint64(0xFFFFFFFF), g >> np.uint64(32),
                    np.full_like(g, shell & 0xFFFFFFFF), np.full_like(g, shell >> 32)], axis=-1)
    key = np.broadcast_to(np.array([seed & 0xFFFFFFFF, seed >> 32], dtype=np.uint64), ctr.shape[:-1] + (2,))
    w = philox4x32_10(ctr.astype(np.uint32), key.astype(np.uint32))
    u1 = (np.float32(w[..., 0]).astype(np.float64) + 0.5) * 2.0**-32
    u2 = (w[..., 1] >> 8).astype(np.float64) * 2.0**-24
    rad = np.sqrt(-2.0 * np.log(u1))
    return rad * np.cos(2 * np.pi * u2) + 1j * rad * np.sin(2 * np.pi * u2)
