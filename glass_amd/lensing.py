"""Convergence -> shear on MI355X, mirroring glass/lensing.py:185-428.

The chain is map2alm (analysis) -> almxfl by the l-dependent factors ->
alm2map / alm2map_spin.  The factors are the reference's:
    psi   = -2 / (l (l+1)) kappa            (l > 0, 0 at l = 0)
    alpha = sqrt(l (l+1)) psi               -> spin-1 synthesis
    gamma = sqrt((l-1)(l+2)) / 2 alpha      (l > 0) [x pw2/pw0 if discretized]
so kappa -> gamma is -sqrt((l-1)(l+2) / (l(l+1))) for l >= 2 and 0 below.
With ``discretized=True`` (the default, as in the reference) the factor is
multiplied by pw2 / pw0 of ``healpix.pixwin(nside, lmax=lmax, pol=True)``,
which glass_amd computes on the GPU from the pixel-window definition; an
explicit ratio can be passed as ``pixwin_ratio`` instead.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import _lib
from . import healpix as hp


def _ell(lmax: int) -> np.ndarray:
    return np.arange(lmax + 1, dtype=np.float64)


def _pw_ratio(nside: int, lmax: int, discretized: bool, pixwin_ratio):
    """pw2 / pw0 up to lmax (glass/lensing.py:353-362, :414-419), or None."""
    if not discretized:
        return None
    if pixwin_ratio is None:
        pw0, pw2 = hp.pixwin(nside, lmax=lmax, pol=True)
        return pw2 / pw0
    r = np.asarray(pixwin_ratio, dtype=np.float64)
    if r.shape[0] < lmax + 1:
        raise ValueError("pixwin_ratio is shorter than lmax + 1")
    return r[: lmax + 1]


def from_convergence(kappa, lmax: int | None = None, *, potential: bool = False,
                     deflection: bool = False, shear: bool = False, discretized: bool = True,
                     pixwin_ratio=None, precision: str = "fp64"):
    """glass/lensing.py:185-371."""
    if not (potential or deflection or shear):
        return ()
    nside = hp.get_nside(kappa)
    if lmax is None:
        lmax = 3 * nside - 1
    alm = hp.map2alm(kappa, lmax=lmax, pol=False, use_pixel_weights=True, precision=precision)
    ell = _ell(lmax)
    results = ()
    fl = np.zeros(lmax + 1)
    fl[1:] = -2.0 / (ell[1:] * (ell[1:] + 1))
    alm = hp.almxfl(alm, fl)
    if potential:
        results += (hp.alm2map(alm, nside, lmax=lmax, pol=False, precision=precision),)
    if not (deflection or shear):
        return results
    zeros = _zeros_like(alm)
    alm = hp.almxfl(alm, np.sqrt(ell * (ell + 1)))
    if deflection:
        q, u = hp.alm2map_spin([alm, zeros], nside, 1, lmax, precision=precision)
        results += (q + 1j * u,)
    if not shear:
        return results
    fl = np.zeros(lmax + 1)
    fl[1:] = np.sqrt((ell[1:] - 1) * (ell[1:] + 2)) / 2
    pw = _pw_ratio(nside, lmax, discretized, pixwin_ratio)
    if pw is not None:
        fl *= pw
    alm = hp.almxfl(alm, fl)
    q, u = hp.alm2map_spin([alm, zeros], nside, 2, lmax, precision=precision)
    results += (q + 1j * u,)
    return results


def shear_from_convergence(kappa, lmax: int | None = None, *, discretized: bool = True,
                           pixwin_ratio=None, precision: str = "fp64"):
    """glass/lensing.py:374-428 (deprecated in the reference): returns [g1, g2]."""
    nside = hp.get_nside(kappa)
    if lmax is None:
        lmax = 3 * nside - 1
    alm = hp.map2alm(kappa, lmax=lmax, pol=False, use_pixel_weights=True, precision=precision)
    ell = _ell(lmax)
    fl = np.sqrt((ell + 2) * (ell + 1) * ell * np.clip(ell - 1, 0, None))
    fl /= np.clip(ell * (ell + 1), 1, None)
    fl *= -1
    pw = _pw_ratio(nside, lmax, discretized, pixwin_ratio)
    if pw is not None:
        fl *= pw
    alm = hp.almxfl(alm, fl)
    return hp.alm2map_spin([alm, _zeros_like(alm)], nside, 2, lmax, precision=precision)


def _zeros_like(a):
    if _lib.is_torch(a):
        return _lib.torch().zeros_like(a)
    return np.zeros_like(a)


# ---------------------------------------------------------------------------
# Multi-plane convergence (glass/lensing.py:430-673)


@dataclasses.dataclass(frozen=True)
class RadialWindow:
    """glass/shells.py:184-230: a window function in redshift (abscissae
    ``za``, values ``wa``) with an effective redshift ``zeff``.  Any object
    with these three attributes (the reference's own RadialWindow included)
    is accepted wherever a window is expected."""

    za: np.ndarray
    wa: np.ndarray
    zeff: float = math.nan


class MultiPlaneConvergence:
    """glass/lensing.py:430-586: convergence planes built iteratively from
    mass planes.  The cosmology (distances, E(z), Omega_m0) is evaluated on
    the host exactly as the reference does; the plane update
    kappa3 = (1 - t) kappa1 + t kappa2 + f delta2 runs on the device
    (``ga_multi_plane_update``, the reference's three float64 steps with the
    same roundings).  The planes stay in device memory: a torch tensor in
    gives torch tensors out (no host round trip); numpy in gives numpy from
    ``kappa`` / ``delta``.  Mass planes passed as numpy are copied to the
    device when added."""

    def __init__(self, cosmo) -> None:
        self.cosmo = cosmo
        self.z2 = 0.0
        self.z3 = 0.0
        self.x3 = 0.0
        self.w3 = 0.0
        self.r23 = 1.0
        self.delta3 = None   # device tensor, or None before the first plane (= 0)
        self.kappa2 = None
        self.kappa3 = None
        self._kind = None

    def add_window(self, delta, w) -> None:
        """glass/lensing.py:488-508: lensing weight from the window function,
        source plane at its effective redshift."""
        zsrc = w.zeff
        za = np.asarray(w.za, dtype=np.float64)
        wa = np.asarray(w.wa, dtype=np.float64)
        lens_weight = float(np.trapezoid(wa, za) / np.interp(zsrc, za, wa))
        self.add_plane(delta, zsrc, lens_weight)

    def add_plane(self, delta, zsrc, wlens: float = 1.0) -> None:
        """glass/lensing.py:510-586."""
        if zsrc <= self.z3:
            raise ValueError("source redshift must be increasing")
        T = _lib.torch()
        if _lib.is_torch(delta):
            d = delta if delta.is_cuda else delta.to("cuda")
            kind = "torch"
        else:
            arr = np.asarray(delta)
            dt = T.float32 if arr.dtype == np.float32 else T.float64
            d, kind = _lib.to_device(arr, dt)
        if d.dtype not in (T.float32, T.float64):
            raise ValueError("mass planes must be float32 or float64")
        d = d.contiguous()
        self._kind = kind

        # cycle mass plane, redshifts of source planes and weights of mass plane
        delta2, self.delta3 = self.delta3, d
        z1, self.z2, self.z3 = self.z2, self.z3, zsrc
        w2, self.w3 = self.w3, wlens

        # extrapolation law
        x2, self.x3 = (self.x3,
                       self.cosmo.transverse_comoving_distance(self.z3) / self.cosmo.hubble_distance)
        r12 = self.r23
        r13, self.r23 = self.cosmo.transverse_comoving_distance(
            [z1, self.z2], self.z3) / (self.cosmo.hubble_distance * self.x3)
        t = r13 / r12

        # lensing weight of the mass plane to be added
        f = 3 * self.cosmo.Omega_m0 / 2
        f *= x2 * self.r23
        f *= (1 + self.z2) / self.cosmo.H_over_H0(self.z2)
        f *= w2

        if self.kappa2 is None:
            self.kappa2 = T.zeros_like(d)
            self.kappa3 = T.zeros_like(d)
        if self.kappa2.shape != d.shape or self.kappa2.dtype != d.dtype:
            raise ValueError("mass planes must all have the same shape and dtype")
        if delta2 is not None and (delta2.shape != d.shape or delta2.dtype != d.dtype):
            raise ValueError("mass planes must all have the same shape and dtype")

        # kappa1, kappa2 = kappa2, kappa3; the new kappa3 overwrites kappa1
        self.kappa2, self.kappa3 = self.kappa3, self.kappa2
        rb = 4 if d.dtype == T.float32 else 8
        _lib.call("ga_multi_plane_update", int(d.numel()), rb, _lib.ptr(self.kappa3),
                  _lib.ptr(self.kappa2), None if delta2 is None else _lib.ptr(delta2),
                  float(t), float(f), _lib.ptr(self.kappa3), _lib.stream_ptr())

    def _out(self, x):
        if x is None or self._kind == "torch":
            return x
        return x.cpu().numpy()

    @property
    def zsrc(self):
        """The redshift of the current convergence plane."""
        return self.z3

    @property
    def kappa(self):
        """The current convergence plane."""
        return self._out(self.kappa3)

    @property
    def delta(self):
        """The current matter plane."""
        return self._out(self.delta3)

    @property
    def wlens(self) -> float:
        """The weight of the current matter plane."""
        return self.w3


def multi_plane_matrix(shells, cosmo) -> np.ndarray:
    """glass/lensing.py:589-616: the matrix of lensing contributions of each
    shell, row i the convergence of source plane i per unit mass plane."""
    n = len(shells)
    mpc = MultiPlaneConvergence(cosmo)
    wmat = np.eye(n)
    for i, w in enumerate(shells):
        mpc.add_window(wmat[i, :].copy(), w)
        wmat[i, :] = mpc.kappa
    return wmat


def multi_plane_weights(weights, shells, cosmo) -> np.ndarray:
    """glass/lensing.py:636-673: relative lensing weights of each shell."""
    weights = np.asarray(weights)
    shape = weights.shape
    if not shape or shape[0] != len(shells):
        raise ValueError("shape mismatch between weights and shells")
    weights = weights / np.sum(weights, axis=0)
    mat = multi_plane_matrix(shells, cosmo)
    return np.matmul(mat.T, weights)
