"""Convergence -> shear on MI355X, mirroring glass/lensing.py:185-428.

The chain is map2alm (analysis) -> almxfl by the l-dependent factors ->
alm2map / alm2map_spin.  The factors are the reference's:
    psi   = -2 / (l (l+1)) kappa            (l > 0, 0 at l = 0)
    alpha = sqrt(l (l+1)) psi               -> spin-1 synthesis
    gamma = sqrt((l-1)(l+2)) / 2 alpha      (l > 0) [x pw2/pw0 if discretized]
so kappa -> gamma is -sqrt((l-1)(l+2) / (l(l+1))) for l >= 2 and 0 below.
With ``discretized=True`` the reference reads healpy's pixel-window tables;
those are not available here, so the ratio pw2/pw0 must be supplied through
``pixwin_ratio`` (an array over l) or ``discretized=False`` must be used.
"""

from __future__ import annotations

import numpy as np

from . import _lib
from . import healpix as hp


def _ell(lmax: int) -> np.ndarray:
    return np.arange(lmax + 1, dtype=np.float64)


def _pw_ratio(nside: int, lmax: int, discretized: bool, pixwin_ratio):
    if not discretized:
        return None
    if pixwin_ratio is None:
        hp._get_healpy_datapath()
        raise NotImplementedError(
            "discretized=True needs the HEALPix pixel-window ratio pw2/pw0; pass it as "
            "pixwin_ratio=... or use discretized=False")
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
    alm = hp.map2alm(kappa, lmax=lmax, pol=False, use_pixel_weights=False, precision=precision)
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
    alm = hp.map2alm(kappa, lmax=lmax, pol=False, use_pixel_weights=False, precision=precision)
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
