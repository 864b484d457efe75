"""HEALPix transforms on MI355X, mirroring ``glass.healpix`` (glass/healpix.py).

The reference wraps healpy with ``numpy_fallback`` (glass/_array_api_utils.py:
569-652): arrays are converted to numpy, healpy runs on the host, results are
converted back to the caller's namespace.  Here the same functions run on the
GPU through the C ABI: numpy inputs come back as numpy arrays, torch tensors
come back as device tensors.  ``precision="fp64"`` (default) is the plain FP64
path; ``precision="fp32"`` runs the batched FP32-MFMA synthesis.
"""

from __future__ import annotations

import math
import os
import pathlib
import warnings
from collections.abc import Sequence

import numpy as np

from . import _lib

__all__ = [
    "alm2map", "alm2map_spin", "almxfl", "get_nside", "map2alm", "npix2nside",
    "nside2npix", "pixwin",
]


def _get_healpy_datapath() -> str | None:
    """glass/healpix.py:31-35."""
    p = os.environ.get("HEALPY_DATAPATH")
    if p is not None and not pathlib.Path(p).is_dir():
        raise ValueError(f"Healpy datapath not found at '{p}'")
    return p


def nside2npix(nside: int) -> int:
    """glass/healpix.py:296-310."""
    return 12 * int(nside) * int(nside)


def npix2nside(npix: int) -> int:
    """glass/healpix.py:279-293."""
    nside = math.isqrt(int(npix) // 12)
    if 12 * nside * nside != int(npix):
        raise ValueError(f"Wrong pixel number (it is not 12*nside**2): {npix}")
    return nside


def get_nside(m) -> int:
    """glass/healpix.py:214-229: nside of a map (last axis)."""
    return npix2nside(m.shape[-1])


def nalm(lmax: int) -> int:
    return (lmax + 1) * (lmax + 2) // 2


def lmax_from_nalm(n: int) -> int:
    lmax = (math.isqrt(8 * int(n) + 1) - 3) // 2
    if nalm(lmax) != int(n):
        raise ValueError(f"Wrong alm size (not a triangular number of lmax): {n}")
    return lmax


def _is_single(alms) -> bool:
    return getattr(alms, "ndim", None) == 1


def _as_list(alms):
    if hasattr(alms, "ndim"):
        return [alms[i] for i in range(alms.shape[0])] if alms.ndim == 2 else [alms]
    return list(alms)


def _synth_f64(alm_dev, nside: int, lmax: int):
    T = _lib.torch()
    plan = _lib.get_plan(nside, lmax)
    ws = _lib.workspace(plan.workspace_bytes(1, 8))
    out = T.empty(nside2npix(nside), dtype=T.float64, device="cuda")
    _lib.call("ga_alm2map_f64", plan.h, _lib.ptr(alm_dev), _lib.ptr(out), _lib.ptr(ws),
              ws.numel(), _lib.stream_ptr())
    return out


def synth_batch_f32(alm_batch, nside: int, lmax: int, epilogues=None, out=None):
    """alm_batch: complex64 device tensor [nalm, S] (shell innermost) -> [S, npix] float32."""
    T = _lib.torch()
    S = int(alm_batch.shape[1])
    plan = _lib.get_plan(nside, lmax)
    ws = _lib.workspace(plan.workspace_bytes(S, 4))
    if out is None:
        out = T.empty((S, nside2npix(nside)), dtype=T.float32, device="cuda")
    epi_ptr = _lib.ptr(epilogues) if epilogues is not None else None
    _lib.call("ga_alm2map_batch_f32", plan.h, S, _lib.ptr(alm_batch), _lib.ptr(out), epi_ptr,
              _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
    return out


def _check_alm_sizes(lst, lmax: int | None) -> int:
    """lmax of a list of 1-D alm (from the first if not given); ValueError
    unless every alm has exactly nalm(lmax) entries."""
    shapes = [tuple(getattr(a, "shape", np.shape(a))) for a in lst]
    if not shapes or any(len(sh) != 1 for sh in shapes):
        raise ValueError("alm must be one-dimensional")
    lm = lmax if lmax is not None else lmax_from_nalm(shapes[0][0])
    if any(sh[0] != nalm(lm) for sh in shapes):
        raise ValueError("alm size does not match lmax (mmax = lmax)")
    return int(lm)


def alm2map(alms, nside: int, *, inplace: bool = False, lmax: int | None = None,
            pixwin: bool = False, pol: bool = True, precision: str = "fp64"):
    """glass/healpix.py:38-78 (healpy.alm2map, RING output, mmax = lmax).

    A single 1-D alm gives one map.  A sequence of three with ``pol=True`` is
    read as (T, E, B) and gives [T, Q, U] (Q + iU = -sum (E + iB) 2Y_lm); any
    other sequence gives one map per alm.
    """
    del inplace  # outputs are always fresh device buffers; inputs are never modified
    single = _is_single(alms)
    lst = [alms] if single else _as_list(alms)
    if pixwin:
        # healpy smooths with its default-length windows (l < 3 nside; alm
        # beyond are zeroed by almxfl): pw0 on T, pw2 on E and B
        pw0, pw2 = (a[: 3 * int(nside)] for a in _window_tables(int(nside)))
        polarised = pol and len(lst) == 3 and not single
        lst = [almxfl(a, pw2 if (polarised and i > 0) else pw0) for i, a in enumerate(lst)]
        if single:
            alms = lst[0]
    if pol and len(lst) == 3 and not single:
        t = alm2map(lst[0], nside, lmax=lmax, pol=False, precision=precision)
        q, u = alm2map_spin(lst[1:], nside, 2, lmax if lmax is not None else
                            lmax_from_nalm(lst[1].shape[-1]), precision=precision)
        return [t, q, u]
    if precision not in ("fp32", "fp64"):
        raise ValueError(f"unknown precision {precision!r}")
    # every alm must hold exactly nalm(lmax) coefficients (mmax = lmax): the
    # kernels read that many; checked on the host, before anything is launched
    lm = _check_alm_sizes(lst, lmax)
    T = _lib.torch()
    outs = []
    if precision == "fp32":
        devs = []
        kind = "numpy"
        for a in lst:
            d, kind = _lib.to_device(a, T.complex128)
            devs.append(d)
        batch = T.stack(devs, dim=1).to(T.complex64).contiguous()
        maps = synth_batch_f32(batch, nside, lm)
        outs = [_lib.from_device(maps[i], kind) for i in range(len(lst))]
    else:
        for a in lst:
            d, kind = _lib.to_device(a, T.complex128)
            outs.append(_lib.from_device(_synth_f64(d, nside, lm), kind))
    return outs[0] if single else outs


_pixel_weight_cache: dict = {}


def pixel_weights(nside: int, *, tol: float = 1e-12, maxiter: int = 1000):
    """Per-pixel quadrature weights of ``nside`` (float64 device tensor of
    npix), computed on the GPU (``ga_pixel_weights``) in place of healpy's
    weight files: the smallest correction of the uniform weights 4 pi / npix
    for which the pixel sum of Y*_lm equals its integral for every
    l <= 2 nside, so the weighted analysis of a map band-limited to nside is
    exact.  Computed once per (nside, device)."""
    T = _lib.torch()
    _lib.require_gpu()
    nside = int(nside)
    key = (nside, T.cuda.current_device())
    if key not in _pixel_weight_cache:
        plan = _lib.get_plan(nside, 2 * nside)
        n = _lib.ctypes.c_size_t()
        _lib.call("ga_pixel_weights_workspace_size", plan.h, _lib.ctypes.byref(n))
        ws = _lib.workspace(n.value)
        w = T.empty(nside2npix(nside), dtype=T.float64, device="cuda")
        it = _lib.ctypes.c_int32()
        res = _lib.ctypes.c_double()
        _lib.call("ga_pixel_weights", plan.h, float(tol), int(maxiter), _lib.ptr(w),
                  _lib.ctypes.byref(it), _lib.ctypes.byref(res), _lib.ptr(ws), ws.numel(),
                  _lib.stream_ptr())
        if res.value > 100 * tol:
            warnings.warn(f"pixel weights for nside {nside}: relative residual {res.value:.2e} "
                          f"after {it.value} iterations", stacklevel=2)
        _pixel_weight_cache[key] = w
    return _pixel_weight_cache[key]


def map2alm(maps, *, lmax: int | None = None, pol: bool = True,
            use_pixel_weights: bool = False, iter: int = 3,  # noqa: A002
            precision: str = "fp64"):
    """glass/healpix.py:232-276 (healpy.map2alm; RING input, mmax = lmax,
    default lmax = 3 nside - 1, ``iter`` Jacobi refinement passes).

    Weights: uniform 4 pi / npix per pixel, or with ``use_pixel_weights`` the
    per-pixel quadrature weights of ``pixel_weights`` (computed on the GPU,
    where healpy reads them from data files; the datapath is still checked
    exactly like the reference).  The weighted analysis runs in FP64 (also for
    ``precision="fp32"``, whose result is then rounded to complex64).
    """
    if use_pixel_weights:
        _get_healpy_datapath()
    T = _lib.torch()
    single = getattr(maps, "ndim", None) == 1
    if not single and pol and len(_as_list(maps)) == 3:
        raise NotImplementedError("polarised (TQU) analysis is not implemented")
    lst = [maps] if single else _as_list(maps)
    nside = npix2nside(lst[0].shape[-1])
    if lmax is None:
        lmax = 3 * nside - 1
    plan = _lib.get_plan(nside, lmax)
    outs = []
    if precision not in ("fp64", "fp32"):
        raise ValueError(f"unknown precision {precision!r}")
    if precision == "fp64" or use_pixel_weights:
        ws = _lib.workspace(plan.workspace_bytes(1, 8, analysis=True))
        w = pixel_weights(nside) if use_pixel_weights else None
        for mp in lst:
            d, kind = _lib.to_device(mp, T.float64)
            alm = T.empty(nalm(lmax), dtype=T.complex128, device="cuda")
            tmp = T.empty_like(d) if (iter > 0 or w is not None) else None
            if w is not None:
                _lib.call("ga_map2alm_weighted_f64", plan.h, _lib.ptr(d), _lib.ptr(w),
                          _lib.ptr(alm), int(iter), _lib.ptr(tmp), _lib.ptr(ws), ws.numel(),
                          _lib.stream_ptr())
            else:
                _lib.call("ga_map2alm_f64", plan.h, _lib.ptr(d), _lib.ptr(alm), int(iter),
                          _lib.ptr(tmp) if tmp is not None else None, _lib.ptr(ws), ws.numel(),
                          _lib.stream_ptr())
            if precision == "fp32":
                alm = alm.to(T.complex64)
            outs.append(_lib.from_device(alm, kind))
    else:
        devs = []
        kind = "numpy"
        for mp in lst:
            d, kind = _lib.to_device(mp, T.float32)
            devs.append(d)
        batch = T.stack(devs, dim=0).contiguous()
        alm = map2alm_batch_f32(batch, lmax, iter)
        outs = [_lib.from_device(alm[:, i].contiguous(), kind) for i in range(len(lst))]
    return outs[0] if single else outs


def map2alm_batch_f32(maps, lmax: int, niter: int = 3):
    """[S, npix] float32 device maps -> [nalm, S] complex64 alm."""
    T = _lib.torch()
    S = int(maps.shape[0])
    nside = npix2nside(maps.shape[-1])
    plan = _lib.get_plan(nside, lmax)
    ws = _lib.workspace(plan.workspace_bytes(S, 4, analysis=True))
    alm = T.empty((nalm(lmax), S), dtype=T.complex64, device="cuda")
    tmp = T.empty_like(maps) if niter > 0 else None
    _lib.call("ga_map2alm_batch_f32", plan.h, S, _lib.ptr(maps), _lib.ptr(alm), int(niter),
              _lib.ptr(tmp) if tmp is not None else None, _lib.ptr(ws), ws.numel(),
              _lib.stream_ptr())
    return alm


def alm2map_spin(alms, nside: int, spin: int, lmax: int, precision: str = "fp64"):
    """glass/healpix.py:81-108: [E, B] -> [Q, U], Q + iU = -sum (E + iB) sY_lm."""
    T = _lib.torch()
    if int(spin) not in (1, 2):
        raise NotImplementedError(f"spin {spin} is not supported (1 or 2)")
    if len(alms) != 2:
        raise ValueError("alm2map_spin takes two alm, [E, B]")
    _check_alm_sizes(list(alms), int(lmax))  # E and B are both read as nalm(lmax) values
    e, kind = _lib.to_device(alms[0], T.complex128)
    b, _ = _lib.to_device(alms[1], T.complex128)
    plan = _lib.get_plan(nside, lmax)
    plan.prepare_spin(int(spin))
    npix = nside2npix(nside)
    if precision == "fp64":
        ws = _lib.workspace(plan.workspace_bytes(2, 8))
        out = T.empty((2, npix), dtype=T.float64, device="cuda")
        _lib.call("ga_alm2map_spin_f64", plan.h, int(spin), _lib.ptr(e), _lib.ptr(b),
                  _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
    elif precision == "fp32":
        out = spin_batch_f32(e.to(T.complex64)[:, None].contiguous(),
                             b.to(T.complex64)[:, None].contiguous(), nside, spin, lmax)[0]
    else:
        raise ValueError(f"unknown precision {precision!r}")
    return [_lib.from_device(out[0], kind), _lib.from_device(out[1], kind)]


def spin_batch_f32(E, B, nside: int, spin: int, lmax: int):
    """E, B: [nalm, S] complex64 device tensors -> [S, 2, npix] float32 (Q, U);
    B = None means B = 0."""
    T = _lib.torch()
    S = int(E.shape[1])
    plan = _lib.get_plan(nside, lmax)
    plan.prepare_spin(int(spin))
    ws = _lib.workspace(plan.workspace_bytes(2 * S, 4))
    out = T.empty((S, 2, nside2npix(nside)), dtype=T.float32, device="cuda")
    _lib.call("ga_alm2map_spin_batch_f32", plan.h, int(spin), S, _lib.ptr(E),
              None if B is None else _lib.ptr(B), _lib.ptr(out), _lib.ptr(ws), ws.numel(),
              _lib.stream_ptr())
    return out


def almxfl(alm, fl, *, inplace: bool = False):
    """glass/healpix.py:111-140: alm_lm * fl[l], zero where fl is not defined."""
    T = _lib.torch()
    is_t = _lib.is_torch(alm)
    dtype = alm.dtype if is_t else np.asarray(alm).dtype
    cdt = T.complex64 if dtype in (T.complex64, np.complex64) else T.complex128
    d, kind = _lib.to_device(alm, cdt)
    f, _ = _lib.to_device(fl, T.float64)
    lmax = lmax_from_nalm(d.shape[-1])
    out = d if (inplace and is_t and d.data_ptr() == alm.data_ptr()) else T.empty_like(d)
    _lib.call("ga_almxfl", lmax, 1, 4 if cdt == T.complex64 else 8, _lib.ptr(d), _lib.ptr(f),
              int(f.numel()), _lib.ptr(out), _lib.stream_ptr())
    res = _lib.from_device(out, kind)
    if inplace and kind == "numpy":
        np.copyto(alm, res)
        return alm
    return res


_pixwin_cache: dict[tuple[int, int], tuple[np.ndarray, np.ndarray]] = {}


def _window_tables(nside: int) -> tuple[np.ndarray, np.ndarray]:
    """(pw0, pw2) for l = 0 .. 4 nside, computed once per (nside, device)."""
    T = _lib.torch()
    _lib.require_gpu()
    key = (nside, T.cuda.current_device())
    if key not in _pixwin_cache:
        lfull = 4 * nside
        n = _lib.ctypes.c_size_t()
        _lib.call("ga_pixwin_workspace_size", nside, lfull, _lib.ctypes.byref(n))
        ws = _lib.workspace(n.value)
        pw0 = T.empty(lfull + 1, dtype=T.float64, device="cuda")
        pw2 = T.empty(lfull + 1, dtype=T.float64, device="cuda")
        _lib.call("ga_pixwin", nside, lfull, 2, _lib.ptr(pw0), _lib.ptr(pw2), _lib.ptr(ws),
                  ws.numel(), _lib.stream_ptr())
        _pixwin_cache[key] = (pw0.cpu().numpy(), pw2.cpu().numpy())
    return _pixwin_cache[key]


def pixwin(nside: int, *, lmax: int | None = None, pol: bool = False, xp=None):
    """glass/healpix.py:313-356: the pixel window function(s) of ``nside``.

    healpy reads these from tabulated package data; here they are computed on
    the GPU from their definition (``ga_pixwin``: the harmonic power of each
    pixel's normalised top hat averaged over the pixels; the polarisation
    window from the spin-2 harmonics), once per nside up to l = 4 nside, and
    sliced.  ``lmax`` defaults to 3 nside - 1 and may not exceed 4 nside (the
    range of the tabulated windows).  Returns float64 arrays of ``xp``
    (numpy by default); with ``pol`` the pair (temperature, polarisation).
    A HEALPY_DATAPATH that names no directory raises as in the reference,
    although no data file is read."""
    _get_healpy_datapath()
    nside = int(nside)
    if nside < 1:
        raise ValueError(f"invalid nside {nside}")
    if lmax is None:
        lmax = 3 * nside - 1
    lmax = int(lmax)
    if lmax < 0 or lmax > 4 * nside:
        raise ValueError(f"lmax {lmax} outside 0 .. 4 nside = {4 * nside}")
    pw0, pw2 = (a[: lmax + 1].copy() for a in _window_tables(nside))
    if xp is not None:
        pw0, pw2 = xp.asarray(pw0, dtype=xp.float64), xp.asarray(pw2, dtype=xp.float64)
    return (pw0, pw2) if pol else pw0


def sequence_of_arrays(x) -> bool:
    return isinstance(x, Sequence) and not isinstance(x, (str, bytes))
