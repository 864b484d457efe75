"""Positions tracing a density contrast, on the device
(glass/points.py:115-160, :163-348, :351-540).

``positions_from_delta`` keeps the reference's signature, argument meaning,
broadcasting of populations and batched yields.  Per population the
expected count per pixel, lambda = max(0, (b(delta) [- mean] + 1) *
ARCMIN2_SPHERE / npix * ngal [* vis]), is formed by a HIP kernel with the
reference's roundings (``ga_points_counts``); the Poisson counts come from

* the numpy generator when ``rng`` is None (seed 42, glass/rng.py:94-102) or
  a numpy ``Generator``: ``rng.poisson(lambda)`` on the host, the reference's
  exact count stream (lambda is copied to the host for it); or
* the counter-based device generator when ``rng`` is a ``DeviceRNG``:
  inversion of the Poisson CDF on Philox4x32-10 uniforms in the same kernel,
  no host work;

and every point is placed uniformly (in area) inside its pixel on the device
(``ga_points_positions``, Philox uniforms in the pixel's equal-area face
coordinates).  The reference places points with the healpix package's
``randang`` from a fresh seed-42 generator, which is not available here, so
positions are statistically, not bitwise, equivalent.

Points are yielded in batches of at most ``batch`` (consecutive points of the
population, not cut at pixel boundaries), with the count of the batch (an
array over the population dimensions, one-hot, when there are several).
numpy inputs give numpy outputs; torch inputs give device tensors.
"""

from __future__ import annotations

import itertools
import math

import numpy as np

from . import _lib
from .rng import SEED, DeviceRNG

ARCMIN2_SPHERE = 60**6 // 100 / math.pi

_MODEL_IDENTITY, _MODEL_LINEAR, _MODEL_LOGLINEAR = 0, 1, 2


def linear_bias(delta, b):
    """glass/points.py:115-134: delta_g = b delta."""
    return b * delta


def loglinear_bias(delta, b):
    """glass/points.py:137-160: ln(1 + delta_g) = b ln(1 + delta)."""
    if _lib.is_torch(delta):
        T = _lib.torch()
        return T.expm1(T.log1p(delta) * b)
    return np.expm1(np.log1p(delta) * b)


def _leading(x, core: int):
    if x is None:
        return None
    if _lib.is_torch(x):
        return tuple(x.shape[: x.dim() - core])
    return np.shape(x)[: np.ndim(x) - core]


def _at(x, k, core: int):
    """x[(*k, ...)] after broadcasting x's leading axes to the population shape."""
    if x is None:
        return None
    shape = _leading(x, core)
    idx = tuple(ki if si != 1 else 0 for ki, si in zip(k[len(k) - len(shape):], shape))
    return x[idx] if idx else x


def positions_from_delta(ngal, delta, bias=None, vis=None, *, bias_model=linear_bias,
                         remove_monopole: bool = False, batch: int = 1_000_000, rng=None):
    """glass/points.py:443-540 (see the module docstring)."""
    if not callable(bias_model):
        raise TypeError("bias_model must be callable")
    T = _lib.torch()
    _lib.require_gpu()
    torch_out = _lib.is_torch(delta)
    ngal_a = ngal if _lib.is_torch(ngal) else np.asarray(ngal, dtype=np.float64)
    bias_a = None if bias is None else (bias if _lib.is_torch(bias)
                                        else np.asarray(bias, dtype=np.float64))
    shapes = [_leading(ngal_a, 0), _leading(delta, 1)]
    if bias_a is not None:
        shapes.append(_leading(bias_a, 0))
    if vis is not None:
        shapes.append(_leading(vis, 1))
    dims = tuple(np.broadcast_shapes(*shapes))
    npix = int(delta.shape[-1])
    nside = math.isqrt(npix // 12)
    if 12 * nside * nside != npix:
        raise ValueError(f"invalid map size {npix}")

    device_rng = isinstance(rng, DeviceRNG)
    seed = rng.seed if device_rng else SEED
    host_rng = None if device_rng else (np.random.default_rng(SEED) if rng is None else rng)
    nbytes = _lib.ctypes.c_size_t()
    _lib.call("ga_points_workspace_size", npix, _lib.ctypes.byref(nbytes))
    ws = _lib.workspace(nbytes.value)
    total = T.zeros(1, dtype=T.int64, device="cuda")
    counts = T.empty(npix, dtype=T.int64, device="cuda")

    for pop, k in enumerate(itertools.product(*map(range, dims))):
        d = _at(delta, k, 1)
        b = _at(bias_a, k, 0)
        g = float(_at(ngal_a, k, 0))
        v = _at(vis, k, 1)
        # the biased contrast: in the kernel for the two reference models
        model, bval = _MODEL_IDENTITY, 0.0
        if b is not None:
            if bias_model is linear_bias:
                model, bval = _MODEL_LINEAR, float(b)
            elif bias_model is loglinear_bias:
                model, bval = _MODEL_LOGLINEAR, float(b)
            else:
                dt, _ = _lib.to_device(d, T.float64)
                d = bias_model(dt, T.as_tensor(float(b), dtype=T.float64))
        if _lib.is_torch(d) and d.dtype == T.float32:
            dd = d.to("cuda").contiguous()
        else:
            dd, _ = _lib.to_device(d, T.float64)
        rb = 4 if dd.dtype == T.float32 else 8
        vd = None if v is None else _lib.to_device(v, T.float64)[0]
        scale = ARCMIN2_SPHERE / npix * g
        vptr = None if vd is None else _lib.ptr(vd)
        if device_rng:
            _lib.call("ga_points_counts", npix, rb, _lib.ptr(dd), model, bval,
                      int(bool(remove_monopole)), scale, vptr, int(seed), pop, None,
                      _lib.ptr(counts), _lib.ptr(total), _lib.ptr(ws), ws.numel(),
                      _lib.stream_ptr())
        else:
            lam = T.empty(npix, dtype=T.float64, device="cuda")
            _lib.call("ga_points_counts", npix, rb, _lib.ptr(dd), model, bval,
                      int(bool(remove_monopole)), scale, vptr, int(seed), pop, _lib.ptr(lam),
                      None, None, _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
            n = host_rng.poisson(lam.cpu().numpy())
            counts.copy_(T.as_tensor(np.asarray(n, dtype=np.int64)))
            _lib.call("ga_points_scan", npix, _lib.ptr(counts), _lib.ptr(total), _lib.ptr(ws),
                      ws.numel(), _lib.stream_ptr())
        ntot = int(total.item())
        if ntot == 0:
            continue
        lon = T.empty(ntot, dtype=T.float64, device="cuda")
        lat = T.empty(ntot, dtype=T.float64, device="cuda")
        _lib.call("ga_points_positions", nside, _lib.ptr(counts), int(seed), pop, _lib.ptr(lon),
                  _lib.ptr(lat), _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
        if dims:
            cmask = np.zeros(dims, dtype=np.int64)
            cmask[k] = 1
        else:
            cmask = 1
        for start in range(0, ntot, batch):
            stop = min(ntot, start + batch)
            lo, la = lon[start:stop], lat[start:stop]
            if not torch_out:
                lo, la = lo.cpu().numpy(), la.cpu().numpy()
            yield lo, la, (stop - start) * cmask
