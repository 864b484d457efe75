"""Random-field generation on MI355X, mirroring glass/fields.py.

``generate`` keeps the reference's signature and lazy per-shell semantics
(glass/fields.py:839-894).  The work of each shell runs on the device:

1. ``cls2cov`` rows (host, cheap) -> device;
2. ``iternorm`` on the device for a batch of shells (one lane per mode l);
3. the correlated alm: sum_j w_t[l, j] z_{t-k+j} in GLASS order, reordered to
   healpy order with the m = 0 fix, in one kernel (glass/fields.py:401-425);
4. synthesis: ``precision="fp64"`` -> plain FP64 path per shell;
   ``precision="fp32"`` -> batched FP32-MFMA synthesis with the transformation
   fused into the ring-FFT epilogue.

Normals: with ``rng=None`` or a numpy Generator they are drawn on the host from
the same PCG64 stream and in the same order as the reference
(``standard_normal((z_size, 2)) @ [1, 1j]`` per shell, seed 42 by default), so
the result is the reference's on identical inputs.  With
``rng=glass_amd.rng.DeviceRNG(seed)`` they come from a counter-based Philox
generator on the device (statistically equivalent, not the numpy stream).
"""

from __future__ import annotations

import math
import warnings

import numpy as np

from . import _lib
from . import grf as _grf
from . import healpix as _hp
from .rng import DeviceRNG

__all__ = [
    "cls2cov", "generate", "generate_gaussian", "generate_lognormal", "getcl", "iternorm",
    "nfields_from_nspectra",
]


def _inv_triangle_number(t: int) -> int:
    """glass/fields.py:70-81."""
    n = math.floor(math.sqrt(2 * t))
    if n * (n + 1) // 2 != t:
        raise ValueError(f"not a triangle number: {t}")
    return n


def nfields_from_nspectra(nspectra: int) -> int:
    """glass/fields.py:84-98."""
    try:
        return _inv_triangle_number(nspectra)
    except ValueError:
        raise ValueError(f"invalid number of spectra: {nspectra}") from None


def _host(a):
    if a is None:
        return None
    if _lib.is_torch(a):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def getcl(cls, i: int, j: int, lmax: int | None = None):
    """glass/fields.py:525-560."""
    if j > i:
        i, j = j, i
    cl = cls[i * (i + 1) // 2 + i - j]
    if lmax is not None:
        if cl.shape[0] > lmax + 1:
            cl = cl[: lmax + 1]
        else:
            h = _host(cl)
            cl = np.pad(h, (0, lmax + 1 - h.shape[0]))
    return cl


def cls2cov(cls, nl: int, nf: int, nc: int):
    """glass/fields.py:191-236: yields ONE reused (nl, nc+1) float64 buffer per
    field: column i holds C_l of (field, field - i), zero-padded, halved."""
    cov = np.zeros((nl, nc + 1))
    end = 0
    for j in range(nf):
        begin, end = end, end + j + 1
        for i, cl in enumerate(cls[begin:end][: nc + 1]):
            c = np.zeros(0) if cl is None else np.asarray(_host(cl), dtype=np.float64)
            if i == 0 and np.any(c < 0):
                raise ValueError("negative values in cl")
            n = c.shape[0]
            cov[:n, i] = c
            cov[n:, i] = 0.0
        cov /= 2
        yield cov


def iternorm(cov):
    """glass/fields.py:101-188 on the device: one lane per covariance row
    (mode), the conditional state carried between rows in HBM."""
    T = _lib.torch()
    state = flag = None
    shape = None
    kind = "numpy"
    kprev = kmax = 0
    nl = 1
    for i, row in enumerate(cov):
        k = row.shape[-1] - 1
        if k < 0:
            raise ValueError("empty covariance matrix")
        if i == 0:
            shape = tuple(row.shape[:-1])
            kind = "torch" if _lib.is_torch(row) else "numpy"
            nl = int(np.prod(shape)) if shape else 1
            flag = T.zeros(1, dtype=T.int32, device="cuda")
        elif tuple(row.shape[:-1]) != shape:
            raise ValueError("shape mismatch in covariance")
        if state is None or k > kmax:
            state = _regrow_state(state, nl, kmax, max(k, 4))
            kmax = max(k, 4)
        r, _ = _lib.to_device(row, T.float64)
        r = r.reshape(-1, k + 1).contiguous()
        w = T.empty_like(r)
        _lib.call("ga_iternorm_batch", nl, kprev, k, kmax, 1, 1 if i == 0 else 0, _lib.ptr(r),
                  _lib.ptr(state), _lib.ptr(w), _lib.ptr(flag), _lib.stream_ptr())
        if int(flag.item()):
            raise ValueError("covariance matrix is not positive definite")
        kprev = k
        yield _lib.from_device(w.reshape(*shape, k + 1), kind)


def _regrow_state(state, nl: int, kold: int, knew: int):
    """Re-lay the per-mode iternorm state [M (k x k), a (k), s, scratch (k)] at a
    larger capacity."""
    T = _lib.torch()
    new = T.zeros((nl, knew * knew + 2 * knew + 1), dtype=T.float64, device="cuda")
    if state is not None:
        old = state.view(nl, kold * kold + 2 * kold + 1)
        Mo = old[:, : kold * kold].view(nl, kold, kold)
        Mn = new[:, : knew * knew].view(nl, knew, knew)
        Mn[:, :kold, :kold] = Mo
        new[:, knew * knew: knew * knew + kold] = old[:, kold * kold: kold * kold + kold]
        new[:, knew * knew + knew] = old[:, kold * kold + kold]
    return new.view(-1)


def _cltovar(cl) -> float:
    """transformcl.cltovar as used at glass/fields.py:890: sum (2l+1) C_l / 4 pi."""
    c = np.asarray(_host(cl), dtype=np.float64)
    return float(np.sum((2 * np.arange(c.shape[-1]) + 1) * c) / (4 * np.pi))


class _GRFEngine:
    """Device state of the iterative sampler for one call of generate."""

    def __init__(self, gls, nside: int, ncorr: int | None, rng, precision: str,
                 batch: int | None):
        T = _lib.torch()
        _lib.require_gpu()
        self.nside = int(nside)
        self.ngrf = nfields_from_nspectra(len(gls))
        self.k = self.ngrf - 1 if ncorr is None else int(ncorr)
        self.nl = max((_host(gl).shape[0] for gl in gls if gl is not None), default=0)
        if self.nl == 0:
            raise ValueError("all gls are empty")
        self.lmax = self.nl - 1
        self.z_size = self.nl * (self.nl + 1) // 2
        self.precision = precision
        self.batch = 1 if precision == "fp64" else int(batch or min(self.ngrf, 32))
        self.cov = cls2cov(gls, self.nl, self.ngrf, self.k)
        nd = _lib.ctypes.c_int32()
        _lib.call("ga_iternorm_state_doubles", self.k, _lib.ctypes.byref(nd))
        self.state = T.empty(self.nl * nd.value, dtype=T.float64, device="cuda")
        self.flag = T.zeros(1, dtype=T.int32, device="cuda")
        if isinstance(rng, DeviceRNG):
            self.host_rng = None
            self.seed = rng.seed
            self.zring = None
            self.nslot = 1
        else:
            self.host_rng = np.random.default_rng(42) if rng is None else rng
            self.seed = 0
            self.nslot = self.batch + self.k
            self.zring = T.empty((self.nslot, self.z_size), dtype=T.complex128, device="cuda")
        self.t = 0

    def next_batch(self, transforms):
        """Maps (device) of the next shells; transforms[i] = (kind, lam, par) or None."""
        T = _lib.torch()
        B = min(self.batch, self.ngrf - self.t)
        rows = np.stack([np.array(next(self.cov)) for _ in range(B)])  # cls2cov reuses its buffer
        cov = T.as_tensor(rows).to("cuda")
        w = T.empty_like(cov)
        _lib.call("ga_iternorm_batch", self.nl, self.k, self.k, self.k, B, 1 if self.t == 0 else 0,
                  _lib.ptr(cov), _lib.ptr(self.state), _lib.ptr(w), _lib.ptr(self.flag),
                  _lib.stream_ptr())
        if int(self.flag.item()):
            raise ValueError("covariance matrix is not positive definite")
        if self.host_rng is not None:
            for j in range(B):
                z = self.host_rng.standard_normal((self.z_size, 2)) @ np.asarray([1, 1j])
                self.zring[(self.t + j) % self.nslot].copy_(T.as_tensor(z))
        zptr = _lib.ptr(self.zring) if self.zring is not None else None
        if self.precision == "fp64":
            alm = T.empty(self.z_size, dtype=T.complex128, device="cuda")
            _lib.call("ga_fields_alm", self.nl, self.k, self.t, 1, _lib.ptr(w), zptr, self.nslot,
                      self.seed, None, 0, _lib.ptr(alm), _lib.stream_ptr())
            maps = [_hp._synth_f64(alm, self.nside, self.lmax)]
        else:
            alm = T.empty((self.z_size, B), dtype=T.complex64, device="cuda")
            _lib.call("ga_fields_alm", self.nl, self.k, self.t, B, _lib.ptr(w), zptr, self.nslot,
                      self.seed, _lib.ptr(alm), 0, None, _lib.stream_ptr())
            epi = (_lib.Epilogue * B)()
            for j in range(B):
                tr = transforms[j] or (_lib.T_NORMAL, 1.0, 0.0)
                epi[j] = _lib.Epilogue(tr[0], 0, tr[1], tr[2])
            epi_dev = T.frombuffer(bytearray(bytes(epi)), dtype=T.uint8).to("cuda")
            out = _hp.synth_batch_f32(alm, self.nside, self.lmax, epilogues=epi_dev)
            maps = [out[j] for j in range(B)]
        self.t += B
        return maps


def _generate(fields, gls, nside, ncorr, rng, precision, batch, fuse: bool):
    n = len(fields)
    kind = "torch" if any(_lib.is_torch(g) for g in gls if g is not None) else "numpy"
    variances = [_cltovar(getcl(gls, i, i)) for i in range(n)]
    eng = _GRFEngine(gls, nside, ncorr, rng, precision, batch)
    if eng.ngrf != n:
        raise ValueError("mismatch between number of fields and gls")
    i = 0
    while i < n:
        B = min(eng.batch, n - i)
        fused = [(_grf.epilogue(fields[i + j], variances[i + j]) if fuse else None)
                 for j in range(B)]
        xs = eng.next_batch(fused if precision == "fp32" else [None] * B)
        for j, x in enumerate(xs):
            t, var = fields[i + j], variances[i + j]
            if precision == "fp32":
                if fused[j] is not None:
                    yield _lib.from_device(x, kind)
                else:
                    yield t(_lib.from_device(x, kind), var)
            else:
                tr = _grf.epilogue(t, var) if fuse else None
                if tr is not None and tr[0] != _lib.T_NORMAL:
                    T = _lib.torch()
                    y = T.empty_like(x)
                    _lib.call("ga_transform_f64", x.numel(), _lib.ptr(x),
                              _lib.Epilogue(tr[0], 0, tr[1], tr[2]), _lib.ptr(y),
                              _lib.stream_ptr())
                    yield _lib.from_device(y, kind)
                elif tr is not None:
                    yield _lib.from_device(x, kind)
                else:
                    yield t(_lib.from_device(x, kind), var)
        i += B


def generate(fields, gls, nside: int, *, ncorr: int | None = None, rng=None,
             precision: str = "fp64", batch: int | None = None):
    """glass/fields.py:839-894: iteratively sample HEALPix maps of transformed
    Gaussian random fields with Gaussian angular power spectra ``gls``.

    Extra keywords: ``precision`` ("fp64" plain path, default; "fp32" batched
    FP32-MFMA path with the transformation fused) and ``batch`` (shells per
    device batch in the fp32 path).

    With ``precision="fp32"`` and host normals (``rng`` None or a numpy
    Generator), the normals of a whole batch of shells are drawn from ``rng``
    when the batch's first map is requested, i.e. ahead of the maps that are
    yielded later.  A caller that draws from the same ``rng`` between maps
    (as the reference's examples do for galaxy positions) therefore gets a
    different interleaving of the stream than the reference; pass ``batch=1``
    to keep the reference's draw order exactly, or use ``DeviceRNG``.
    """
    # a generator function, like the reference: the checks below run at the
    # first next(), not at the call (glass/fields.py:884-887)
    n = len(fields)
    if len(gls) != n * (n + 1) // 2:
        raise ValueError("mismatch between number of fields and gls")
    if precision not in ("fp64", "fp32"):
        raise ValueError(f"unknown precision {precision!r}")
    yield from _generate(fields, gls, nside, ncorr, rng, precision, batch, fuse=True)


def _generate_grf(gls, nside: int, *, ncorr: int | None = None, rng=None,
                  precision: str = "fp64"):
    """glass/fields.py:334-429: the Gaussian fields themselves."""
    n = nfields_from_nspectra(len(gls))
    return _generate([_grf.Normal() for _ in range(n)], gls, nside, ncorr, rng, precision,
                     None, fuse=True)


def generate_gaussian(gls, nside: int, *, ncorr: int | None = None, rng=None):
    """glass/fields.py:432-482 (deprecated in the reference)."""
    warnings.warn("use glass.generate() instead", DeprecationWarning, stacklevel=2)
    n = nfields_from_nspectra(len(gls))
    return generate([_grf.Normal() for _ in range(n)], gls, nside, ncorr=ncorr, rng=rng)


def generate_lognormal(gls, nside: int, shift: float = 1.0, *, ncorr: int | None = None,
                       rng=None):
    """glass/fields.py:485-522 (deprecated in the reference)."""
    warnings.warn("use glass.generate() instead", DeprecationWarning, stacklevel=2)
    n = nfields_from_nspectra(len(gls))
    return generate([_grf.Lognormal(shift) for _ in range(n)], gls, nside, ncorr=ncorr, rng=rng)


def _glass_to_healpix_alm(alm):
    """glass/fields.py:943-962 on the device."""
    T = _lib.torch()
    a, kind = _lib.to_device(alm, T.complex128)
    nl = _inv_triangle_number(int(a.numel()))
    out = T.empty_like(a)
    _lib.call("ga_glass_to_healpix_alm_f64", nl, _lib.ptr(a), _lib.ptr(out), _lib.stream_ptr())
    return _lib.from_device(out, kind)



# ---------------------------------------------------------------------------
# Spectra preparation (glass/fields.py:239-300, :562-620, :745-836)


def enumerate_spectra(entries):
    """glass/fields.py:563-580: (i, j, entry) in standard order."""
    for k, cl in enumerate(entries):
        i = int((2 * k + 0.25) ** 0.5 - 0.5)
        j = i * (i + 3) // 2 - k
        yield i, j, cl


def spectra_indices(n: int, *, xp=None):
    """glass/fields.py:583-605: rows (i, j) in standard order."""
    i, j = np.tril_indices(n)
    return np.transpose([i, i - j])


def compute_gaussian_spectra(fields, spectra):
    """glass/fields.py:745-775: band-limited Gaussian spectra (grf.compute)."""
    n = len(fields)
    if len(spectra) != n * (n + 1) // 2:
        raise ValueError("mismatch between number of fields and spectra")
    gls = []
    for i, j, cl in enumerate_spectra(spectra):
        cl = np.asarray(cl)
        gls.append(_grf.compute(cl, fields[i], fields[j]) if cl.shape[0] > 0 else 0 * cl)
    return gls


def solve_gaussian_spectra(fields, spectra):
    """glass/fields.py:778-836: Gaussian spectra from the iterative solver
    (grf.solve with zero padding 2 n, the monopole fixed to zero where the
    target monopole is zero; a warning where it did not converge)."""
    n = len(fields)
    if len(spectra) != n * (n + 1) // 2:
        raise ValueError("mismatch between number of fields and spectra")
    if len(spectra) == 0:
        return []
    gls = []
    for i, j, cl in enumerate_spectra(spectra):
        cl = np.asarray(cl, dtype=np.float64)
        if cl.shape[0] > 0:
            monopole = 0.0 if cl[0] == 0 else None
            gl, _cl_out, info = _grf.solve(cl, fields[i], fields[j], pad=2 * cl.shape[0],
                                           monopole=monopole)
            if info == 0:
                warnings.warn(f"Gaussian spectrum for fields ({i}, {j}) did not converge",
                              stacklevel=2)
        else:
            gl = 0 * cl
        gls.append(gl)
    return gls


def discretized_cls(cls, *, lmax: int | None = None, ncorr: int | None = None,
                    nside: int | None = None):
    """glass/fields.py:239-300: truncate to lmax, keep ncorr correlations, and
    multiply by the squared pixel window of nside (healpix.pixwin, computed on
    the GPU from its definition)."""
    if len(cls) == 0:
        return []
    if ncorr is not None:
        n = nfields_from_nspectra(len(cls))
        cls = [cls[i * (i + 1) // 2 + j] if j <= ncorr else np.asarray([])
               for i in range(n) for j in range(i + 1)]
    pw = _hp.pixwin(nside, lmax=lmax) if nside is not None else None
    gls = []
    for cl in cls:
        cl = np.asarray(cl)
        if cl.shape[0] > 0:
            if lmax is not None:
                cl = cl[: lmax + 1]
            if pw is not None:
                n = min(cl.shape[0], pw.shape[0])
                cl = cl[:n] * pw[:n] ** 2
        gls.append(cl)
    return gls
