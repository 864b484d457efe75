"""Device-resident batched generation of correlated lognormal shells.

This is the throughput form of ``glass.generate`` (glass/fields.py:839-894)
for S shells that share one (nside, lmax): everything from the covariance rows
to the transformed maps stays in HBM and runs as five kernels per batch:

    iternorm (scaling vectors, one lane per l)
    -> correlated alm (Philox normals, multalm sum, healpy reorder, m = 0 fix)
    -> Legendre stage on FP32 MFMA (FP64 recurrence)
    -> ring FFT stage with the lognormal transform fused into the store.

Normals come from the counter-based device generator (``DeviceRNG``), keyed by
(seed, shell index, alm index), so a shell's map does not depend on which
batch it is computed in.
"""

from __future__ import annotations

import numpy as np

from . import _lib
from . import fields as _fields
from . import healpix as _hp


def shell_spectra(nshell: int, lmax: int, *, amp=lambda i: 1.0, rho=(0.6, 0.3, 0.1),
                  scale: float = 1500.0, tilt: float = -1.2):
    """Gaussian spectra in GLASS order for a chain of shells with neighbour
    correlation coefficients ``rho`` (the synthetic family of the workloads):
    g_l^i = amp(i) 1e-4 (l+1)^tilt exp(-(l/scale)^2), g_l^ij = rho_|i-j| sqrt(g^i g^j)."""
    ell = np.arange(lmax + 1, dtype=np.float64)
    auto = [amp(i) * 1e-4 * (ell + 1.0) ** tilt * np.exp(-((ell / scale) ** 2))
            for i in range(nshell)]
    gls = []
    for i in range(nshell):
        for j in range(i, -1, -1):
            d = i - j
            if d == 0:
                gls.append(auto[i])
            elif d <= len(rho):
                gls.append(rho[d - 1] * np.sqrt(auto[i] * auto[j]))
            else:
                gls.append(np.zeros(0))
    return gls


def shear_factor(lmax: int, pixwin_ratio=None) -> np.ndarray:
    """kappa -> gamma E-mode factor of shear_from_convergence
    (glass/lensing.py:414-417, and the product of the three almxfl factors of
    from_convergence, :316-363): -sqrt((l+2)(l+1) l (l-1)) / (l (l+1)), 0 for l < 2,
    times pw2/pw0 when a pixel-window ratio is given."""
    ell = np.arange(lmax + 1, dtype=np.float64)
    f = np.sqrt((ell + 2) * (ell + 1) * ell * np.clip(ell - 1, 0, None))
    f /= np.clip(ell * (ell + 1), 1, None)
    f *= -1
    if pixwin_ratio is not None:
        f *= np.asarray(pixwin_ratio, dtype=np.float64)[: lmax + 1]
    return f


class ShearShells:
    """Batched convergence -> shear for S maps of one geometry, device resident:
    FP32 analysis with the pixel quadrature weights and Jacobi refinement
    (map2alm(use_pixel_weights=True), iter passes; ``pixel_weights=False``:
    uniform weights), the kappa -> gamma factor (almxfl), spin-2 synthesis
    with B = 0.  The batched form of shear_from_convergence(kappa, lmax,
    discretized=False) (glass/lensing.py:374-428); gamma[s] = (gamma1, gamma2)
    of map s."""

    def __init__(self, nside: int, lmax: int, nshell: int, *, niter: int = 3,
                 pixwin_ratio=None, pixel_weights: bool = True):
        T = _lib.torch()
        _lib.require_gpu()
        self.nside, self.lmax, self.S, self.niter = int(nside), int(lmax), int(nshell), int(niter)
        self.npix = 12 * self.nside**2
        self.nalm = (self.lmax + 1) * (self.lmax + 2) // 2
        self.plan = _lib.get_plan(self.nside, self.lmax)
        self.plan.prepare_spin(2)
        # one workspace serves both steps (the analysis needs ring buffer +
        # partial sums for S maps, the spin synthesis a ring buffer for 2 S)
        nbytes = max(self.plan.workspace_bytes(self.S, 4, analysis=True),
                     self.plan.workspace_bytes(2 * self.S, 4))
        self.ws = _lib.workspace(nbytes)
        self.alm = T.empty((self.nalm, self.S), dtype=T.complex64, device="cuda")
        self.tmp = T.empty((self.S, self.npix), dtype=T.float32, device="cuda")
        self.fl = T.as_tensor(shear_factor(self.lmax, pixwin_ratio)).to("cuda")
        self.gamma = T.empty((self.S, 2, self.npix), dtype=T.float32, device="cuda")
        self.w = _hp.pixel_weights(self.nside) if pixel_weights else None

    def stage_analysis(self, kappa):
        if tuple(kappa.shape) != (self.S, self.npix) or not kappa.is_contiguous():
            raise ValueError("kappa must be a contiguous [nshell, npix] float32 device tensor")
        if self.w is not None:
            _lib.call("ga_map2alm_weighted_batch_f32", self.plan.h, self.S, _lib.ptr(kappa),
                      _lib.ptr(self.w), _lib.ptr(self.alm), self.niter, _lib.ptr(self.tmp),
                      _lib.ptr(self.ws), self.ws.numel(), _lib.stream_ptr())
        else:
            _lib.call("ga_map2alm_batch_f32", self.plan.h, self.S, _lib.ptr(kappa),
                      _lib.ptr(self.alm), self.niter, _lib.ptr(self.tmp), _lib.ptr(self.ws),
                      self.ws.numel(), _lib.stream_ptr())
        _lib.call("ga_almxfl", self.lmax, self.S, 4, _lib.ptr(self.alm), _lib.ptr(self.fl),
                  int(self.fl.numel()), _lib.ptr(self.alm), _lib.stream_ptr())

    def stage_spin(self):
        # almB = NULL: B = 0, one contraction
        _lib.call("ga_alm2map_spin_batch_f32", self.plan.h, 2, self.S, _lib.ptr(self.alm),
                  None, _lib.ptr(self.gamma), _lib.ptr(self.ws), self.ws.numel(),
                  _lib.stream_ptr())

    def run(self, kappa):
        self.stage_analysis(kappa)
        self.stage_spin()
        return self.gamma


class LognormalShells:
    """S correlated lognormal shells generated in one device batch (with
    ``shear=True`` also the shear of every shell, its map taken as the
    convergence: ``gamma`` [S][2][npix])."""

    def __init__(self, gls, nside: int, *, ncorr: int, shift: float = 1.0, seed: int = 0,
                 lmax: int | None = None, shear: bool = False, shear_lmax: int | None = None):
        T = _lib.torch()
        _lib.require_gpu()
        self.S = _fields.nfields_from_nspectra(len(gls))
        self.nside = int(nside)
        nl = max(np.asarray(g).shape[0] for g in gls if g is not None)
        self.lmax = nl - 1 if lmax is None else int(lmax)
        self.nl = self.lmax + 1
        self.k = int(ncorr)
        self.seed = int(seed)
        self.nalm = self.nl * (self.nl + 1) // 2
        self.npix = 12 * self.nside**2
        rows = [np.array(c) for c in _fields.cls2cov(gls, self.nl, self.S, self.k)]
        self.cov = T.as_tensor(np.stack(rows)).to("cuda")
        self.var = [_fields._cltovar(_fields.getcl(gls, i, i)) for i in range(self.S)]
        epi = (_lib.Epilogue * self.S)()
        for i, v in enumerate(self.var):
            epi[i] = _lib.Epilogue(_lib.T_LOGNORMAL, 0, float(shift), float(v))
        self.epi = T.frombuffer(bytearray(bytes(epi)), dtype=T.uint8).to("cuda")
        nd = _lib.ctypes.c_int32()
        _lib.call("ga_iternorm_state_doubles", self.k, _lib.ctypes.byref(nd))
        self.state = T.empty(self.nl * nd.value, dtype=T.float64, device="cuda")
        self.w = T.empty_like(self.cov)
        self.flag = T.zeros(1, dtype=T.int32, device="cuda")
        self.alm = T.empty((self.nalm, self.S), dtype=T.complex64, device="cuda")
        # per-block column maxima written by the field kernel, so that the
        # operand preparation does not read the whole alm batch for them
        self.colmax_fused = self.k + 1 <= 8
        self.nparts = 0
        self.colpart = None
        if self.colmax_fused:
            rows = _lib.ctypes.c_int64()
            _lib.call("ga_fields_colmax_rows", self.nl, _lib.ctypes.byref(rows))
            self.nparts = rows.value
            self.colpart = T.empty((self.nparts, 2 * self.S), dtype=T.float32, device="cuda")
        self.plan = _lib.get_plan(self.nside, self.lmax)
        self.F = _lib.workspace(self.plan.workspace_bytes(self.S, 4))
        self.maps = T.empty((self.S, self.npix), dtype=T.float32, device="cuda")
        self.shear = None
        if shear:
            self.shear = ShearShells(self.nside, self.lmax if shear_lmax is None else shear_lmax,
                                     self.S)

    @property
    def gamma(self):
        return None if self.shear is None else self.shear.gamma

    # the stages, separately callable so that each can be timed on the stream
    def stage_fields(self):
        _lib.call("ga_iternorm_batch", self.nl, self.k, self.k, self.k, self.S, 1,
                  _lib.ptr(self.cov), _lib.ptr(self.state), _lib.ptr(self.w),
                  _lib.ptr(self.flag), _lib.stream_ptr())
        if self.colmax_fused:
            _lib.call("ga_fields_alm_colmax", self.nl, self.k, 0, self.S, _lib.ptr(self.w),
                      self.seed, _lib.ptr(self.alm), 0, _lib.ptr(self.colpart),
                      _lib.stream_ptr())
        else:
            _lib.call("ga_fields_alm", self.nl, self.k, 0, self.S, _lib.ptr(self.w), None, 1,
                      self.seed, _lib.ptr(self.alm), 0, None, _lib.stream_ptr())

    def stage_legendre(self):
        self.stage_operands()
        self.stage_contract()

    def stage_operands(self):
        if self.colmax_fused:
            _lib.call("ga_legendre_prepare_colmax_batch_f32", self.plan.h, self.S,
                      _lib.ptr(self.alm), _lib.ptr(self.colpart), self.nparts, _lib.ptr(self.F),
                      self.F.numel(), _lib.stream_ptr())
        else:
            _lib.call("ga_legendre_prepare_batch_f32", self.plan.h, self.S, _lib.ptr(self.alm),
                      _lib.ptr(self.F), self.F.numel(), _lib.stream_ptr())

    def stage_contract(self):
        _lib.call("ga_legendre_contract_batch_f32", self.plan.h, self.S, _lib.ptr(self.alm),
                  _lib.ptr(self.F), self.F.numel(), _lib.stream_ptr())

    def stage_rings(self):
        _lib.call("ga_ring_synth_batch_f32", self.plan.h, self.S, _lib.ptr(self.F),
                  _lib.ptr(self.maps), _lib.ptr(self.epi), _lib.stream_ptr())

    def stage_shear(self):
        if self.shear is not None:
            self.shear.run(self.maps)

    def run(self):
        self.stage_fields()
        self.stage_legendre()
        self.stage_rings()
        self.stage_shear()
        return self.maps

    def check(self) -> None:
        if int(self.flag.item()):
            raise ValueError("covariance matrix is not positive definite")

    def fp64_reference_map(self, shell: int):
        """The same shell through the plain FP64 path (same device normals)."""
        T = _lib.torch()
        nd = _lib.ctypes.c_int32()
        _lib.call("ga_iternorm_state_doubles", self.k, _lib.ctypes.byref(nd))
        state = T.empty(self.nl * nd.value, dtype=T.float64, device="cuda")
        w = T.empty((shell + 1, self.nl, self.k + 1), dtype=T.float64, device="cuda")
        flag = T.zeros(1, dtype=T.int32, device="cuda")
        _lib.call("ga_iternorm_batch", self.nl, self.k, self.k, self.k, shell + 1, 1,
                  _lib.ptr(self.cov), _lib.ptr(state), _lib.ptr(w), _lib.ptr(flag),
                  _lib.stream_ptr())
        alm = T.empty(self.nalm, dtype=T.complex128, device="cuda")
        _lib.call("ga_fields_alm", self.nl, self.k, shell, 1, _lib.ptr(w[shell]), None, 1,
                  self.seed, None, 0, _lib.ptr(alm), _lib.stream_ptr())
        x = _hp._synth_f64(alm, self.nside, self.lmax)
        y = T.empty_like(x)
        _lib.call("ga_transform_f64", x.numel(), _lib.ptr(x),
                  _lib.Epilogue(_lib.T_LOGNORMAL, 0, 1.0, float(self.var[shell])), _lib.ptr(y),
                  _lib.stream_ptr())
        return y
