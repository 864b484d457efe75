"""Pointwise transformations of Gaussian random fields, mirroring
glass/grf/_transformations.py, and the Gaussian-spectra solver of
glass/grf/_core.py and glass/grf/_solver.py (below).

On float64 device tensors ``__call__`` runs one HIP kernel; numpy arrays and
other dtypes use their own namespace, as the reference does.  In the batched
FP32 path the same transform is fused into the synthesis epilogue (see
``epilogue``).
"""

from __future__ import annotations

import typing
from dataclasses import dataclass

import numpy as np

from . import _lib


class Transformation(typing.Protocol):
    """glass/grf/_core.py:18-35: t(x, var) -> transformed field."""

    def __call__(self, x, var: float, /): ...


def _device_transform(x, kind: int, lamda: float, param: float):
    T = _lib.torch()
    if isinstance(x, T.Tensor) and x.is_cuda and x.dtype == T.float64 and x.is_contiguous():
        out = T.empty_like(x)
        epi = _lib.Epilogue(kind, 0, float(lamda), float(param))
        _lib.call("ga_transform_f64", x.numel(), _lib.ptr(x), epi, _lib.ptr(out),
                  _lib.stream_ptr())
        return out
    return None


class _CorrMixin:
    """corr / icorr / dcorr against another transformation
    (glass/grf/_transformations.py), NotImplemented for unknown pairs."""

    def corr(self, other, x, /):
        return _corr_impl(self, other, x, "corr")

    def icorr(self, other, x, /):
        return _corr_impl(self, other, x, "icorr")

    def dcorr(self, other, x, /):
        return _corr_impl(self, other, x, "dcorr")


@dataclass
class Normal(_CorrMixin):
    """glass/grf/_transformations.py:13-28: t(X) = X."""

    def __call__(self, x, _var: float, /):
        return x


@dataclass
class Lognormal(_CorrMixin):
    """glass/grf/_transformations.py:61-89: t(X) = lambda [exp(X - var/2) - 1]."""

    lamda: float = 1.0

    def __call__(self, x, var: float, /):
        y = _device_transform(x, _lib.T_LOGNORMAL, self.lamda, var)
        if y is not None:
            return y
        if _lib.is_torch(x):
            T = _lib.torch()
            y = T.expm1(x - var / 2)
        else:
            y = np.expm1(x - var / 2)
        if self.lamda != 1.0:
            y = self.lamda * y
        return y


@dataclass
class SquaredNormal(_CorrMixin):
    """glass/grf/_transformations.py:140-175: t(X) = lambda [(X - a)^2 - 1]."""

    a: float
    lamda: float = 1.0

    def __call__(self, x, _var: float, /):
        y = _device_transform(x, _lib.T_SQUARED_NORMAL, self.lamda, self.a)
        if y is not None:
            return y
        y = (x - self.a) ** 2 - 1
        if self.lamda != 1.0:
            y = self.lamda * y
        return y


def epilogue(t, var: float):
    """The fused-epilogue form of a transformation, or None if it has none.

    Recognises this module's classes and, by name and fields, the reference's
    glass.grf classes, so reference-constructed field lists keep the fused path.
    """
    name = type(t).__name__
    if name == "Normal":
        return (_lib.T_NORMAL, 1.0, 0.0)
    if name == "Lognormal" and hasattr(t, "lamda"):
        return (_lib.T_LOGNORMAL, float(t.lamda), float(var))
    if name == "SquaredNormal" and hasattr(t, "a"):
        return (_lib.T_SQUARED_NORMAL, float(getattr(t, "lamda", 1.0)), float(t.a))
    return None


# ---------------------------------------------------------------------------
# Correlation functions of the transformations and the Gaussian-spectra
# solver (glass/grf/_transformations.py corr / icorr / dcorr,
# glass/grf/_core.py:63-184, glass/grf/_solver.py:27-148).  The transforms
# between C_l and C(theta) run on the GPU (ga_cl2corr / ga_corr2cl on
# Gauss-Legendre nodes, exact for l < n); the solver's few elementwise steps
# run on the host, as the reference's numpy code does.


def _corr_impl(t, other, x, kind: str):
    """The corr / icorr / dcorr of transformation t against other, or
    NotImplemented (glass/grf/_transformations.py)."""
    a, b = type(t).__name__, type(other).__name__
    if a == "Normal" and b == "Normal":
        return {"corr": x, "icorr": x, "dcorr": 1.0 + 0 * x}[kind]
    if a == "Lognormal":
        if b == "Lognormal":
            ll = t.lamda * other.lamda
            if kind == "corr":
                return ll * np.expm1(x)
            if kind == "icorr":
                return np.log1p(x / ll)
            return ll * np.exp(x)
        if b == "Normal":
            if kind == "corr":
                return t.lamda * x
            if kind == "icorr":
                return x / t.lamda
            return t.lamda + 0.0 * x
    if a == "SquaredNormal" and b == "SquaredNormal":
        aa = t.a * other.a
        ll = t.lamda * other.lamda
        if kind == "corr":
            return 2 * ll * x * (x + 2 * aa)
        if kind == "icorr":
            return np.sqrt(x / (2 * ll) + aa**2) - aa
        return 4 * ll * (x + aa)
    return NotImplemented


def _dispatch(kind: str, t1, t2, x):
    for a, b in ((t1, t2), (t2, t1)):
        r = getattr(a, kind)(b, x)
        if r is not NotImplemented:
            return r
    raise NotImplementedError(f"{type(t1).__name__} x {type(t2).__name__}")


def corr(t1, t2, x, /):
    """glass/grf/_core.py:63-86."""
    return _dispatch("corr", t1, t2, np.asarray(x, dtype=np.float64))


def icorr(t1, t2, x, /):
    """glass/grf/_core.py:89-112."""
    return _dispatch("icorr", t1, t2, np.asarray(x, dtype=np.float64))


def dcorr(t1, t2, x, /):
    """glass/grf/_core.py:115-138."""
    return _dispatch("dcorr", t1, t2, np.asarray(x, dtype=np.float64))


_nodes_cache: dict = {}


def _nodes(n: int):
    T = _lib.torch()
    key = (int(n), T.cuda.current_device())
    if key not in _nodes_cache:
        x = T.empty(n, dtype=T.float64, device="cuda")
        w = T.empty(n, dtype=T.float64, device="cuda")
        _lib.call("ga_legendre_nodes", int(n), _lib.ptr(x), _lib.ptr(w), _lib.stream_ptr())
        _nodes_cache[key] = (x, w)
    return _nodes_cache[key]


def cltocorr(cl) -> np.ndarray:
    """C(theta) of ``cl`` at the len(cl) Gauss-Legendre nodes (ascending
    cos theta); inverse of ``corrtocl``."""
    T = _lib.torch()
    _lib.require_gpu()
    c = T.as_tensor(np.ascontiguousarray(cl, dtype=np.float64)).cuda()
    n = int(c.numel())
    x, _ = _nodes(n)
    out = T.empty(n, dtype=T.float64, device="cuda")
    _lib.call("ga_cl2corr", n, _lib.ptr(c), n, _lib.ptr(x), _lib.ptr(out), _lib.stream_ptr())
    return out.cpu().numpy()


def corrtocl(corr) -> np.ndarray:
    """C_l, l < n, of C(theta) given at the n Gauss-Legendre nodes."""
    T = _lib.torch()
    _lib.require_gpu()
    c = T.as_tensor(np.ascontiguousarray(corr, dtype=np.float64)).cuda()
    n = int(c.numel())
    x, w = _nodes(n)
    nb = _lib.ctypes.c_size_t()
    _lib.call("ga_corr2cl_workspace_size", n, n, _lib.ctypes.byref(nb))
    ws = _lib.workspace(nb.value)
    out = T.empty(n, dtype=T.float64, device="cuda")
    _lib.call("ga_corr2cl", n, _lib.ptr(x), _lib.ptr(w), _lib.ptr(c), n, _lib.ptr(out),
              _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
    return out.cpu().numpy()


def compute(cl, t1, t2=None):
    """glass/grf/_core.py:141-184: band-limited Gaussian spectrum for cl."""
    if t2 is None:
        t2 = t1
    return corrtocl(icorr(t1, t2, cltocorr(cl)))


def _relerr(dx, x) -> float:
    q = np.divide(dx, x, where=(dx != 0), out=np.zeros_like(dx))
    return float(np.fabs(q).max())


def solve(cl, t1, t2=None, *, pad: int = 0, initial=None, cltol: float = 1e-5,
          gltol: float = 1e-5, maxiter: int = 20, monopole: float | None = None):
    """glass/grf/_solver.py:27-148: the Gauss-Newton solver of arXiv:2302.01942
    for a Gaussian spectrum whose transformed spectrum is cl.  Returns
    (gl, realised cl of length len(cl) + pad, info)."""
    if t2 is None:
        t2 = t1
    cl = np.asarray(cl, dtype=np.float64)
    n = cl.shape[0]
    if pad < 0:
        raise ValueError("pad must be a positive integer")
    if initial is None:
        gl = corrtocl(icorr(t1, t2, cltocorr(cl)))
    else:
        gl = np.zeros(n)
        gl[: initial.shape[0]] = initial[:n]
    if monopole is not None:
        gl[0] = monopole
    gt = cltocorr(np.pad(gl, (0, pad)))
    rl = corrtocl(corr(t1, t2, gt))
    fl = rl[:n] - cl
    if monopole is not None:
        fl[0] = 0
    clerr = _relerr(fl, cl)
    info = 0
    for _ in range(maxiter):
        if clerr <= cltol:
            info |= 1
        if info > 0:
            break
        ft = cltocorr(np.pad(fl, (0, pad)))
        dt = dcorr(t1, t2, gt)
        xl = -corrtocl(ft / dt)[:n]
        if monopole is not None:
            xl[0] = 0
        while True:
            gl_ = gl + xl
            gt_ = cltocorr(np.pad(gl_, (0, pad)))
            rl_ = corrtocl(corr(t1, t2, gt_))
            fl_ = rl_[:n] - cl
            if monopole is not None:
                fl_[0] = 0
            clerr_ = _relerr(fl_, cl)
            if clerr_ <= clerr:
                break
            xl /= 2
        if _relerr(xl, gl) <= gltol:
            info |= 2
        gl, gt, rl, fl, clerr = gl_, gt_, rl_, fl_, clerr_
    return gl, rl, info
