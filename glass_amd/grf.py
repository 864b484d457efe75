"""Pointwise transformations of Gaussian random fields, mirroring
glass/grf/_transformations.py (the ``__call__`` part, which is on the
generation path; the corr/icorr/dcorr functions belong to the spectra solver).

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


@dataclass
class Normal:
    """glass/grf/_transformations.py:13-28: t(X) = X."""

    def __call__(self, x, _var: float, /):
        return x


@dataclass
class Lognormal:
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
class SquaredNormal:
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
