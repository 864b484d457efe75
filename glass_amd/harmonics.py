"""glass/harmonics.py on MI355X."""

from __future__ import annotations

from . import _lib


def multalm(alm, bl):
    """glass/harmonics.py:17-47: GLASS-order alm[l(l+1)/2 + m] * bl[l]."""
    T = _lib.torch()
    a, kind = _lib.to_device(alm, T.complex128)
    b, _ = _lib.to_device(bl, T.float64)
    nl = int(b.numel())
    if a.numel() != nl * (nl + 1) // 2:
        raise ValueError(f"alm of size {a.numel()} does not match bl of size {nl}")
    out = T.empty_like(a)
    if nl:
        _lib.call("ga_multalm_f64", nl, _lib.ptr(a), _lib.ptr(b), _lib.ptr(out), _lib.stream_ptr())
    res = _lib.from_device(out, kind)
    if kind == "numpy" and not _is_complex(alm):
        return res.real.astype(getattr(alm, "dtype", res.real.dtype), copy=False)
    return res


def _is_complex(x) -> bool:
    import numpy as np

    T = _lib.torch()
    if isinstance(x, T.Tensor):
        return x.is_complex()
    return np.iscomplexobj(x)
