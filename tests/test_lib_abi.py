"""CPU checks of the drop-in boundary: the HIP library loads here (no GPU
needed to dlopen it) and exports exactly what include/glass_amd.h declares."""

import ctypes
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "glass_amd.h"
LIB = ROOT / "glass_amd" / "_lib" / "libglass_amd.so"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ga_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from glass_amd import _lib

    assert set(declared()) == set(_lib.exported_symbols())


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_host_side_errors_without_gpu():
    """Argument checks run on the host: they report GA_ERR_INVALID and a message."""
    from glass_amd import _lib

    L = _lib.lib()
    assert L.ga_version().startswith(b"glass_amd")
    n = ctypes.c_int32()
    assert L.ga_iternorm_state_doubles(3, ctypes.byref(n)) == 0 and n.value == 16
    assert L.ga_iternorm_state_doubles(-1, ctypes.byref(n)) == _lib.GA_ERR_INVALID
    rc = L.ga_fields_alm(4, 1, 0, 1, None, None, 0, 0, None, 0, None, None)
    assert rc == _lib.GA_ERR_INVALID
    assert b"ga_fields_alm" in L.ga_last_error()
    with pytest.raises(ValueError, match="ga_plan_create"):
        _lib.call("ga_plan_create", 4, 8, None)


def test_api_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import numpy as np

    import glass_amd

    with pytest.raises(RuntimeError, match="GPU"):
        glass_amd.healpix.alm2map(np.zeros(6, complex), 2)
