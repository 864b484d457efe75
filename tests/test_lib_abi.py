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


def test_alm_size_mismatch_raises_before_any_launch():
    """alm whose length is not nalm(lmax) are rejected on the host (ValueError,
    no device touched): FP32 and FP64 alm2map, and E or B of alm2map_spin."""
    import numpy as np

    from glass_amd import healpix as hp

    good = np.zeros(hp.nalm(8), complex)
    short = np.zeros(hp.nalm(7), complex)
    for precision in ("fp32", "fp64"):
        with pytest.raises(ValueError, match="alm size does not match lmax"):
            hp.alm2map(good, 4, lmax=9, pol=False, precision=precision)
        with pytest.raises(ValueError, match="alm size does not match lmax"):
            hp.alm2map([good, short], 4, pol=False, precision=precision)
    with pytest.raises(ValueError, match="alm size does not match lmax"):
        hp.alm2map_spin([good, short], 4, 2, 8)
    with pytest.raises(ValueError, match="alm size does not match lmax"):
        hp.alm2map_spin([short, good], 4, 2, 8)
    with pytest.raises(ValueError, match="one-dimensional"):
        hp.alm2map(np.zeros((2, 2, hp.nalm(3)), complex), 4, pol=False)


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_tuning_default_is_checked_and_restorable():
    """The process default of the tuning knobs rejects out-of-range values
    with GA_ERR_INVALID and keeps the old value; plan setters reject a NULL
    plan (plan creation itself needs a device)."""
    from glass_amd import _lib

    L = _lib.lib()
    v = ctypes.c_int32()
    assert L.ga_get_tuning(_lib.TUNE_MFMA_PATH, ctypes.byref(v)) == 0
    old = v.value
    assert L.ga_set_tuning(_lib.TUNE_MFMA_PATH, 9) == _lib.GA_ERR_INVALID
    assert L.ga_get_tuning(_lib.TUNE_MFMA_PATH, ctypes.byref(v)) == 0 and v.value == old
    assert L.ga_set_tuning(99, 0) == _lib.GA_ERR_INVALID
    # accumulation forms: one chain (-1), two chunks (-2, -3), master sums (n > 0)
    assert L.ga_set_tuning(_lib.TUNE_F16_FLUSH, -4) == _lib.GA_ERR_INVALID
    assert L.ga_set_tuning(_lib.TUNE_CHUNK_MIN, -1) == _lib.GA_ERR_INVALID
    assert L.ga_set_tuning(_lib.TUNE_M_GROUP, 2) == _lib.GA_ERR_INVALID
    assert L.ga_plan_set_tuning(None, _lib.TUNE_MFMA_PATH, 0) == _lib.GA_ERR_INVALID
    assert L.ga_plan_get_tuning(None, _lib.TUNE_MFMA_PATH, ctypes.byref(v)) == _lib.GA_ERR_INVALID


def test_generate_raises_lazily_like_the_reference():
    """generate is a generator function: a count mismatch raises at the first
    next(), not at the call (glass/fields.py:884-887)."""
    import numpy as np

    import glass_amd

    it = glass_amd.generate([glass_amd.grf.Normal()] * 2, [np.ones(4)], 4)
    with pytest.raises(ValueError, match="mismatch between number of fields and gls"):
        next(it)
