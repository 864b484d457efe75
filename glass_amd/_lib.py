"""ctypes binding of ``libglass_amd.so`` (the C ABI declared in
``include/glass_amd.h``) plus the small amount of device plumbing the host
layer needs: plans, workspaces, streams and array-namespace round trips.

PyTorch-ROCm is used only for device memory and the current stream.  There is
no CPU fallback: every transform runs through the HIP library, and a missing
library or GPU raises immediately.
"""

from __future__ import annotations

import ctypes
import os
import pathlib
import threading

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
# GLASS_AMD_LIBRARY selects an alternate build of the same library (A/B timing)
LIB_PATH = pathlib.Path(os.environ.get("GLASS_AMD_LIBRARY", _HERE / "_lib" / "libglass_amd.so"))

GA_OK = 0
GA_ERR_INVALID = 1
GA_ERR_HIP = 2
GA_ERR_NOMEM = 3
GA_ERR_UNSUPPORTED = 4
GA_ERR_NOT_POSDEF = 5
GA_ERR_WORKSPACE = 6

TUNE_MFMA_TILES = 1
TUNE_MFMA_CHUNK = 2
TUNE_MFMA_PATH = 3
TUNE_RING_PATH = 4
TUNE_SPIN_PATH = 5
TUNE_ANAL_PATH = 6
TUNE_F16_FLUSH = 7
TUNE_M_GROUP = 8
TUNE_CHUNK_MIN = 9

T_NORMAL = 0
T_LOGNORMAL = 1
T_SQUARED_NORMAL = 2


class Epilogue(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("lamda", ctypes.c_double), ("param", ctypes.c_double)]


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_sz = ctypes.c_size_t

# name -> (argtypes); every entry point returns int status
_SIGNATURES = {
    "ga_set_tuning": [_i32, _i32],
    "ga_get_tuning": [_i32, ctypes.POINTER(_i32)],
    "ga_plan_create": [_i32, _i32, ctypes.POINTER(_vp)],
    "ga_plan_destroy": [_vp],
    "ga_plan_table_bytes": [_vp, ctypes.POINTER(_sz)],
    "ga_plan_set_tuning": [_vp, _i32, _i32],
    "ga_plan_get_tuning": [_vp, _i32, ctypes.POINTER(_i32)],
    "ga_plan_prepare_spin": [_vp, _i32],
    "ga_workspace_size": [_vp, _i32, _i32, ctypes.POINTER(_sz)],
    "ga_map2alm_workspace_size": [_vp, _i32, _i32, ctypes.POINTER(_sz)],
    "ga_alm2map_f64": [_vp, _vp, _vp, _vp, _sz, _vp],
    "ga_alm2map_batch_f32": [_vp, _i32, _vp, _vp, _vp, _vp, _sz, _vp],
    "ga_noalloc_scratch_bytes": [ctypes.POINTER(_sz)],
    "ga_alm2map_batch_f32_noalloc": [_vp, _i32, _vp, _vp, _vp, _vp],
    "ga_legendre_synth_batch_f32": [_vp, _i32, _vp, _vp, _sz, _vp],
    "ga_legendre_prepare_batch_f32": [_vp, _i32, _vp, _vp, _sz, _vp],
    "ga_legendre_contract_batch_f32": [_vp, _i32, _vp, _vp, _sz, _vp],
    "ga_legendre_prepare_colmax_batch_f32": [_vp, _i32, _vp, _vp, _i64, _vp, _sz, _vp],
    "ga_ring_synth_batch_f32": [_vp, _i32, _vp, _vp, _vp, _vp],
    "ga_legendre_contract_grouped_f32": [_i32, ctypes.POINTER(_vp), ctypes.POINTER(_i32),
                                         ctypes.POINTER(_vp), ctypes.POINTER(_sz), _vp],
    "ga_graph_begin": [_vp],
    "ga_graph_end": [_vp, ctypes.POINTER(_vp)],
    "ga_graph_launch": [_vp, _vp],
    "ga_graph_destroy": [_vp],
    "ga_map2alm_f64": [_vp, _vp, _vp, _i32, _vp, _vp, _sz, _vp],
    "ga_map2alm_batch_f32": [_vp, _i32, _vp, _vp, _i32, _vp, _vp, _sz, _vp],
    "ga_map2alm_weighted_f64": [_vp, _vp, _vp, _vp, _i32, _vp, _vp, _sz, _vp],
    "ga_map2alm_weighted_batch_f32": [_vp, _i32, _vp, _vp, _vp, _i32, _vp, _vp, _sz, _vp],
    "ga_pixel_weights_workspace_size": [_vp, ctypes.POINTER(_sz)],
    "ga_pixel_weights": [_vp, ctypes.c_double, _i32, _vp, ctypes.POINTER(_i32),
                         ctypes.POINTER(ctypes.c_double), _vp, _sz, _vp],
    "ga_alm2map_spin_f64": [_vp, _i32, _vp, _vp, _vp, _vp, _sz, _vp],
    "ga_alm2map_spin_batch_f32": [_vp, _i32, _i32, _vp, _vp, _vp, _vp, _sz, _vp],
    "ga_almxfl": [_i32, _i32, _i32, _vp, _vp, _i32, _vp, _vp],
    "ga_multalm_f64": [_i32, _vp, _vp, _vp, _vp],
    "ga_glass_to_healpix_alm_f64": [_i32, _vp, _vp, _vp],
    "ga_iternorm_state_doubles": [_i32, ctypes.POINTER(_i32)],
    "ga_iternorm_batch": [_i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp],
    "ga_fields_alm": [_i32, _i32, _i32, _i32, _vp, _vp, _i32, _u64, _vp, _i32, _vp, _vp],
    "ga_fields_colmax_rows": [_i32, ctypes.POINTER(_i64)],
    "ga_fields_alm_colmax": [_i32, _i32, _i32, _i32, _vp, _u64, _vp, _i32, _vp, _vp],
    "ga_transform_f64": [_i64, _vp, ctypes.POINTER(Epilogue), _vp, _vp],
    "ga_pixwin_workspace_size": [_i32, _i32, ctypes.POINTER(_sz)],
    "ga_pixwin": [_i32, _i32, _i32, _vp, _vp, _vp, _sz, _vp],
    "ga_multi_plane_update": [_i64, _i32, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double, _vp,
                              _vp],
    "ga_points_workspace_size": [_i64, ctypes.POINTER(_sz)],
    "ga_points_counts": [_i64, _i32, _vp, _i32, ctypes.c_double, _i32, ctypes.c_double, _vp, _u64,
                         _i64, _vp, _vp, _vp, _vp, _sz, _vp],
    "ga_points_scan": [_i64, _vp, _vp, _vp, _sz, _vp],
    "ga_points_positions": [_i32, _vp, _u64, _i64, _vp, _vp, _vp, _sz, _vp],
    "ga_legendre_nodes": [_i32, _vp, _vp, _vp],
    "ga_cl2corr": [_i32, _vp, _i32, _vp, _vp, _vp],
    "ga_corr2cl_workspace_size": [_i32, _i32, ctypes.POINTER(_sz)],
    "ga_corr2cl": [_i32, _vp, _vp, _vp, _i32, _vp, _vp, _sz, _vp],
}

_lib = None
_lock = threading.Lock()


class LibraryMissing(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not LIB_PATH.exists():
                    raise LibraryMissing(
                        f"glass_amd HIP library not found at {LIB_PATH}; "
                        "build it with `make -C glass_amd/csrc` (or __graft_entry__.build())")
                L = ctypes.CDLL(str(LIB_PATH))
                for name, args in _SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.argtypes = args
                    fn.restype = ctypes.c_int
                L.ga_last_error.restype = ctypes.c_char_p
                L.ga_version.restype = ctypes.c_char_p
                _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return sorted(list(_SIGNATURES) + ["ga_last_error", "ga_version"])


def check(rc: int, what: str = "") -> None:
    if rc == GA_OK:
        return
    msg = lib().ga_last_error().decode(errors="replace")
    if what:
        msg = f"{what}: {msg}"
    if rc in (GA_ERR_INVALID, GA_ERR_NOT_POSDEF):
        raise ValueError(msg)
    if rc == GA_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == GA_ERR_NOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


# ---------------------------------------------------------------------------
# torch plumbing


def torch():
    import torch as _torch

    return _torch


def require_gpu() -> None:
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("glass_amd needs a ROCm GPU (MI355X); none is visible")
    lib()


def stream_ptr():
    return ctypes.c_void_p(torch().cuda.current_stream().cuda_stream)


def ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


class Plan:
    """Device tables for one (nside, lmax) on one device."""

    def __init__(self, nside: int, lmax: int, device: int):
        self.nside, self.lmax, self.device = int(nside), int(lmax), int(device)
        self.h = ctypes.c_void_p()
        with torch().cuda.device(self.device):
            call("ga_plan_create", self.nside, self.lmax, ctypes.byref(self.h))
        self._spin = set()

    def prepare_spin(self, spin: int) -> None:
        if spin not in self._spin:
            with torch().cuda.device(self.device):
                call("ga_plan_prepare_spin", self.h, int(spin))
            self._spin.add(spin)

    def set_tuning(self, key: int, value: int) -> None:
        """This plan's knob ``key`` (``TUNE_*``); launches on the plan read it."""
        call("ga_plan_set_tuning", self.h, int(key), int(value))

    def get_tuning(self, key: int) -> int:
        v = ctypes.c_int32()
        call("ga_plan_get_tuning", self.h, int(key), ctypes.byref(v))
        return v.value

    def table_bytes(self) -> int:
        n = ctypes.c_size_t()
        call("ga_plan_table_bytes", self.h, ctypes.byref(n))
        return n.value

    def workspace_bytes(self, nbatch: int, real_bytes: int, analysis: bool = False) -> int:
        n = ctypes.c_size_t()
        fn = "ga_map2alm_workspace_size" if analysis else "ga_workspace_size"
        call(fn, self.h, int(nbatch), int(real_bytes), ctypes.byref(n))
        return n.value

    def __del__(self):
        try:
            if self.h and _lib is not None:
                _lib.ga_plan_destroy(self.h)
        except Exception:  # noqa: BLE001
            pass


class Graph:
    """Work recorded from one stream through the library's graph entry points
    (``ga_graph_begin`` / ``ga_graph_end``) and replayed with one call."""

    def __init__(self):
        self.h = ctypes.c_void_p()

    def capture(self, fn):
        """Record what ``fn()`` enqueues on the current torch stream (which
        must not be the default stream) and on the streams it joins."""
        s = stream_ptr()
        call("ga_graph_begin", s)
        try:
            fn()
        finally:
            rc = lib().ga_graph_end(s, ctypes.byref(self.h))
        check(rc, "ga_graph_end")
        return self

    def launch(self):
        call("ga_graph_launch", self.h, stream_ptr())

    def __del__(self):
        try:
            if self.h and _lib is not None:
                _lib.ga_graph_destroy(self.h)
        except Exception:  # noqa: BLE001
            pass


def contract_grouped(plans, nshells, workspaces) -> None:
    """``ga_legendre_contract_grouped_f32`` over up to 8 (plan, S, prepared
    workspace) triples, on the current stream."""
    n = len(plans)
    hp = (_vp * n)(*[p.h.value for p in plans])
    ns = (_i32 * n)(*[int(x) for x in nshells])
    wp = (_vp * n)(*[w.data_ptr() for w in workspaces])
    wb = (_sz * n)(*[int(w.numel()) for w in workspaces])
    call("ga_legendre_contract_grouped_f32", n, hp, ns, wp, wb, stream_ptr())


_plans: dict[tuple[int, int, int], Plan] = {}


def get_plan(nside: int, lmax: int, device: int | None = None) -> Plan:
    require_gpu()
    if device is None:
        device = torch().cuda.current_device()
    key = (int(nside), int(lmax), int(device))
    p = _plans.get(key)
    if p is None:
        p = Plan(nside, lmax, device)
        _plans[key] = p
    return p


def workspace(nbytes: int, device=None):
    t = torch()
    return t.empty(max(int(nbytes), 1), dtype=t.uint8, device=device or "cuda")


# ---------------------------------------------------------------------------
# array namespaces: numpy in -> numpy out, torch in -> torch out


def is_torch(x) -> bool:
    t = torch()
    return isinstance(x, t.Tensor)


def to_device(x, dtype):
    """(device tensor, kind) for a numpy array / torch tensor / sequence."""
    t = torch()
    require_gpu()
    if isinstance(x, t.Tensor):
        return x.to(device="cuda", dtype=dtype).contiguous(), "torch"
    # a fresh C-ordered copy: views with negative strides are not accepted by torch
    arr = np.array(x, order="C", copy=True)
    return t.as_tensor(arr).to(device="cuda", dtype=dtype), "numpy"


def from_device(x, kind):
    if kind == "torch":
        return x
    return x.cpu().numpy()


def complex_view(x):
    """Complex tensor -> interleaved real view (no copy)."""
    return torch().view_as_real(x)
