"""ctypes binding of libholo.so (the C ABI declared in include/holo.h).

There is no fallback: if the library is missing or cannot be loaded, every call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent          # particleholography.jl_amd/
LIB_PATH = Path(os.environ.get("HOLO_LIB", PKG_ROOT / "lib" / "libholo.so"))

HOLO_OK, HOLO_ERR_ARG, HOLO_ERR_SIZE, HOLO_ERR_TYPE, HOLO_ERR_HIP, HOLO_ERR_NOMEM, HOLO_ERR_BUDGET = range(7)
HOLO_N0F8, HOLO_F16, HOLO_F32, HOLO_F64, HOLO_C64 = range(5)
HOLO_TF_REFERENCE, HOLO_TF_FP64_D, HOLO_TF_FP64 = range(3)
HOLO_LPF_RECTANGLE, HOLO_LPF_SUPER_GAUSSIAN = range(2)
HOLO_BACKGROUND_MODE, HOLO_BACKGROUND_MEAN = range(2)
HOLO_CCL_F32, HOLO_CCL_BINARY, HOLO_CCL_N0F8_LE = range(3)


class ArgumentError(ValueError):
    """The reference's ArgumentError (e.g. an unsupported return_type, holofunc.jl:164)."""


class HoloError(RuntimeError):
    """HIP runtime / allocation failure inside libholo."""


class holo_tf(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("z", ctypes.c_double), ("wavlen", ctypes.c_double),
                ("dx", ctypes.c_double), ("generated", ctypes.c_int)]


class holo_job(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("src", ctypes.c_void_p),
                ("src_real", ctypes.c_int), ("front", holo_tf), ("dz", holo_tf),
                ("slices", ctypes.c_int), ("proj", ctypes.c_void_p)]


class holo_gs_job(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("holo1", ctypes.c_void_p),
                ("holo2", ctypes.c_void_p), ("fwd", holo_tf), ("bwd", holo_tf),
                ("iters", ctypes.c_int), ("out", ctypes.c_void_p)]


class holo_box(ctypes.Structure):
    _fields_ = [("label", ctypes.c_uint32), ("i0", ctypes.c_int32), ("j0", ctypes.c_int32),
                ("i1", ctypes.c_int32), ("j1", ctypes.c_int32)]


_vp, _i, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_tfp = ctypes.POINTER(holo_tf)

# name -> argtypes (restype is int for all but holo_last_error)
SIGNATURES = {
    "holo_create": [ctypes.POINTER(_vp), _i],
    "holo_destroy": [_vp],
    "holo_last_error": [],
    "holo_version": [],
    "holo_supported_length": [_i],
    "holo_set_tuning": [_vp, _i, _i, _i, _i, _i],
    "holo_transfer_sqrt": [_vp, _vp, _i, _i, _d, _d, _i, _vp],
    "holo_transfer": [_vp, _vp, _i, _i, _d, _d, _vp, _i, _i, _d, _vp],
    "holo_gabor_wavefront": [_vp, _vp, _i, _i, _vp, _vp],
    "holo_reconst_xyprojection": [_vp, _vp, _i, _i, _vp, _tfp, _tfp, _i, _vp],
    "holo_reconst_xyprojection_holo": [_vp, _vp, _i, _i, _vp, _tfp, _tfp, _i, _vp],
    "holo_reconst_xyprojection_noalloc": [_vp, _vp, _i, _i, _vp, _tfp, _tfp, _i, _vp],
    "holo_reconst_vol": [_vp, _vp, _i, _i, _vp, _tfp, _tfp, _i, _i, _vp, _vp],
    "holo_2d_pad": [_vp, _vp, _i, _i, _vp, _vp],
    "holo_reconst_vol_and_xyprojection_padded": [_vp, _vp, _i, _i, _vp, _tfp, _tfp, _i, _vp, _vp],
    "holo_phase_retrieval": [_vp, _vp, _i, _i, _vp, _vp, _tfp, _tfp, _i, _vp, _i],
    "holo_asm_prop": [_vp, _vp, _i, _i, _vp, _vp, _tfp],
    "holo_fft2": [_vp, _vp, _i, _i, _vp, _vp, _i],
    "holo_lowpass_filter": [_vp, _vp, _i, _d, _d, _i, _d, _vp],
    "holo_apply_low_pass_filter": [_vp, _vp, _i, _i, _vp, _vp, _vp],
    "holo_ingest_u8": [_vp, _vp, _i, _i, _i, _vp, _vp, _d, _vp],
    "holo_make_background": [_vp, _vp, _i, _i, _i, _vp, _i, _vp],
    "holo_ccl_label": [_vp, _vp, _i, _i, _i, _vp, _i, _d, _vp],
    "holo_ccl_boxes": [_vp, _vp, _i, _i, _i, _vp, _i, _d, _i, _vp, _vp],
    "holo_label_boxes": [_vp, _vp, _i, _i, _i, _vp, _i, _vp, _vp],
    "holo_dilate": [_vp, _vp, _i, _i, _i, _vp, _vp],
    "holo_link_rectangles": [_vp, _vp, _i, _i, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_int64)],
    "holo_reconst_batch": [_vp, _vp, ctypes.POINTER(holo_job), _i],
    "holo_phase_retrieval_batch": [_vp, _vp, ctypes.POINTER(holo_gs_job), _i],
    "holo_set_tuning_for_size": [_vp, _i, _i, _i, _i, _i, _i, _i],
    "holo_set_workspace_limit": [_vp, ctypes.c_size_t],
    "holo_reserve_workspace": [_vp, ctypes.c_size_t],
    "holo_workspace_size": [_vp, ctypes.POINTER(ctypes.c_size_t)],
    "holo_get_tuning_for_size": [_vp, _i, _i, _i, ctypes.POINTER(ctypes.c_int)],
    "holo_size_variants": [_vp, _i, _i, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _i],
    "holo_autotune": [_vp, _vp, _i, _i, _d, _d, ctypes.POINTER(ctypes.c_int)],
    "holo_set_profiling": [_vp, _i],
    "holo_set_gs_graphs": [_vp, _i],
    "holo_profile_read": [_vp, _i, ctypes.POINTER(_d), ctypes.POINTER(ctypes.c_longlong),
                          ctypes.POINTER(ctypes.c_longlong), _i],
}
HOLO_PASS_RECON_COLS, HOLO_PASS_RECON_ROWS, HOLO_PASS_FWD_FFT, HOLO_PASS_PROP_COLS, HOLO_PASS_GS_ROWS = range(5)

_lib = None
_lock = threading.Lock()


def load() -> ctypes.CDLL:
    """Load libholo.so once; raise if it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise ImportError(f"libholo.so not found at {LIB_PATH}; build it with "
                              f"`make -C {PKG_ROOT / 'csrc'}`")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_char_p if name == "holo_last_error" else ctypes.c_int
        _lib = lib
        return lib


def check(code: int) -> None:
    if code == HOLO_OK:
        return
    msg = load().holo_last_error().decode(errors="replace")
    if code == HOLO_ERR_ARG:
        raise AssertionError(msg)
    if code == HOLO_ERR_TYPE:
        raise ArgumentError(msg)
    if code == HOLO_ERR_SIZE:
        raise ValueError(msg)
    if code == HOLO_ERR_BUDGET:
        raise MemoryError(msg)
    raise HoloError(f"libholo error {code}: {msg}")


class Handle:
    """One libholo context (plans, workspace, graph cache) on one device."""

    def __init__(self, device: int = 0):
        lib = load()
        h = ctypes.c_void_p()
        check(lib.holo_create(ctypes.byref(h), device))
        self.ptr = h
        self.device = device
        self.lib = lib

    def close(self):
        if self.ptr:
            self.lib.holo_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, batch=0, spg_gen=0, spg_out=0, cols_per_wg=0, rows_per_wg=0):
        check(self.lib.holo_set_tuning(self.ptr, batch, spg_gen, spg_out, cols_per_wg, rows_per_wg))

    def set_workspace_limit(self, nbytes: int):
        check(self.lib.holo_set_workspace_limit(self.ptr, int(nbytes)))

    def reserve_workspace(self, nbytes: int):
        """Allocates the handle's device workspace up front (at least ``nbytes``)."""
        check(self.lib.holo_reserve_workspace(self.ptr, int(nbytes)))

    def workspace_size(self) -> int:
        out = ctypes.c_size_t()
        check(self.lib.holo_workspace_size(self.ptr, ctypes.byref(out)))
        return out.value

    def set_gs_graphs(self, enable: bool):
        """Batched GS: run the chunk-uniform iterations as cached hipGraphs (default) or directly."""
        check(self.lib.holo_set_gs_graphs(self.ptr, 1 if enable else 0))

    def set_profiling(self, enable: bool):
        check(self.lib.holo_set_profiling(self.ptr, 1 if enable else 0))

    def profile_read(self, pass_id: int, reset: bool = True):
        """(total_ms, launches, units) of one pass since the last reset."""
        ms, n, u = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_longlong()
        check(self.lib.holo_profile_read(self.ptr, pass_id, ctypes.byref(ms), ctypes.byref(n),
                                         ctypes.byref(u), 1 if reset else 0))
        return ms.value, n.value, u.value


_handles: dict = {}


def handle(device: int = 0) -> Handle:
    """Per (thread, device) handle: a libholo handle is not thread safe."""
    key = (threading.get_ident(), device)
    h = _handles.get(key)
    if h is None:
        h = Handle(device)
        _handles[key] = h
    return h
