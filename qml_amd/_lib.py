"""ctypes binding of libqmlhip.so (the C ABI declared in include/qmlhip.h).

The library is the only compute path: if it is missing or no GPU is visible, every entry
point raises; nothing falls back to the CPU. Device buffers are torch tensors on the current
HIP device (torch is used for allocation and streams only).
"""

from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
# QMLH_LIBRARY may name another build of the same C ABI (e.g. a tuning build).
LIB_PATH = os.environ.get("QMLH_LIBRARY", os.path.join(_HERE, "libqmlhip.so"))

QMLH_OK = 0
PAIR_LAPLACIAN, PAIR_GAUSSIAN, PAIR_MANHATTAN, PAIR_L2 = 0, 1, 2, 3
PAIR_SARGAN, PAIR_MATERN_L2, PAIR_P_INT, PAIR_P_REAL = 4, 5, 6, 7
SYMMETRIC = 1
LOWER = 2
CM_MAX_SIZE = 256
ACM_MAX_ATOMS = 256
ACM_ROW_NORM, ACM_DISTANCE = 0, 1

_lock = threading.Lock()
_lib = None
_inited_devices: set[int] = set()


class QmlhError(RuntimeError):
    """A libqmlhip call returned a non-zero status."""


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_f64 = ctypes.c_double
_sz = ctypes.c_size_t

class Problem(ctypes.Structure):
    """struct qmlh_problem."""

    _fields_ = [("n", ctypes.c_int64), ("x_row0", ctypes.c_int64), ("y_offset", ctypes.c_int64),
                ("alpha_offset", ctypes.c_int64), ("sigma", ctypes.c_double),
                ("lambda_", ctypes.c_double), ("kind", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


# name -> (restype, argtypes); mirrors include/qmlhip.h
_SIGNATURES = {
    "qmlh_version": (ctypes.c_int, []),
    "qmlh_last_error": (ctypes.c_char_p, []),
    "qmlh_init": (ctypes.c_int, []),
    "qmlh_cm_batch": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i32, _i32, _vp, _i64, _vp, _vp]),
    "qmlh_atomic_cm_batch": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _i32, _i32,
                                            _f64, _f64, _f64, _f64, _vp, _i64, _vp, _vp]),
    "qmlh_atomic_cm_large_workspace": (_sz, [_i64, _i32]),
    "qmlh_atomic_cm_large": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _i32, _i32,
                                            _f64, _f64, _f64, _f64, _vp, _i64, _vp, _i64, _vp,
                                            _sz, _vp]),
    "qmlh_local_kernels": (ctypes.c_int, [_i32, _vp, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _i64,
                                          _i64, _vp, _vp, _vp, _i64, _i64, _vp, _i32, _vp, _i64,
                                          _i64, _i32, _vp]),
    "qmlh_local_panels": (_i64, [_vp, _i64, _vp]),
    "qmlh_pairwise": (ctypes.c_int, [_i32, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _f64, _vp,
                                     _i64, _i32, _vp]),
    "qmlh_row_sqnorms": (ctypes.c_int, [_vp, _i64, _i64, _i64, _vp, _vp]),
    "qmlh_pairwise_ext": (ctypes.c_int, [_i32, _vp, _i64, _i64, _vp, _i64, _i64, _i64, _vp, _i32,
                                         _vp, _i64, _i32, _vp]),
    "qmlh_linear_mfma": (ctypes.c_int, [_vp, _i64, _i64, _vp, _i64, _i64, _i64, _vp, _i64, _i32,
                                        _vp]),
    "qmlh_gaussian_mfma": (ctypes.c_int, [_vp, _i64, _i64, _vp, _vp, _i64, _i64, _vp, _i64,
                                          _f64, _vp, _i64, _i32, _vp]),
    "qmlh_potrf": (ctypes.c_int, [_vp, _i64, _i64, _vp, _vp]),
    "qmlh_potrs_workspace": (_sz, [_i64]),
    "qmlh_potrs": (ctypes.c_int, [_vp, _i64, _i64, _vp, _vp, _sz, _vp]),
    "qmlh_cho_invert_workspace": (_sz, [_i64]),
    "qmlh_cho_invert": (ctypes.c_int, [_vp, _i64, _i64, _vp, _vp, _sz, _vp]),
    "qmlh_bkf_workspace": (_sz, [_i64]),
    "qmlh_bkf_factor": (ctypes.c_int, [_vp, _i64, _i64, _vp, _vp, _vp, _sz, _vp]),
    "qmlh_bkf_solve": (ctypes.c_int, [_vp, _i64, _i64, _vp, _vp, _i64, _i64, _vp]),
    "qmlh_spotrf": (ctypes.c_int, [_vp, _i64, _i64, _vp, _vp]),
    "qmlh_spotrf_f16x3_workspace": (_sz, [_i64]),
    "qmlh_spotrf_f16x3": (ctypes.c_int, [_vp, _i64, _i64, _vp, _vp, _sz, _vp]),
    "qmlh_gaussian_deflated": (ctypes.c_int, [_vp, _i64, _i64, _vp, _i64, _f64, _f64, _f64, _vp,
                                              _i64, _vp, _i64, _vp, _i64, _vp, _vp]),
    "qmlh_woodbury_workspace": (_sz, [_i64]),
    "qmlh_spotrf_lowrank": (ctypes.c_int, [_vp, _i64, _i64, _vp, _i64, _i64, _i32, _f64, _vp,
                                           _vp, _sz, _vp]),
    "qmlh_spotrf_lowrank_shift": (ctypes.c_int, [_vp, _i64, _i64, _vp, _i64, _i64, _i32, _f64,
                                                 _f64, _f64, _vp, _vp, _vp, _sz, _vp]),
    "qmlh_woodbury_capacitance_workspace": (_sz, [_i64, _i64]),
    "qmlh_woodbury_capacitance": (ctypes.c_int, [_vp, _i64, _i64, _i64, _vp, _i64, _vp, _vp, _sz,
                                                 _vp]),
    "qmlh_woodbury_setup": (ctypes.c_int, [_vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _vp,
                                           _vp, _sz, _vp]),
    "qmlh_lower_to_f32": (ctypes.c_int, [_vp, _i64, _i64, _vp, _i64, _f64, _vp]),
    "qmlh_pcg_workspace": (_sz, [_i64, _i64]),
    "qmlh_pcg_solve": (ctypes.c_int, [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _i64, _i32, _vp,
                                      _i64, _vp, _vp, _f64, _i32, _vp, _sz, ctypes.POINTER(_i32),
                                      ctypes.POINTER(_f64), ctypes.POINTER(_f64), _vp]),
    "qmlh_woodbury_capacitance_f64": (ctypes.c_int, [_vp, _i64, _i64, _i64, _vp, _i64, _vp, _vp,
                                                     _sz, _vp]),
    "qmlh_gaussian_lowrank": (ctypes.c_int, [_vp, _i64, _i64, _i64, _f64, _vp, _vp, _i64, _vp]),
    "qmlh_lowrank_hybrid_workspace": (_sz, [_i64]),
    "qmlh_lowrank_hybrid": (ctypes.c_int, [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _i64, _i64, _vp,
                                           _sz, _vp]),
    "qmlh_trsm_f32l_workspace": (_sz, [_i64]),
    "qmlh_trsm_f32l": (ctypes.c_int, [_vp, _i64, _i64, _vp, _i64, _i64, _vp, _sz, _vp]),
    "qmlh_add_diagonal": (ctypes.c_int, [_vp, _i64, _i64, _f64, _vp]),
    "qmlh_predict_workspace": (_sz, [_i64, _i64]),
    "qmlh_krr_predict": (ctypes.c_int, [_i32, _vp, _i64, _i64, _vp, _vp, _i64, _i64, _vp, _vp,
                                        _i64, _f64, _vp, _vp, _sz, _vp]),
    "qmlh_krr_predict_noalloc_capacity": (_sz, []),
    "qmlh_krr_predict_noalloc": (ctypes.c_int, [_i32, _vp, _i64, _i64, _vp, _vp, _i64, _i64, _vp,
                                                _vp, _i64, _f64, _vp, _vp]),
    "qmlh_krr_batched_workspace": (_sz, [_vp, _i32, _i32, _i32]),
    "qmlh_krr_batched_create": (ctypes.c_int, [_vp, _i32, _i32, _i32, _vp, _i64, _i64, _i64, _vp, _vp,
                                               _i64, _vp, _i64, _vp, _vp, _sz,
                                               ctypes.POINTER(_vp)]),
    "qmlh_krr_batched_run": (ctypes.c_int, [_vp, _vp]),
    "qmlh_krr_batched_capture": (ctypes.c_int, [_vp, _vp]),
    "qmlh_krr_batched_launch": (ctypes.c_int, [_vp, _vp]),
    "qmlh_krr_batched_launches": (_i64, [_vp]),
    "qmlh_krr_batched_factor": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_i64),
                                               ctypes.POINTER(_i64)]),
    "qmlh_krr_batched_destroy": (None, [_vp]),
    "qmlh_timer_enable": (ctypes.c_int, [_i32]),
    "qmlh_timer_read": (ctypes.c_int, [_i32, ctypes.POINTER(_f64), ctypes.POINTER(_i64),
                                       ctypes.POINTER(_f64)]),
}

TIMER_CLASSES = {"syrk": 0, "gemm": 1, "gauss": 2, "pair": 3, "predict": 4, "potrf": 5,
                 "potrs": 6, "cm": 7, "local": 8, "pcg": 9}


def timer_enable(on: bool = True) -> None:
    check(load_library().qmlh_timer_enable(1 if on else 0))


def timer_read(name: str):
    """(total_ms, launches, work) of a timing class since timer_enable(True)."""
    ms, n, w = _f64(), _i64(), _f64()
    check(load_library().qmlh_timer_read(TIMER_CLASSES[name], ctypes.byref(ms), ctypes.byref(n),
                                         ctypes.byref(w)))
    return ms.value, n.value, w.value


def exported_symbols():
    return sorted(_SIGNATURES)


def load_library(path: str = LIB_PATH):
    """dlopen libqmlhip.so and attach signatures (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise QmlhError(
                    f"libqmlhip.so not found at {path}; build it with python -m qml_amd.build")
            L = ctypes.CDLL(path)
            for name, (res, args) in _SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


def lib():
    """The library, initialised on the current device (requires a visible GPU)."""
    import torch

    L = load_library()
    if not torch.cuda.is_available():
        raise QmlhError("no HIP device visible: qml_amd runs only on the GPU")
    dev = torch.cuda.current_device()
    if dev not in _inited_devices:
        with _lock:
            if dev not in _inited_devices:
                check(L.qmlh_init())
                _inited_devices.add(dev)
    return L


def check(rc: int) -> None:
    if rc != QMLH_OK:
        msg = load_library().qmlh_last_error().decode(errors="replace")
        raise QmlhError(f"libqmlhip status {rc}: {msg}")


def stream_handle(stream=None) -> int:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def host_int(t, stream=None) -> int:
    """Read a one-element device tensor after waiting for ``stream`` with the GIL released
    (a bare ``.item()`` holds the GIL for the whole wait, stalling other host threads)."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    s.synchronize()
    return int(t.item())


def ptr(t) -> int:
    return t.data_ptr() if t is not None else 0
