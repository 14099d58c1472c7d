"""ctypes binding of libflct_hip.so (include/flct_hip.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C pyflct_amd/csrc``).  Loading fails loudly when it is missing: there
is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FLCT_HIP_LIB selects another build of the library (tuning variants)
LIB_PATH = os.environ.get("FLCT_HIP_LIB") or os.path.join(HERE, "libflct_hip.so")

FLCT_HIP_OK = 0
FLCT_HIP_E_ARG = 1
FLCT_HIP_E_SKIP_NONPOS = 2
FLCT_HIP_E_PQ_RANGE = 3
FLCT_HIP_E_SKIP_TOO_BIG = 4
FLCT_HIP_E_BOX = 5
FLCT_HIP_E_UNSUPPORTED = 6
FLCT_HIP_E_NODEVICE = 7
FLCT_HIP_E_ALLOC = 8
FLCT_HIP_E_HIP = 9
FLCT_HIP_E_SHAPE = 10
FLCT_HIP_E_NOSCRATCH = 11

FP64 = 0
FP32 = 1
SPLIT_BF16 = 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "split_bf16": SPLIT_BF16, "bf16x3": SPLIT_BF16}

# entry points declared in include/flct_hip.h
EXPORTS = (
    "flct_hip_opts_default",
    "flct_hip_check_opts",
    "flct_hip_strerror",
    "flct_hip_create",
    "flct_hip_destroy",
    "flct_hip_track",
    "flct_hip_track_device",
    "flct_hip_track_device_noalloc",
    "flct_hip_track_batch",
    "flct_hip_last_stats",
    "flct_hip_last_refined",
    "flct_hip_last_kernel_ms",
    "flct_hip_last_path",
    "flct_hip_autotune",
    "flct_hip_set_path",
    "flct_hip_set_config",
    "flct_hip_last_shape",
    "flct_hip_warp",
    "flct_hip_warp_device",
    "flct_hip_set_warp_method",
    "flct_hip_last_warp_method",
    "flct_hip_shift",
    "flct_hip_malloc",
    "flct_hip_free",
    "flct_hip_upload",
    "flct_hip_download",
)


class FlctHipError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = lib().flct_hip_strerror(status).decode()
        super().__init__(f"{where}: {msg} (status {status})" if where else msg)


class Opts(ctypes.Structure):
    _fields_ = [
        ("deltat", ctypes.c_double),
        ("deltas", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("thresh", ctypes.c_double),
        ("thresh_abs", ctypes.c_int),
        ("skip", ctypes.c_int),
        ("poff", ctypes.c_int),
        ("qoff", ctypes.c_int),
        ("interp", ctypes.c_int),
        ("filter", ctypes.c_int),
        ("kr", ctypes.c_double),
        ("bias", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("quiet", ctypes.c_int),
        ("hires", ctypes.c_int),
    ]


class Pair(ctypes.Structure):
    """flct_hip_pair: one image pair of a batch (device arrays)."""
    _fields_ = [
        ("f1", ctypes.c_void_p),
        ("f2", ctypes.c_void_p),
        ("vx", ctypes.c_void_p),
        ("vy", ctypes.c_void_p),
        ("vm", ctypes.c_void_p),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("opts", Opts),
        ("tracked", ctypes.c_longlong),
        ("refined", ctypes.c_longlong),
        ("status", ctypes.c_int),
    ]


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    """Load libflct_hip.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.flct_hip_opts_default.argtypes = [ctypes.POINTER(Opts)]
        L.flct_hip_opts_default.restype = None
        L.flct_hip_strerror.argtypes = [i]
        L.flct_hip_strerror.restype = ctypes.c_char_p
        L.flct_hip_check_opts.argtypes = [i, i, ctypes.POINTER(Opts), ctypes.POINTER(i)]
        L.flct_hip_create.argtypes = [ctypes.POINTER(vp), i]
        L.flct_hip_destroy.argtypes = [vp]
        L.flct_hip_track.argtypes = [vp, _dp, _dp, i, i, ctypes.POINTER(Opts), _dp, _dp, _dp]
        L.flct_hip_track_device.argtypes = [vp, vp, vp, i, i, ctypes.POINTER(Opts), vp, vp, vp]
        L.flct_hip_track_device_noalloc.argtypes = [vp, vp, vp, i, i, ctypes.POINTER(Opts), vp, vp,
                                                    vp]
        L.flct_hip_track_batch.argtypes = [vp, ctypes.POINTER(Pair), i]
        L.flct_hip_last_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong),
                                          ctypes.POINTER(d), ctypes.POINTER(d)]
        L.flct_hip_last_kernel_ms.argtypes = [vp, ctypes.POINTER(d), ctypes.POINTER(d)]
        L.flct_hip_last_path.argtypes = [vp]
        L.flct_hip_last_path.restype = ctypes.c_char_p
        L.flct_hip_autotune.argtypes = [vp, vp, vp, i, i, ctypes.POINTER(Opts)]
        L.flct_hip_set_path.argtypes = [vp, i, i, ctypes.c_char_p]
        L.flct_hip_set_config.argtypes = [vp, i, i, ctypes.c_char_p, i, i]
        L.flct_hip_last_shape.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i)]
        L.flct_hip_warp.argtypes = [vp, _dp, _dp, _dp, _dp, i, i, i, i]
        L.flct_hip_warp_device.argtypes = [vp, vp, vp, vp, vp, i, i, i]
        L.flct_hip_set_warp_method.argtypes = [vp, i]
        L.flct_hip_last_warp_method.argtypes = [vp]
        L.flct_hip_shift.argtypes = [vp, _dp, d, d, _dp, i, i, i, i]
        L.flct_hip_malloc.argtypes = [vp, ctypes.POINTER(vp), ctypes.c_longlong]
        L.flct_hip_free.argtypes = [vp, vp]
        L.flct_hip_upload.argtypes = [vp, vp, vp, ctypes.c_longlong]
        L.flct_hip_download.argtypes = [vp, vp, vp, ctypes.c_longlong]
        for name in EXPORTS:
            if name not in ("flct_hip_opts_default", "flct_hip_strerror", "flct_hip_last_path"):
                getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def check(status, where=""):
    if status != FLCT_HIP_OK:
        raise FlctHipError(status, where)


def ptr(a):
    return a.ctypes.data_as(_dp)


def f64c(a):
    return np.ascontiguousarray(a, dtype=np.float64)
