"""MI355X-native FLCT local correlation tracker and Fourier warp.

The product path is libflct_hip.so (HIP kernels for gfx950 behind the C ABI in
include/flct_hip.h); this package is its Python binding plus vcimage I/O.
"""
import importlib

from . import vcimage  # noqa: F401

__version__ = "0.1.0"

_API_NAMES = ("flct", "warp", "shift_frac2d", "Context", "FlctHipError", "parse_skip",
              "parse_thresh")


def __getattr__(name):
    # the binding loads libflct_hip.so lazily so that vcimage works without it
    if name == "api":
        return importlib.import_module(".api", __name__)
    if name in _API_NAMES:
        return getattr(importlib.import_module(".api", __name__), name)
    raise AttributeError(name)
