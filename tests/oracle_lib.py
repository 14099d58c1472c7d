"""ctypes binding of the CPU parity oracle (oracle/liboracle_flct.so).

Test infrastructure: only tests/, the smoke check and the CPU-baseline leg
of bench.py load this.  The oracle restates the reference's algorithm; it is
the checker, never the thing measured as the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle_flct.so")

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_longlong)
_lib = None


def _ptr(a):
    return a.ctypes.data_as(_dp)


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(LIB_PATH)
    return _lib


class using:
    """Temporarily route every oracle call to another build of the oracle
    (used by the planted-mistake tests)."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _lib
        self.saved = _lib
        _lib = _bind(self.path)
        return self

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def _bind(path):
    L = ctypes.CDLL(path)
    i, d = ctypes.c_int, ctypes.c_double
    L.oracle_flct.argtypes = [_dp, _dp, i, i, d, d, d, d, i, i, i, i, i, i, d, i, i,
                              _dp, _dp, _dp]
    L.oracle_flct_pixels.argtypes = [_dp, _dp, i, i, d, d, d, d, i, i, i, i, i,
                                     d, i, i, _lp, ctypes.c_longlong, _dp, _dp, _dp]
    L.oracle_cross_cor.argtypes = [_dp, _dp, i, i, i, d, _dp, _dp, _dp]
    L.oracle_peak_fit.argtypes = [_dp, i, i, _dp, _dp]
    L.oracle_peak_hires.argtypes = [_dp, i, i, _dp, _dp]
    L.oracle_interpcc2d.argtypes = [_dp, d, i, i, _dp, i, _dp, i, _dp]
    L.oracle_fft2.argtypes = [_dp, i, i, i]
    L.oracle_make_freq.argtypes = [_dp, i]
    L.oracle_box_nsize.argtypes = [d, i, i]
    L.oracle_warp_frac2d.argtypes = [_dp, _dp, _dp, _dp, i, i, i]
    L.oracle_warp_frac2d_rows.argtypes = [_dp, _dp, _dp, _dp, i, i, i, i, i]
    L.oracle_shift_frac2d.argtypes = [_dp, d, d, _dp, i, i, i]
    L.oracle_resolve_thresh.argtypes = [_dp, _dp, i, d, i, _dp]
    L.oracle_bias_pass.argtypes = [_dp, _dp, i, i, d, i, d, i, i, i, d, d, _dp, _dp]
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def flct(f1, f2, deltat, deltas, sigma, thresh=0.0, thresh_abs=False, skip=0,
         poff=0, qoff=0, interp=False, kr=None, bias=False, hires=False):
    """Whole flct run on C-space arrays; returns (vx, vy, vm), 1 x 1 for
    sigma <= 0."""
    f1, f2 = _f64(f1), _f64(f2)
    nx, ny = f1.shape
    shape = (nx, ny) if sigma > 0 else (1, 1)
    vx, vy, vm = (np.empty(shape) for _ in range(3))
    rc = lib().oracle_flct(_ptr(f1), _ptr(f2), nx, ny, deltat, deltas, sigma, thresh,
                           int(thresh_abs), skip, poff, qoff, int(interp),
                           int(kr is not None), float(kr or 0.0), int(bias), int(hires),
                           _ptr(vx), _ptr(vy), _ptr(vm))
    if rc:
        raise RuntimeError(f"oracle_flct status {rc}")
    return vx, vy, vm


def flct_pixels(f1, f2, deltat, deltas, sigma, pix, thresh=0.0, thresh_abs=False,
                skip=0, poff=0, qoff=0, kr=None, bias=False, hires=False):
    """Track only the C-space linear pixel indices ``pix`` (no skip fill)."""
    f1, f2 = _f64(f1), _f64(f2)
    nx, ny = f1.shape
    pix = np.ascontiguousarray(pix, dtype=np.int64)
    n = pix.size
    vx, vy, vm = (np.empty(n) for _ in range(3))
    rc = lib().oracle_flct_pixels(_ptr(f1), _ptr(f2), nx, ny, deltat, deltas, sigma,
                                  thresh, int(thresh_abs), skip, poff, qoff,
                                  int(kr is not None), float(kr or 0.0), int(bias),
                                  int(hires), pix.ctypes.data_as(_lp), n, _ptr(vx), _ptr(vy),
                                  _ptr(vm))
    if rc:
        raise RuntimeError(f"oracle_flct_pixels status {rc}")
    return vx, vy, vm


def cross_cor(g1, g2, kr=None):
    g1, g2 = _f64(g1), _f64(g2)
    nx, ny = g1.shape
    sx, sy = ctypes.c_double(), ctypes.c_double()
    surf = np.empty((nx, ny))
    lib().oracle_cross_cor(_ptr(g1), _ptr(g2), nx, ny, int(kr is not None),
                           float(kr or 0.0), ctypes.byref(sx), ctypes.byref(sy),
                           _ptr(surf))
    return sx.value, sy.value, surf


def peak_fit(surf):
    surf = _f64(surf)
    sx, sy = ctypes.c_double(), ctypes.c_double()
    lib().oracle_peak_fit(_ptr(surf), surf.shape[0], surf.shape[1], ctypes.byref(sx),
                          ctypes.byref(sy))
    return sx.value, sy.value


def peak_hires(surf):
    surf = _f64(surf)
    sx, sy = ctypes.c_double(), ctypes.c_double()
    lib().oracle_peak_hires(_ptr(surf), surf.shape[0], surf.shape[1], ctypes.byref(sx),
                            ctypes.byref(sy))
    return sx.value, sy.value


def interpcc2d(fdata, xmiss, xwant, ywant):
    fdata = _f64(fdata)
    xwant, ywant = _f64(xwant), _f64(ywant)
    out = np.empty((xwant.size, ywant.size))
    lib().oracle_interpcc2d(_ptr(fdata), xmiss, fdata.shape[0], fdata.shape[1],
                            _ptr(xwant), xwant.size, _ptr(ywant), ywant.size, _ptr(out))
    return out


def fft2(a, sign=-1):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    buf = a.view(np.float64).copy()
    lib().oracle_fft2(_ptr(buf), a.shape[0], a.shape[1], sign)
    return buf.view(np.complex128).reshape(a.shape)


def make_freq(n):
    k = np.empty(n)
    lib().oracle_make_freq(_ptr(k), n)
    return k


def box_nsize(sigma, nx, ny):
    return lib().oracle_box_nsize(float(sigma), nx, ny)


def warp_frac2d(arr, delx, dely, transp=1, rows=None):
    arr, delx, dely = _f64(arr), _f64(delx), _f64(dely)
    nx, ny = arr.shape
    out = np.full((nx, ny), np.nan)
    r0, r1 = (0, nx) if rows is None else rows
    lib().oracle_warp_frac2d_rows(_ptr(arr), _ptr(delx), _ptr(dely), _ptr(out), nx, ny,
                                  transp, r0, r1)
    return out


def shift_frac2d(arr, delx, dely, transp=1):
    arr = _f64(arr)
    out = np.empty_like(arr)
    lib().oracle_shift_frac2d(_ptr(arr), float(delx), float(dely), _ptr(out),
                              arr.shape[0], arr.shape[1], transp)
    return out


def bias_pass(f1, f2, sigma, i, j, d, kr=None, hires=False):
    """The -b second pass alone at pixel (i, j) with the offset d = (rows,
    columns) in surface units; returns the surface shifts (rows, columns)."""
    f1, f2 = _f64(f1), _f64(f2)
    sx, sy = ctypes.c_double(), ctypes.c_double()
    rc = lib().oracle_bias_pass(_ptr(f1), _ptr(f2), f1.shape[0], f1.shape[1], float(sigma),
                                int(kr is not None), float(kr or 0.0), int(hires), int(i), int(j),
                                float(d[0]), float(d[1]), ctypes.byref(sx), ctypes.byref(sy))
    if rc:
        raise RuntimeError(f"oracle_bias_pass status {rc}")
    return sx.value, sy.value


def resolve_thresh(f1, f2, thresh, thresh_abs=False):
    f1, f2 = _f64(f1), _f64(f2)
    out = ctypes.c_double()
    lib().oracle_resolve_thresh(_ptr(f1), _ptr(f2), f1.size, thresh, int(thresh_abs),
                                ctypes.byref(out))
    return out.value
