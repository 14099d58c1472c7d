"""Python interface of the MI355X FLCT tracker and Fourier shifter.

Mirrors the reference programs' operator interface:

* ``flct(f1, f2, deltat, deltas, sigma, ...)`` is ``flct ifile ofile deltat
  deltas sigma [-t thr[a]] [-k kr] [-s N[pP][qQ][i]] [-b]`` (flct.c:304,
  flct.c:370-487) on in-memory C-space arrays, returning ``(vx, vy, vm)``;
* ``warp(arr, delx, dely)`` is ``warp_frac2d`` (warp.c:883-1011);
* ``shift_frac2d(arr, delx, dely)`` is ``shift_frac2d`` (warp.c:757-881).

Option errors raise :class:`FlctHipError` whose ``status`` names the
reference's exit(1) site (see include/flct_hip.h).  Everything runs on the GPU
through libflct_hip.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import FlctHipError, Opts, check, f64c, ptr

__all__ = ["Context", "flct", "warp", "shift_frac2d", "parse_skip", "parse_thresh",
           "FlctHipError"]


def parse_thresh(text):
    """``-t`` value: a trailing 'a' forces absolute units (flct.c:392-404).

    Returns ``(thresh, thresh_abs)``.
    """
    s = str(text)
    absolute = "a" in s
    if absolute:
        s = s[: s.index("a")]
    return _atof(s), absolute


def parse_skip(text):
    """``-s N[pP][qQ][i]`` sub-grammar (flct.c:405-456), parsed right to left.

    Returns ``(skip, poff, qoff, interp)``; raises FlctHipError for the
    reference's fatal cases (skip <= 0, |p| or |q| >= skip).
    """
    s = str(text)
    interp = False
    p = q = 0
    if "i" in s:
        interp = True
        s = s[: s.index("i")]
    if "q" in s:
        k = s.index("q")
        q = _atoi(s[k + 1:])
        s = s[:k]
    if "p" in s:
        k = s.index("p")
        p = _atoi(s[k + 1:])
        s = s[:k]
    skip = _atoi(s)
    if skip <= 0:
        raise FlctHipError(_lib.FLCT_HIP_E_SKIP_NONPOS, "parse_skip")
    if abs(p) >= skip or abs(q) >= skip:
        raise FlctHipError(_lib.FLCT_HIP_E_PQ_RANGE, "parse_skip")
    return skip, p, q, interp


def _atoi(s):
    """C atoi: optional sign then leading digits, 0 when none."""
    s = s.strip()
    k = 0
    if k < len(s) and s[k] in "+-":
        k += 1
    while k < len(s) and s[k].isdigit():
        k += 1
    try:
        return int(s[:k])
    except ValueError:
        return 0


def _atof(s):
    """C atof: the longest leading float literal, 0.0 when none."""
    s = s.strip()
    for end in range(len(s), 0, -1):
        try:
            return float(s[:end])
        except ValueError:
            continue
    return 0.0


class Context:
    """One GPU, one HIP stream, device buffers reused across calls."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        check(_lib.lib().flct_hip_create(ctypes.byref(self._h), int(device)), "flct_hip_create")

    def close(self):
        if self._h:
            _lib.lib().flct_hip_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ------------------------------------------------------------ options
    @staticmethod
    def make_opts(deltat, deltas, sigma, thresh=0.0, thresh_abs=False, kr=None, skip=0,
                  poff=0, qoff=0, interp=False, bias=False, precision="fp64", quiet=True,
                  hires=False):
        o = Opts()
        _lib.lib().flct_hip_opts_default(ctypes.byref(o))
        o.deltat, o.deltas, o.sigma = float(deltat), float(deltas), float(sigma)
        o.thresh, o.thresh_abs = float(thresh), int(bool(thresh_abs))
        o.skip, o.poff, o.qoff, o.interp = int(skip or 0), int(poff), int(qoff), int(bool(interp))
        o.filter = int(kr is not None)
        o.kr = float(kr or 0.0)
        o.bias = int(bool(bias))
        o.precision = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        o.quiet = int(bool(quiet))
        o.hires = int(bool(hires))
        return o

    @staticmethod
    def box_size(nx, ny, sigma):
        """Sliding-box size 2*nsize for an nx x ny image (flct.c:567-580)."""
        o = Context.make_opts(1.0, 1.0, sigma)
        nbox = ctypes.c_int(0)
        check(_lib.lib().flct_hip_check_opts(int(nx), int(ny), ctypes.byref(o), ctypes.byref(nbox)),
              "flct_hip_check_opts")
        return nbox.value

    # ------------------------------------------------------------ tracking
    def track(self, f1, f2, deltat, deltas, sigma, **kw):
        """Track on host arrays; returns (vx, vy, vm) float64 arrays.

        With ``sigma <= 0`` the two whole images are correlated once and the
        arrays are 1 x 1 (flct.c:357-365, 581-586).
        """
        f1, f2 = f64c(f1), f64c(f2)
        if f1.ndim != 2 or f1.shape != f2.shape:
            raise ValueError("f1 and f2 must be 2-D arrays of the same shape")
        nx, ny = f1.shape
        shape = (nx, ny) if sigma > 0 else (1, 1)
        vx, vy, vm = (np.empty(shape) for _ in range(3))
        o = self.make_opts(deltat, deltas, sigma, **kw)
        check(_lib.lib().flct_hip_track(self._h, ptr(f1), ptr(f2), nx, ny, ctypes.byref(o),
                                        ptr(vx), ptr(vy), ptr(vm)), "flct_hip_track")
        return vx, vy, vm

    def track_device(self, d_f1, d_f2, nx, ny, d_vx, d_vy, d_vm, opts):
        check(_lib.lib().flct_hip_track_device(self._h, d_f1, d_f2, nx, ny, ctypes.byref(opts),
                                               d_vx, d_vy, d_vm), "flct_hip_track_device")

    def track_device_noalloc(self, d_f1, d_f2, nx, ny, d_vx, d_vy, d_vm, opts):
        """flct_hip_track_device without device allocation: its scratch is the
        library's static 256 MiB device array (FLCT_HIP_E_NOSCRATCH beyond)."""
        check(_lib.lib().flct_hip_track_device_noalloc(self._h, d_f1, d_f2, nx, ny,
                                                       ctypes.byref(opts), d_vx, d_vy, d_vm),
              "flct_hip_track_device_noalloc")

    def track_batch(self, pairs, **common):
        """Track several image pairs in one graph launch (flct_hip_track_batch).

        ``pairs`` is a sequence of ``(f1, f2, deltat, deltas, sigma)`` or
        ``(f1, f2, deltat, deltas, sigma, options)`` tuples of host arrays;
        ``options`` (a dict) overrides ``common``.  Returns a list of
        ``(vx, vy, vm)``.  Buffers are allocated per call; use
        :meth:`track_batch_device` to reuse device arrays (and the captured
        graph) across calls.
        """
        bufs, descs = [], []
        try:
            for item in pairs:
                f1, f2, dt, ds, sigma = item[:5]
                kw = dict(common)
                if len(item) > 5:
                    kw.update(item[5])
                f1, f2 = f64c(f1), f64c(f2)
                if f1.ndim != 2 or f1.shape != f2.shape:
                    raise ValueError("f1 and f2 must be 2-D arrays of the same shape")
                nx, ny = f1.shape
                d = [self.malloc(f1.nbytes) for _ in range(5)]
                bufs.append(d)
                self.upload(d[0], f1)
                self.upload(d[1], f2)
                descs.append((d, nx, ny, self.make_opts(dt, ds, sigma, **kw)))
            self.track_batch_device([(d[0], d[1], d[2], d[3], d[4], nx, ny, o)
                                     for d, nx, ny, o in descs])
            out = []
            for d, nx, ny, _ in descs:
                out.append(tuple(self.download(np.empty((nx, ny)), d[k]) for k in (2, 3, 4)))
            return out
        finally:
            for d in bufs:
                for p in d:
                    self.free(p)

    def track_batch_device(self, pairs):
        """``pairs``: sequence of (d_f1, d_f2, d_vx, d_vy, d_vm, nx, ny, opts)
        with device pointers and :class:`Opts`.  Returns per-pair
        ``{"tracked", "refined"}``; raises FlctHipError with the first failing
        pair's status."""
        arr = (_lib.Pair * len(pairs))()
        for k, (f1, f2, vx, vy, vm, nx, ny, o) in enumerate(pairs):
            arr[k].f1, arr[k].f2 = _vp(f1), _vp(f2)
            arr[k].vx, arr[k].vy, arr[k].vm = _vp(vx), _vp(vy), _vp(vm)
            arr[k].nx, arr[k].ny = int(nx), int(ny)
            arr[k].opts = o
        rc = _lib.lib().flct_hip_track_batch(self._h, arr, len(pairs))
        if rc != _lib.FLCT_HIP_OK:
            bad = [k for k in range(len(pairs)) if arr[k].status != _lib.FLCT_HIP_OK]
            raise FlctHipError(rc, f"flct_hip_track_batch (pairs {bad})")
        return [{"tracked": arr[k].tracked, "refined": arr[k].refined} for k in range(len(pairs))]

    def autotune(self, d_f1, d_f2, nx, ny, opts):
        """Time every kernel path that meets opts' precision form on these
        device inputs, then the winner's runtime shapes (slots per block,
        segment width), and keep the fastest for that form and box size."""
        check(_lib.lib().flct_hip_autotune(self._h, d_f1, d_f2, int(nx), int(ny),
                                           ctypes.byref(opts)), "flct_hip_autotune")
        return self.last_path()

    def set_path(self, precision, nbox, path):
        """Pin the kernel path of a precision form and box size
        (flct_hip_set_path); ``path=None`` restores the default."""
        prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        check(_lib.lib().flct_hip_set_path(self._h, prec, int(nbox),
                                           None if path is None else path.encode()),
              "flct_hip_set_path")

    def set_config(self, precision, nbox, path, slots=0, segpx=0):
        """Pin a kernel path and its runtime shape (flct_hip_set_config):
        slots per block (0: compiled count) and pixels per row segment
        (0: 64); ``path=None`` restores the default."""
        prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        check(_lib.lib().flct_hip_set_config(self._h, prec, int(nbox),
                                             None if path is None else path.encode(),
                                             int(slots), int(segpx)),
              "flct_hip_set_config")

    def last_shape(self):
        """(slots per block, pixels per segment) of the last call's main
        kernel; 0 stands for the compiled shape."""
        a, b = ctypes.c_int(), ctypes.c_int()
        check(_lib.lib().flct_hip_last_shape(self._h, ctypes.byref(a), ctypes.byref(b)),
              "flct_hip_last_shape")
        return a.value, b.value

    def last_path(self):
        """Kernel path of the last track call ("fp64-fft", "fp32-fft",
        "fp64-rowseg", "fp32-rowseg", "bf16-segment", "bf16-pixel",
        "fp64-plane")."""
        return _lib.lib().flct_hip_last_path(self._h).decode()

    def last_stats(self):
        n = ctypes.c_longlong()
        k = ctypes.c_double()
        t = ctypes.c_double()
        check(_lib.lib().flct_hip_last_stats(self._h, ctypes.byref(n), ctypes.byref(k),
                                             ctypes.byref(t)))
        rf = ctypes.c_longlong()
        check(_lib.lib().flct_hip_last_refined(self._h, ctypes.byref(rf)))
        mm = ctypes.c_double()
        rm = ctypes.c_double()
        check(_lib.lib().flct_hip_last_kernel_ms(self._h, ctypes.byref(mm), ctypes.byref(rm)))
        return {"tracked": n.value, "kernel_ms": k.value, "total_ms": t.value,
                "refined": rf.value, "main_ms": mm.value, "refine_ms": rm.value}

    # ------------------------------------------------------------ warp
    def warp(self, arr, delx, dely, transp=True):
        arr, delx, dely = f64c(arr), f64c(delx), f64c(dely)
        if delx.shape != arr.shape or dely.shape != arr.shape:
            raise FlctHipError(_lib.FLCT_HIP_E_SHAPE, "warp")
        nx, ny = arr.shape
        out = np.empty((nx, ny))
        check(_lib.lib().flct_hip_warp(self._h, ptr(arr), ptr(delx), ptr(dely), ptr(out), nx,
                                       ny, int(bool(transp)), 1), "flct_hip_warp")
        return out

    WARP_METHODS = {"auto": 0, "direct": 1, "nufft": 2}

    def set_warp_method(self, method):
        """"auto" (NUFFT for 16 <= nx, ny <= 4096), "direct" (the reference's
        per-pixel sum) or "nufft" (flct_hip_set_warp_method)."""
        check(_lib.lib().flct_hip_set_warp_method(self._h, self.WARP_METHODS[method]),
              "flct_hip_set_warp_method")

    def last_warp_method(self):
        m = _lib.lib().flct_hip_last_warp_method(self._h)
        return {1: "direct", 2: "nufft"}.get(m, "none")

    def warp_device(self, d_arr, d_delx, d_dely, d_out, nx, ny, transp=True):
        """warp_frac2d on device arrays (flct_hip_warp_device)."""
        check(_lib.lib().flct_hip_warp_device(self._h, d_arr, d_delx, d_dely, d_out, int(nx),
                                              int(ny), int(bool(transp))), "flct_hip_warp_device")

    def shift(self, arr, delx, dely, transp=True):
        arr = f64c(arr)
        nx, ny = arr.shape
        out = np.empty((nx, ny))
        check(_lib.lib().flct_hip_shift(self._h, ptr(arr), float(delx), float(dely), ptr(out),
                                        nx, ny, int(bool(transp)), 1), "flct_hip_shift")
        return out

    # ------------------------------------------------------------ memory
    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        check(_lib.lib().flct_hip_malloc(self._h, ctypes.byref(p), int(nbytes)), "flct_hip_malloc")
        return p

    def free(self, p):
        check(_lib.lib().flct_hip_free(self._h, p), "flct_hip_free")

    def upload(self, dptr, a):
        a = np.ascontiguousarray(a)
        check(_lib.lib().flct_hip_upload(self._h, dptr, a.ctypes.data_as(ctypes.c_void_p),
                                         a.nbytes), "flct_hip_upload")

    def download(self, a, dptr):
        check(_lib.lib().flct_hip_download(self._h, a.ctypes.data_as(ctypes.c_void_p), dptr,
                                           a.nbytes), "flct_hip_download")
        return a


def _vp(p):
    return p.value if isinstance(p, ctypes.c_void_p) else p


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


def flct(f1, f2, deltat, deltas, sigma, thresh=0.0, thresh_abs=False, kr=None, skip=0, poff=0,
         qoff=0, interp=False, bias=False, precision="fp64", ctx=None):
    """Local correlation tracking of f2 against f1 (C-space arrays).

    Returns (vx, vy, vm) with vx along the column index j (IDL x) and vy
    along the row index i (IDL y), in units of deltas/deltat (flct.c:901-905).
    """
    ctx = ctx or default_context()
    return ctx.track(f1, f2, deltat, deltas, sigma, thresh=thresh, thresh_abs=thresh_abs, kr=kr,
                     skip=skip, poff=poff, qoff=qoff, interp=interp, bias=bias,
                     precision=precision)


def warp(arr, delx, dely, transp=True, ctx=None):
    """warp_frac2d: evaluate arr's Fourier series at displaced coordinates."""
    return (ctx or default_context()).warp(arr, delx, dely, transp)


def shift_frac2d(arr, delx, dely, transp=True, ctx=None):
    """Uniform sub-pixel Fourier shift (warp.c:757-881)."""
    return (ctx or default_context()).shift(arr, delx, dely, transp)
