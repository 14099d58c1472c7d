"""Readers and writers for the "vcimage" binary file format of flct/warp.

Format (flct.c:1158-1268 read2images, flct.c:2016-2109 write3images,
IDL-io-procedures/vcimage2out.pro:50-68, vcimage3in.pro:37-65):

* big-endian int32 magic 2136967593 (0x7F5F89A9);
* big-endian int32 ``nx_idl`` then ``ny_idl``;
* ``k`` big-endian float32 arrays of nx_idl * ny_idl values, IDL x fastest.

flct and warp work in "C-space" (transp = 1, flct.c:291): an IDL array
``a(nx_idl, ny_idl)`` is the C row-major array with ``nx = ny_idl`` rows and
``ny = nx_idl`` columns.  The functions here return and accept numpy arrays
of shape ``(nx, ny) = (ny_idl, nx_idl)`` in C order, which is exactly that
C-space layout; values are widened to float64 on read as flct does
(flct.c:1253-1257) and narrowed to float32 on write (flct.c:2079-2084).
"""
from __future__ import annotations

import numpy as np

MAGIC = 2136967593


class VcimageError(ValueError):
    """Raised for files that are not vcimage files (flct.c:1184-1188)."""


def read(path, count=None):
    """Read every array of a vcimage file.

    Returns a list of float64 arrays of shape ``(ny_idl, nx_idl)``.  ``count``
    (optional) checks the number of arrays the file must hold.
    """
    with open(path, "rb") as fh:
        head = np.frombuffer(fh.read(12), dtype=">i4")
        if head.size < 3 or int(head[0]) != MAGIC:
            raise VcimageError(f"{path}: input file is not a vel_ccor i/o file")
        nx_idl, ny_idl = int(head[1]), int(head[2])
        if nx_idl <= 0 or ny_idl <= 0:
            raise VcimageError(f"{path}: bad dimensions {nx_idl} x {ny_idl}")
        body = np.frombuffer(fh.read(), dtype=">f4")
    n = nx_idl * ny_idl
    if body.size % n:
        raise VcimageError(f"{path}: payload is not a whole number of arrays")
    k = body.size // n
    if count is not None and k < count:
        raise VcimageError(f"{path}: holds {k} arrays, {count} needed")
    return [body[a * n:(a + 1) * n].astype(np.float64).reshape(ny_idl, nx_idl)
            for a in range(k)]


def write(path, *arrays):
    """Write one or more same-shape arrays (C-space ``(nx, ny)``) as float32."""
    if not arrays:
        raise ValueError("nothing to write")
    shape = np.shape(arrays[0])
    if len(shape) != 2:
        raise ValueError("arrays must be 2-D")
    for a in arrays:
        if np.shape(a) != shape:
            raise ValueError("all arrays must have the same shape")
    nx, ny = shape
    with open(path, "wb") as fh:
        # header holds IDL dims: nx_idl = C ny, ny_idl = C nx (flct.c:2057-2061)
        fh.write(np.array([MAGIC, ny, nx], dtype=">i4").tobytes())
        for a in arrays:
            fh.write(np.ascontiguousarray(a, dtype=np.float64).astype(">f4").tobytes())


def read1(path):
    return read(path, 1)[0]


def read2(path):
    a = read(path, 2)
    return a[0], a[1]


def read3(path):
    a = read(path, 3)
    return a[0], a[1], a[2]
