"""Host-side checks that need no GPU: vcimage format, CLI sub-grammars,
the C ABI's exported symbols and its host-only option validation."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HDR = os.path.join(ROOT, "include", "flct_hip.h")
LIB = os.path.join(ROOT, "pyflct_amd", "libflct_hip.so")


def declared_functions():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(flct_hip_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (flct_hip_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    from pyflct_amd import _lib
    assert sorted(_lib.EXPORTS) == declared_functions()


def test_library_loads_and_strerror():
    from pyflct_amd import _lib
    L = _lib.lib()
    assert L.flct_hip_strerror(0) == b"ok"
    assert b"skip" in L.flct_hip_strerror(_lib.FLCT_HIP_E_SKIP_NONPOS)


def _check(nx, ny, **kw):
    from pyflct_amd import _lib, api
    o = api.Context.make_opts(**kw)
    nbox = ctypes.c_int(-1)
    rc = _lib.lib().flct_hip_check_opts(nx, ny, ctypes.byref(o), ctypes.byref(nbox))
    return rc, nbox.value


def test_option_checks_match_reference_exit_sites():
    from pyflct_amd import _lib as L
    assert _check(128, 128, deltat=1, deltas=1, sigma=8) == (0, 32)
    assert _check(4096, 4096, deltat=1, deltas=1, sigma=15) == (0, 64)
    assert _check(128, 128, deltat=1, deltas=1, sigma=8, skip=-2)[0] == L.FLCT_HIP_E_SKIP_NONPOS
    assert _check(128, 128, deltat=1, deltas=1, sigma=8, skip=4, poff=4)[0] == L.FLCT_HIP_E_PQ_RANGE
    assert _check(128, 128, deltat=1, deltas=1, sigma=8, skip=4, qoff=-4)[0] == L.FLCT_HIP_E_PQ_RANGE
    assert _check(20, 128, deltat=1, deltas=1, sigma=8, skip=20)[0] == L.FLCT_HIP_E_SKIP_TOO_BIG
    assert _check(5, 5, deltat=1, deltas=1, sigma=8)[0] == L.FLCT_HIP_E_BOX
    # sigma <= 0 correlates the whole image (no box, flct.c:357-365, 567-586)
    assert _check(64, 64, deltat=1, deltas=1, sigma=0) == (0, 0)
    assert _check(7, 9, deltat=1, deltas=1, sigma=0, skip=3)[0] == 0
    assert _check(64, 64, deltat=1, deltas=1, sigma=-3) == (0, 0)
    assert _check(64, 64, deltat=1, deltas=1, sigma=0, skip=64)[0] == L.FLCT_HIP_E_SKIP_TOO_BIG
    # boxes above 96 run the plane kernel up to 1024 (sigma 24 gives 100,
    # 30 gives 128, 50 gives 212); larger boxes are refused
    assert _check(400, 400, deltat=1, deltas=1, sigma=24) == (0, 100)
    assert _check(400, 400, deltat=1, deltas=1, sigma=30) == (0, 128)
    assert _check(400, 400, deltat=1, deltas=1, sigma=50) == (0, 212)
    assert _check(4000, 4000, deltat=1, deltas=1, sigma=239) == (0, 1024)
    assert _check(4000, 4000, deltat=1, deltas=1, sigma=240)[0] == L.FLCT_HIP_E_UNSUPPORTED  # 1028


def test_skip_and_threshold_grammar():
    from pyflct_amd import api
    assert api.parse_skip("10p5q5") == (10, 5, 5, False)
    assert api.parse_skip("5i") == (5, 0, 0, True)
    assert api.parse_skip("4p-1q2i") == (4, -1, 2, True)
    assert api.parse_skip("6q3") == (6, 0, 3, False)
    with pytest.raises(api.FlctHipError):
        api.parse_skip("0")
    with pytest.raises(api.FlctHipError):
        api.parse_skip("4p4")
    assert api.parse_thresh("0.1") == (0.1, False)
    assert api.parse_thresh("0.1a") == (0.1, True)
    assert api.parse_thresh("25") == (25.0, False)


def test_vcimage_round_trip_and_header(tmp_path, golden):
    from pyflct_amd import vcimage
    p = tmp_path / "x.dat"
    a = np.arange(12.0).reshape(3, 4)
    b = -a
    vcimage.write(p, a, b)
    raw = open(p, "rb").read()
    assert raw[:4] == bytes.fromhex("7f5f89a9")
    assert int.from_bytes(raw[4:8], "big") == 4  # nx_idl = C-space columns
    assert int.from_bytes(raw[8:12], "big") == 3
    r = vcimage.read2(p)
    assert np.array_equal(r[0], a) and np.array_equal(r[1], b)
    assert golden["f1"].shape == (101, 201)  # IDL (201, 101) -> C (101, 201)
    bad = tmp_path / "bad.dat"
    bad.write_bytes(b"\0" * 16)
    with pytest.raises(vcimage.VcimageError):
        vcimage.read(bad)
