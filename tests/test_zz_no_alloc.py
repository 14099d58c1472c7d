"""flct_hip_track_device_noalloc (GPU): the same tracking with every scratch
buffer and table in the library's static device array instead of device
allocations.  Called through the C ABI; its outputs must equal
flct_hip_track_device's on the same inputs (same kernels and tables, so bit
for bit, which is within the FP64 form's 1e-8 px), on consecutive calls of
decreasing size that reuse the array, and it must refuse what does not fit."""
import numpy as np
import pytest

from conftest import smooth_pair
from pyflct_amd import _lib, api, synth
from pyflct_amd._lib import FlctHipError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nctx():
    c = api.Context(0)
    yield c
    c.close()


def _both(c, f1, f2, sigma, **kw):
    """Outputs of flct_hip_track_device and of the no-allocation entry point
    on the same device inputs."""
    nx, ny = f1.shape
    d = [c.malloc(f1.nbytes) for _ in range(8)]
    try:
        c.upload(d[0], f1)
        c.upload(d[1], f2)
        o = c.make_opts(1.0, 1.0, sigma, **kw)
        c.track_device(d[0], d[1], nx, ny, d[2], d[3], d[4], o)
        ref = [c.download(np.empty((nx, ny)), d[k]) for k in (2, 3, 4)]
        st_ref = c.last_stats()
        c.track_device_noalloc(d[0], d[1], nx, ny, d[5], d[6], d[7], o)
        got = [c.download(np.empty((nx, ny)), d[k]) for k in (5, 6, 7)]
        st = c.last_stats()
        assert st["tracked"] == st_ref["tracked"] and st["refined"] == st_ref["refined"]
        return ref, got
    finally:
        for p in d:
            c.free(p)


def _same(ref, got, where):
    for k in range(3):
        assert np.array_equal(ref[k], got[k]), (where, k)


# consecutive calls, each smaller than the one before (the second and later
# reuse the array the first filled): FP64 full disk, split-BF16 (row-segment
# kernel with its edge list and FP64 re-evaluation list), FP32 with
# threshold, skip fill and filter, and a box above 96 (block planes)
CASES = [
    ("fp64 1024^2 box 64", 1024, 15.0, {}),
    ("split-bf16 768^2 box 64", 768, 15.0, {"precision": "split_bf16"}),
    ("fp32 400^2 box 40 -t -s -k", 400, 10.0,
     {"precision": "fp32", "thresh": 0.2, "skip": 3, "poff": 1, "interp": True, "kr": 0.5}),
    ("fp64 210^2 box 100 -b", 210, 24.0, {"bias": True}),
    ("fp64 96^2 box 24 -h", 96, 6.0, {"hires": True}),
    ("fp32 512^2 box 64 -k (row-shared, FP32 row tasks)", 512, 15.0, {"precision": "fp32", "kr": 0.5}),
]


def test_noalloc_matches_track_device(nctx):
    for where, n, sigma, kw in CASES:
        f1, f2, _ = synth.cfg4(n=n)
        ref, got = _both(nctx, f1, f2, sigma, **kw)
        _same(ref, got, where)


def test_noalloc_two_calls_second_smaller(nctx):
    rng = np.random.default_rng(77)
    a, b = smooth_pair(rng, 640, 700, shift=(0.8, -0.3), width=2.5)
    c_, d_ = smooth_pair(rng, 150, 170, shift=(-1.1, 0.6), width=2.0)
    r1, g1 = _both(nctx, a, b, 12.0, thresh=0.3)
    r2, g2 = _both(nctx, c_, d_, 5.0, skip=4, interp=True)
    _same(r1, g1, "first call")
    _same(r2, g2, "second, smaller call")


def test_noalloc_refuses_what_does_not_fit(nctx):
    n = 4096  # row sums and means alone need 4 x 4096 x 4033 doubles > 256 MiB
    d = [nctx.malloc(8 * n * n) for _ in range(5)]
    try:
        with pytest.raises(FlctHipError) as e:
            nctx.track_device_noalloc(d[0], d[1], n, n, d[2], d[3], d[4],
                                      nctx.make_opts(1.0, 1.0, 15.0))
        assert e.value.status == _lib.FLCT_HIP_E_NOSCRATCH
        with pytest.raises(FlctHipError) as e:
            nctx.track_device_noalloc(d[0], d[1], 64, 64, d[2], d[3], d[4],
                                      nctx.make_opts(1.0, 1.0, 0.0))
        assert e.value.status == _lib.FLCT_HIP_E_UNSUPPORTED
    finally:
        for p in d:
            nctx.free(p)
    # the context still works afterwards
    f1, f2, _ = synth.cfg4(n=128)
    ref, got = _both(nctx, f1, f2, 8.0)
    _same(ref, got, "after the refusals")
