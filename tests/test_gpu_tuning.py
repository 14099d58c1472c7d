"""Runtime kernel shapes and the autotuner (GPU).

The per-pixel FFT kernels (fp64-fft, fp32-fft) track each pixel the same way
whatever the number of pixels per block, so a shape change must leave every
output bit-identical.  The row-segment kernels (fp64-rowseg, bf16-segment)
share the box-row transforms of a segment: in FP64 about a segment
reference mean, so a narrower segment changes the rounding only and the
result stays within 1e-8 px of the oracle with the vm mask bit-exact; in
split-BF16 about per-column means, so the result does not change at all.
"""
import numpy as np
import pytest

from pyflct_amd import synth
from pyflct_amd._lib import FlctHipError
import oracle_lib

pytestmark = pytest.mark.gpu


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def _rms(a, b):
    return float(np.sqrt(np.mean((a[0] - b[0]) ** 2 + (a[1] - b[1]) ** 2)))


@pytest.mark.parametrize("form,path,sigma", [
    ("fp64", "fp64-fft", 7.0),    # box 28
    ("fp64", "fp64-fft", 15.0),   # box 64
    ("fp32", "fp32-fft", 5.0),    # box 20
    ("fp32", "fp32-fft", 19.0),   # box 80
])
def test_slots_per_block_bit_identical(form, path, sigma):
    from pyflct_amd import api
    c = api.Context(0)
    try:
        f1, f2, _ = synth.cfg4(n=288)
        nbox = c.box_size(288, 288, sigma)
        c.set_config(form, nbox, path)
        ref = c.track(f1, f2, 1.0, 1.0, sigma, precision=form)
        assert c.last_path() == path and c.last_shape() == (0, 0)
        ran = 0
        for slots in (1, 2, 3):
            try:
                c.set_config(form, nbox, path, slots=slots)
            except FlctHipError:
                continue  # more slots than the compiled count
            got = c.track(f1, f2, 1.0, 1.0, sigma, precision=form)
            assert c.last_path() == path and c.last_shape() == (slots, 0)
            assert _same(got, ref), (path, nbox, slots)
            ran += 1
        assert ran >= 1
    finally:
        c.close()


def test_rowseg_shapes_against_oracle():
    from pyflct_amd import api
    c = api.Context(0)
    try:
        f1, f2, _ = synth.cfg4(n=320)
        ref = c.track(f1, f2, 1.0, 1.0, 15.0)
        assert c.last_path() == "fp64-rowseg"
        rng = np.random.default_rng(5)
        pix = np.concatenate([rng.choice(320 * 320, 150, replace=False),
                              np.arange(0, 320), np.arange(319 * 320, 320 * 320)])
        ox, oy, _ = oracle_lib.flct_pixels(f1, f2, 1.0, 1.0, 15.0, pix)
        for slots, segpx in ((1, 0), (0, 32), (1, 16), (0, 8)):
            c.set_config("fp64", 64, "fp64-rowseg", slots=slots, segpx=segpx)
            got = c.track(f1, f2, 1.0, 1.0, 15.0)
            assert c.last_path() == "fp64-rowseg" and c.last_shape() == (slots, segpx)
            assert np.array_equal(got[2], ref[2])
            err = max(np.max(np.abs(got[0].ravel()[pix] - ox)), np.max(np.abs(got[1].ravel()[pix] - oy)))
            assert err <= 1e-8, (slots, segpx, err)
    finally:
        c.close()


@pytest.mark.parametrize("sigma", [7.0, 11.0, 15.0])  # boxes 28, 46 -> 44, 64
def test_bf16_segment_widths(sigma):
    """Every segment width gives the same split-BF16 result: the shared
    box-row transforms and the per-pixel mean correction do not depend on
    where a segment starts.  The 320^2 full-disk pair at box 64 is also the
    regression case of the mean correction, whose column vector was once
    formed in FP32 about the segment's first box mean and lost digits where
    that mean was far from the pixel's (0.042 px at rows 0..4, column 279)."""
    from pyflct_amd import api
    c = api.Context(0)
    try:
        f1, f2, _ = synth.cfg4(n=320)
        ref = c.track(f1, f2, 1.0, 1.0, sigma)
        nbox = c.box_size(320, 320, sigma)
        base = c.track(f1, f2, 1.0, 1.0, sigma, precision="split_bf16")
        assert c.last_path() == "bf16-segment"
        assert np.array_equal(base[2], ref[2])
        e = np.hypot(base[0] - ref[0], base[1] - ref[1])
        assert float(np.sqrt(np.mean(e ** 2))) <= 1e-4 and e.max() <= 5e-3, (e.max(),)
        for segpx in (48, 32, 16, 8):
            c.set_config("split_bf16", nbox, "bf16-segment", segpx=segpx)
            got = c.track(f1, f2, 1.0, 1.0, sigma, precision="split_bf16")
            assert c.last_shape() == (0, segpx)
            assert _same(got, base), segpx
    finally:
        c.close()


def test_config_rules():
    from pyflct_amd import api
    c = api.Context(0)
    try:
        for args in (("fp64", 64, "fp64-fft", 99, 0),     # more slots than compiled
                     ("fp32", 40, "fp32-fft", 0, 32),     # no segment axis
                     ("split_bf16", 40, "bf16-segment", 0, 4),
                     ("split_bf16", 40, "bf16-segment", 2, 0),  # no slot axis
                     ("fp64", 40, "fp64-rowseg", -1, 0),
                     ("fp64", 40, "bf16-segment", 0, 0)):  # outside the FP64 contract
            with pytest.raises(FlctHipError):
                c.set_config(*args[:3], slots=args[3], segpx=args[4])
        c.set_config("fp64", 40, "fp64-rowseg", slots=1, segpx=64)
        c.set_config("fp64", 40, None)
        f1, f2, _ = synth.cfg4(n=160)
        c.track(f1, f2, 1.0, 1.0, 9.0)
        assert c.last_shape() == (0, 0)
    finally:
        c.close()


def test_autotune_keeps_a_shape():
    """The tuner's configuration (path and shape) serves the later calls of
    that form and box size, within the form's contract."""
    from pyflct_amd import api
    c = api.Context(0)
    try:
        n = 448
        f1, f2, _ = synth.cfg4(n=n)
        d = [c.malloc(f1.nbytes) for _ in range(2)]
        try:
            c.upload(d[0], f1)
            c.upload(d[1], f2)
            for form, sigma in (("fp32", 9.0), ("split_bf16", 15.0), ("fp64", 12.0)):
                best = c.autotune(d[0], d[1], n, n, c.make_opts(1.0, 1.0, sigma, precision=form))
                shape = c.last_shape()
                got = c.track(f1, f2, 1.0, 1.0, sigma, precision=form)
                assert (c.last_path(), c.last_shape()) == (best, shape)
                ref = c.track(f1, f2, 1.0, 1.0, sigma)
                assert np.array_equal(got[2], ref[2])
                assert _rms(got, ref) <= 1e-4, (form, best, shape)
        finally:
            for p_ in d:
                c.free(p_)
    finally:
        c.close()
