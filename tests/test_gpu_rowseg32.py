"""The FP32 row-shared path ("fp32-rowseg", GPU).

k_track_rs<M, float>: the box-row transforms shared by a row segment stay
FP64; each row task forms its two unpacked spectra in FP64, scales each image
by a power of two and rounds to FP32 once; transforms, cross power, argmax
and the fit's inputs are FP32.  Pixels whose FP32 surface cannot settle the
peak (a second cell within 1e-5 of the maximum, an ill-conditioned fit) and
degenerate boxes are tracked again by the FP64 kernels.  The contract is the
FP32 form's: vm bit-exact, RMS <= 1e-4 px and max <= 1e-2 px against FP64.
"""
import numpy as np
import pytest

from pyflct_amd import synth
from conftest import smooth_pair
from test_gpu_bf16 import BOXES, _box_case
import oracle_lib

pytestmark = pytest.mark.gpu


def _errs(a, b):
    e = np.hypot(b[0] - a[0], b[1] - a[1])
    return float(np.sqrt(np.mean(e ** 2))), float(e.max())


@pytest.fixture(scope="module")
def ctx32():
    from pyflct_amd import api
    c = api.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("nbox", BOXES, ids=[f"N{b}" for b in BOXES])
def test_every_box_size(ctx32, nbox):
    c = ctx32
    sigma, nx, ny = _box_case(nbox)
    rng = np.random.default_rng(300 + nbox)
    a, b = smooth_pair(rng, nx, ny, shift=(0.6, -0.9), width=max(1.5, nbox / 16))
    ref = c.track(a, b, 1.0, 1.0, sigma)
    c.set_config("fp32", nbox, "fp32-rowseg")
    try:
        got = c.track(a, b, 1.0, 1.0, sigma, precision="fp32")
        assert c.last_path() == "fp32-rowseg"
    finally:
        c.set_config("fp32", nbox, None)
    assert np.array_equal(got[2], ref[2])
    rms, mx = _errs(ref, got)
    assert rms <= 1e-4 and mx <= 1e-2, (nbox, rms, mx)


@pytest.mark.parametrize("sigma,thresh", [(8.0, 0.0), (15.0, 0.0), (11.0, 0.2)])
def test_full_disk_pairs(ctx32, sigma, thresh):
    """Full-disk pairs with a zero limb (degenerate and near-degenerate
    boxes) and a relative threshold."""
    c = ctx32
    f1, f2, _ = synth.cfg4(n=448)
    nbox = c.box_size(448, 448, sigma)
    ref = c.track(f1, f2, 1.0, 1.0, sigma, thresh=thresh)
    c.set_config("fp32", nbox, "fp32-rowseg")
    try:
        got = c.track(f1, f2, 1.0, 1.0, sigma, thresh=thresh, precision="fp32")
        st = c.last_stats()
        assert c.last_path() == "fp32-rowseg"
        # without a threshold the limb's constant boxes are re-evaluated
        assert thresh > 0 or st["refined"] > 0
    finally:
        c.set_config("fp32", nbox, None)
    assert np.array_equal(got[2], ref[2])
    rms, mx = _errs(ref, got)
    assert rms <= 1e-4 and mx <= 1e-2, (rms, mx)


def test_against_oracle_and_shapes(ctx32):
    """cfg1 (128^2, box 32) against the oracle on every pixel, and the same
    result within the contract for other slot counts and segment widths."""
    c = ctx32
    f1, f2 = synth.cfg1()
    o = oracle_lib.flct(f1, f2, 1.0, 1.0, 8.0)
    for slots, segpx in ((0, 0), (1, 0), (0, 16), (1, 8)):
        c.set_config("fp32", 32, "fp32-rowseg", slots=slots, segpx=segpx)
        try:
            got = c.track(f1, f2, 1.0, 1.0, 8.0, precision="fp32")
            assert c.last_path() == "fp32-rowseg" and c.last_shape() == (slots, segpx)
        finally:
            c.set_config("fp32", 32, None)
        assert np.array_equal(got[2], o[2])
        rms, mx = _errs(o, got)
        assert rms <= 1e-4 and mx <= 1e-2, (slots, segpx, rms, mx)


def test_split_bf16_may_run_it(ctx32):
    """The split-BF16 form accepts the FP32 row-shared path (within its
    contract) and pins it per box size."""
    c = ctx32
    sigma, nx, ny = _box_case(80)
    rng = np.random.default_rng(80)
    a, b = smooth_pair(rng, nx, ny, shift=(0.3, 0.5), width=5.0)
    ref = c.track(a, b, 1.0, 1.0, sigma)
    c.set_config("split_bf16", 80, "fp32-rowseg")
    try:
        got = c.track(a, b, 1.0, 1.0, sigma, precision="split_bf16")
        assert c.last_path() == "fp32-rowseg"
    finally:
        c.set_config("split_bf16", 80, None)
    assert np.array_equal(got[2], ref[2])
    rms, mx = _errs(ref, got)
    assert rms <= 1e-4 and mx <= 1e-2, (rms, mx)


def test_filter_and_batch(ctx32):
    """-k (the filter applied to both spectra of every row task) and the
    batch graph: each pair of a batch equals its single call bit for bit."""
    c = ctx32
    rng = np.random.default_rng(41)
    items = []
    for nx, ny, sigma, kw in ((300, 260, 15.0, {"kr": 0.5}), (200, 230, 9.0, {}),
                              (260, 300, 19.0, {"kr": 0.7})):
        a, b = smooth_pair(rng, nx, ny, shift=(0.5, -0.8), width=3.0)
        items.append((a, b, 1.0, 1.0, sigma, dict(kw, precision="fp32")))
    for a, b, dt, ds, s, kw in items:
        nbox = c.box_size(a.shape[0], a.shape[1], s)
        c.set_config("fp32", nbox, "fp32-rowseg")
    try:
        got = c.track_batch(items)
        for (a, b, dt, ds, s, kw), g in zip(items, got):
            single = c.track(a, b, dt, ds, s, **kw)
            assert c.last_path() == "fp32-rowseg"
            for k in range(3):
                assert np.array_equal(g[k], single[k]), (a.shape, s, k)
            fkw = {k: v for k, v in kw.items() if k != "precision"}
            ref = c.track(a, b, dt, ds, s, **fkw)
            assert np.array_equal(single[2], ref[2])
            rms, mx = _errs(ref, single)
            assert rms <= 1e-4 and mx <= 1e-2, (a.shape, s, rms, mx)
    finally:
        for a, b, dt, ds, s, kw in items:
            c.set_config("fp32", c.box_size(a.shape[0], a.shape[1], s), None)
