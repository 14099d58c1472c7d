"""The row-shared FP64 path ("fp64-rowseg", k_track_rs) against the CPU
oracle: the same 1e-8 px and bit-exact vm as the FP64 plain form, at every
box size 4..96, with thresholds, the filter and a large constant level (GPU).
Calls it does not serve (skip lattices, -b, -h) fall back to k_track."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import smooth_pair
from test_gpu_track import assert_parity

pytestmark = pytest.mark.gpu


@pytest.fixture
def rowseg(ctx):
    pinned = set()

    def pin(nbox):
        ctx.set_path("fp64", nbox, "fp64-rowseg")
        pinned.add(nbox)
    yield pin
    for nbox in pinned:
        ctx.set_path("fp64", nbox, None)


def _sample(nx, ny, rng, nrand):
    i, j = np.mgrid[0:nx, 0:ny]
    edge = np.flatnonzero((i < 2) | (i >= nx - 2) | (j < 2) | (j >= ny - 2))
    return np.unique(np.concatenate([edge[::5], rng.choice(nx * ny, nrand, replace=False)]))


@pytest.mark.parametrize("sigma", list(range(1, 24)))
def test_rowseg_every_box_size(ctx, rowseg, sigma):
    rng = np.random.default_rng(300 + sigma)
    nbox = 2 * O.box_nsize(sigma, 4096, 4096)
    nx, ny = int(1.6 * nbox) + 3, int(2.3 * nbox) + 70  # several column bands, odd sizes
    a, b = smooth_pair(rng, nx, ny, shift=(1.1, -0.7), width=2.5)
    rowseg(nbox)
    g = ctx.track(a, b, 1.0, 1.0, float(sigma))
    assert ctx.last_path() == "fp64-rowseg"
    pix = _sample(nx, ny, rng, 150 if nbox > 48 else 400)
    o = O.flct_pixels(a, b, 1.0, 1.0, float(sigma), pix)
    assert_parity(tuple(v.ravel()[pix] for v in g), o, where=f"rowseg box {nbox}")


def test_rowseg_box_56_by_clamp(ctx, rowseg):
    rng = np.random.default_rng(356)
    nx, ny = 85, 190
    assert ctx.box_size(nx, ny, 14.0) == 56
    a, b = smooth_pair(rng, nx, ny, shift=(0.9, -0.4), width=3.0)
    rowseg(56)
    g = ctx.track(a, b, 1.0, 1.0, 14.0)
    assert ctx.last_path() == "fp64-rowseg"
    o = O.flct(a, b, 1.0, 1.0, 14.0)
    assert_parity(g, o, where="rowseg box 56")


@pytest.mark.parametrize("kw", [{"thresh": 0.3}, {"thresh": 0.4, "thresh_abs": True}, {"kr": 0.5},
                                {"thresh": 0.2, "kr": 0.7, "deltat": 2.0, "deltas": 0.5}],
                         ids=["rel", "abs", "k", "mixed"])
def test_rowseg_options(ctx, rowseg, kw):
    kw = dict(kw)
    dt, ds = kw.pop("deltat", 1.0), kw.pop("deltas", 1.0)
    rng = np.random.default_rng(310)
    a, b = smooth_pair(rng, 150, 211, shift=(0.5, 1.2), width=2.0)
    rowseg(40)
    g = ctx.track(a, b, dt, ds, 10.0, **kw)
    assert ctx.last_path() == "fp64-rowseg"
    o = O.flct(a, b, dt, ds, 10.0, **kw)
    assert_parity(g, o, where=f"rowseg {kw}")


def test_rowseg_pedestal_and_zero_boxes(ctx, rowseg):
    """A large constant level (the shared transform subtracts the segment's
    reference mean, the pixel adds back its own offset) and all-zero boxes
    (the max == 0 branch) match the oracle."""
    rng = np.random.default_rng(311)
    a, b = smooth_pair(rng, 200, 230, shift=(-0.6, 0.8), width=2.5)
    s = 100.0 / np.std(a)
    a, b = 1e4 + s * a, 1e4 + s * b
    rowseg(64)
    g = ctx.track(a, b, 1.0, 1.0, 15.0)
    pix = _sample(200, 230, rng, 200)
    o = O.flct_pixels(a, b, 1.0, 1.0, 15.0, pix)
    assert_parity(tuple(v.ravel()[pix] for v in g), o, where="rowseg pedestal")
    z = np.zeros((120, 150))
    z[:, 100:] = a[:120, 100:150]
    g = ctx.track(z, z * 0.5, 1.0, 1.0, 15.0)
    assert_parity(g, O.flct(z, z * 0.5, 1.0, 1.0, 15.0), where="rowseg zero boxes")


def test_rowseg_falls_back_where_it_does_not_apply(ctx, rowseg):
    rng = np.random.default_rng(312)
    a, b = smooth_pair(rng, 120, 130, shift=(0.4, 0.3))
    rowseg(24)
    for kw in ({"skip": 3, "interp": True}, {"bias": True}, {"hires": True}):
        g = ctx.track(a, b, 1.0, 1.0, 6.0, **kw)
        assert ctx.last_path() == "fp64-fft", kw
        assert_parity(g, O.flct(a, b, 1.0, 1.0, 6.0, **kw), where=f"fallback {kw}")


def test_rowseg_in_a_batch(ctx, rowseg):
    rng = np.random.default_rng(313)
    items = []
    for k, (nx, ny, sigma, kw) in enumerate([(130, 150, 8.0, {}), (200, 260, 15.0, {"thresh": 0.3}),
                                             (90, 100, 4.0, {"kr": 0.5})]):
        a, b = smooth_pair(rng, nx, ny, shift=(0.7, -0.2))
        items.append((a, b, 1.0, 1.0, sigma, kw))
    for nbox in (32, 64, 16):
        rowseg(nbox)
    got = ctx.track_batch(items)
    for (a, b, dt, ds, s, kw), g in zip(items, got):
        ref = ctx.track(a, b, dt, ds, s, **kw)
        assert ctx.last_path() == "fp64-rowseg"
        for k in range(3):
            assert np.array_equal(g[k], ref[k]), (a.shape, s, k)


def test_rowseg_degenerate_boxes_in_mixed_segments(ctx, rowseg):
    """All-zero boxes in segments whose reference mean is not zero get the
    zero surface's result (v = 0) as in the reference; boxes constant at a
    value whose box mean is inexact are re-evaluated by k_track; both match
    the oracle."""
    rng = np.random.default_rng(314)
    nx, ny = 140, 300
    a, b = smooth_pair(rng, nx, ny, shift=(0.8, 0.5), width=2.5)
    a[:, 100:] = 0.0
    b[:, 100:] = 0.0
    a[:, 200:270] = 0.1  # a constant plateau; 4096 x 0.1 does not sum exactly
    b[:, 200:270] = 0.1
    rowseg(64)
    g = ctx.track(a, b, 1.0, 1.0, 15.0)
    assert ctx.last_path() == "fp64-rowseg"
    assert ctx.last_stats()["refined"] > 0
    cols = np.arange(ny)
    pix = np.concatenate([r * ny + cols[::3] for r in (0, 37, 70, 139)])
    o = O.flct_pixels(a, b, 1.0, 1.0, 15.0, pix)
    assert_parity(tuple(v.ravel()[pix] for v in g), o, where="rowseg degenerate boxes")
    # boxes of columns 131..167 (box columns j - 31 .. j + 32) hold only zeros
    assert np.all(g[0][:, 131:168] == 0) and np.all(g[1][:, 131:168] == 0)


def test_rowseg_full_disk_with_limb(ctx, rowseg):
    from pyflct_amd import synth
    f1, f2, _ = synth.cfg4(n=512)
    rowseg(64)
    g = ctx.track(f1, f2, 1.0, 1.0, 15.0)
    assert ctx.last_path() == "fp64-rowseg"
    ctx.set_path("fp64", 64, "fp64-fft")  # the fixture unpins box 64 afterwards
    ref = ctx.track(f1, f2, 1.0, 1.0, 15.0)
    assert ctx.last_path() == "fp64-fft"
    rng = np.random.default_rng(315)
    pix = np.unique(np.concatenate([rng.choice(512 * 512, 300, replace=False),
                                    256 * 512 + np.arange(0, 512, 7), np.arange(0, 512, 9)]))
    o = O.flct_pixels(f1, f2, 1.0, 1.0, 15.0, pix)
    assert_parity(tuple(v.ravel()[pix] for v in g), o, where="rowseg full disk")
    assert_parity(g, ref, where="rowseg against fp64-fft, every pixel")
