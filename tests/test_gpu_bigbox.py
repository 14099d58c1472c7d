"""Boxes above 96 (sigma >= 24): the reference bounds the box only by the
image size (flct.c:567-580).  k_track_big keeps each pixel's N x N spectrum
in device memory and runs in FP64 for every precision form; checked against
the CPU oracle at 1e-8 px with vm bit-exact (GPU)."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import smooth_pair
from test_gpu_track import assert_parity

pytestmark = pytest.mark.gpu


def _pair(nx, ny, seed, shift=(1.3, -0.8)):
    return smooth_pair(np.random.default_rng(seed), nx, ny, shift=shift, width=4.0)


def _sample(nx, ny, rng, nrand):
    i, j = np.mgrid[0:nx, 0:ny]
    edge = np.flatnonzero((i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1))
    return np.unique(np.concatenate([edge[:: max(1, edge.size // 24)],
                                     [0, ny - 1, (nx - 1) * ny, nx * ny - 1],
                                     rng.choice(nx * ny, nrand, replace=False)]))


def _sampled(g, pix):
    return tuple(v.ravel()[pix] for v in g)


# (sigma, nx, ny): box 100 (Bluestein, length 256), 128 (radix-2) and 212
# (Bluestein, length 512) on the 400^2 image
@pytest.mark.parametrize("sigma,nx,ny", [(24.0, 170, 183), (30.0, 200, 211), (50.0, 400, 400)])
def test_big_box_against_oracle(ctx, sigma, nx, ny):
    a, b = _pair(nx, ny, int(sigma))
    nbox = ctx.box_size(nx, ny, sigma)
    assert nbox > 96
    g = ctx.track(a, b, 1.0, 1.0, sigma)
    assert ctx.last_path() == "fp64-plane"
    assert ctx.last_stats()["tracked"] == nx * ny
    rng = np.random.default_rng(int(sigma) + 1)
    pix = _sample(nx, ny, rng, 40 if nbox > 200 else 120)
    o = O.flct_pixels(a, b, 1.0, 1.0, sigma, pix)
    assert_parity(_sampled(g, pix), o, where=f"sigma {sigma} box {nbox}")
    # the interior recovers the shift (the field moves by (1.3, -0.8))
    c = (slice(nx // 3, 2 * nx // 3), slice(ny // 3, 2 * ny // 3))
    assert abs(np.median(g[0][c]) - 1.3) < 0.3 and abs(np.median(g[1][c]) + 0.8) < 0.3


@pytest.mark.parametrize("kw", [{"kr": 0.5}, {"bias": True}, {"hires": True},
                                {"bias": True, "hires": True, "kr": 0.7},
                                {"deltat": 2.0, "deltas": 0.6}],
                         ids=["k", "b", "h", "bhk", "scale"])
def test_big_box_options(ctx, kw):
    kw = dict(kw)
    dt, ds = kw.pop("deltat", 1.0), kw.pop("deltas", 1.0)
    a, b = _pair(170, 176, 7, shift=(0.6, 1.1))
    g = ctx.track(a, b, dt, ds, 24.0, **kw)
    rng = np.random.default_rng(8)
    pix = _sample(170, 176, rng, 60)
    o = O.flct_pixels(a, b, dt, ds, 24.0, pix, **kw)
    assert_parity(_sampled(g, pix), o, where=f"box 100 {kw}")


def test_big_box_threshold_and_skip_fill(ctx):
    """A relative threshold and a skip lattice with cubic-convolution fill
    (the whole flct pipeline) against the oracle's full run."""
    a, b = _pair(150, 160, 9, shift=(0.9, 0.4))
    a = a * np.hypot(*np.mgrid[-1:1:150j, -1:1:160j])  # weak centre, so -t drops pixels
    b = b * np.hypot(*np.mgrid[-1:1:150j, -1:1:160j])
    kw = dict(thresh=0.25, skip=6, poff=2, qoff=-1, interp=True)
    g = ctx.track(a, b, 1.0, 1.0, 24.0, **kw)
    assert 0 < ctx.last_stats()["tracked"] < (150 // 6 + 1) * (160 // 6 + 1)
    o = O.flct(a, b, 1.0, 1.0, 24.0, **kw)
    assert_parity(g, o, where="box 100 -t -s")


def test_big_box_every_form_runs_fp64(ctx):
    a, b = _pair(170, 180, 10)
    assert ctx.box_size(170, 180, 24.0) == 100
    ref = ctx.track(a, b, 1.0, 1.0, 24.0)
    for form in ("fp32", "split_bf16"):
        g = ctx.track(a, b, 1.0, 1.0, 24.0, precision=form)
        assert ctx.last_path() == "fp64-plane" and ctx.last_stats()["refined"] == 0
        for k in range(3):
            assert np.array_equal(g[k], ref[k]), (form, k)


def test_big_box_in_a_batch(ctx):
    """Boxes 100 and 128 next to quad-FFT boxes in one graph launch: each
    pair equals its single call bit for bit, on two replays."""
    items = []
    for k, (nx, ny, sigma, kw) in enumerate([(170, 180, 24.0, {}), (120, 130, 6.0, {}),
                                             (200, 205, 30.0, {"kr": 0.6, "bias": True}),
                                             (150, 150, 12.0, {"precision": "split_bf16"}),
                                             (160, 170, 24.0, {"thresh": 0.3, "precision": "fp32"})]):
        a, b = _pair(nx, ny, 20 + k)
        items.append((a, b, 1.0, 1.0, sigma, kw))
    first = ctx.track_batch(items)
    again = ctx.track_batch(items)
    for (a, b, dt, ds, s, kw), g, h in zip(items, first, again):
        ref = ctx.track(a, b, dt, ds, s, **kw)
        for k in range(3):
            assert np.array_equal(g[k], ref[k]) and np.array_equal(h[k], ref[k]), (a.shape, s, kw, k)
