"""GPU parity of flct_hip_track against the CPU oracle (FP64 plain form:
|dv| <= 1e-8 px, vm and the thresholded/skipped sets bit-exact)."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import f32, smooth_pair

pytestmark = pytest.mark.gpu
TOL = 1e-8  # px, FP64 plain form


def assert_parity(g, o, tol=TOL, where=""):
    vx, vy, vm = g
    ox, oy, om = o
    assert np.array_equal(vm, om), f"{where}: vm differs at {np.argwhere(vm != om)[:5]}"
    ex = np.max(np.abs(vx - ox)) if vx.size else 0.0
    ey = np.max(np.abs(vy - oy)) if vy.size else 0.0
    assert ex <= tol and ey <= tol, f"{where}: max|dvx| {ex:.3e} max|dvy| {ey:.3e}"


def test_reference_fixture(ctx, golden):
    """hashgauss.dat with "1 1 5 -k 0.5" -> testgaussvel.dat (README.md:207-208)."""
    g = ctx.track(golden["f1"], golden["f2"], 1.0, 1.0, 5.0, kr=0.5)
    o = O.flct(golden["f1"], golden["f2"], 1.0, 1.0, 5.0, kr=0.5)
    assert_parity(g, o, where="hashgauss")
    assert np.array_equal(g[2], golden["vm"])
    for a, r in ((g[0], golden["vx"]), (g[1], golden["vy"])):
        assert np.max(np.abs(f32(a) - r)) < 1e-11
        assert np.sum(f32(a) == r) >= 0.998 * r.size


SIGMAS_FULL = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11]          # boxes 4..44
SIGMAS_SAMPLED = [12, 13, 14, 15, 16, 17, 18, 19, 20, 22, 23]  # 48..96 (no integer sigma gives 56)


FP64_PATHS = ["fp64-fft", "fp64-rowseg"]  # quad-FFT k_track; row-shared k_track_rs


@pytest.mark.parametrize("path", FP64_PATHS)
@pytest.mark.parametrize("sigma,kr", [(15, 0.4), (19, 0.6)])
def test_kspace_filter_each_path(ctx, sigma, kr, path):
    """-k through each FP64 kernel at boxes 64 and 80 (the row-shared kernel
    runs its row tasks as one shared-transform form at 64 and with a branch
    per task kind at 80), sampled pixels against the oracle; then the FP32
    row-shared kernel against the FP64 result."""
    rng = np.random.default_rng(11 + sigma)
    n = 2 * O.box_nsize(sigma, 4096, 4096)
    a, b = smooth_pair(rng, int(1.6 * n) + 2, int(1.6 * n) + 7)
    ctx.set_path("fp64", n, path)
    try:
        g = ctx.track(a, b, 1.0, 1.0, sigma, kr=kr)
        assert ctx.last_path() == path
    finally:
        ctx.set_path("fp64", n, None)
    nx, ny = a.shape
    pix = rng.choice(nx * ny, 150, replace=False)
    o = O.flct_pixels(a, b, 1.0, 1.0, sigma, pix, kr=kr)
    assert_parity((g[0].ravel()[pix], g[1].ravel()[pix], g[2].ravel()[pix]), o,
                  where=f"-k {kr} N {n} {path}")
    if path == "fp64-rowseg":
        ctx.set_path("fp32", n, "fp32-rowseg")
        try:
            f = ctx.track(a, b, 1.0, 1.0, sigma, kr=kr, precision="fp32")
            assert ctx.last_path() == "fp32-rowseg"
        finally:
            ctx.set_path("fp32", n, None)
        assert np.array_equal(f[2], g[2])
        e = np.hypot(f[0] - g[0], f[1] - g[1])
        assert np.sqrt(np.mean(e ** 2)) <= 1e-4 and e.max() <= 5e-3, (np.sqrt(np.mean(e ** 2)), e.max())


@pytest.mark.parametrize("path", FP64_PATHS)
@pytest.mark.parametrize("sigma", SIGMAS_FULL)
def test_box_sizes_full_image(ctx, sigma, path):
    rng = np.random.default_rng(100 + sigma)
    n = 2 * O.box_nsize(sigma, 4096, 4096)
    nx, ny = int(1.6 * n) + 3, int(1.9 * n) + 5  # odd, non-square, edge-clamped boxes
    a, b = smooth_pair(rng, nx, ny)
    ctx.set_path("fp64", n, path)
    try:
        g = ctx.track(a, b, 1.0, 1.0, sigma)
        assert ctx.last_path() == path
    finally:
        ctx.set_path("fp64", n, None)
    o = O.flct(a, b, 1.0, 1.0, sigma)
    assert_parity(g, o, where=f"sigma {sigma} N {n} {path}")


@pytest.mark.parametrize("path", FP64_PATHS)
@pytest.mark.parametrize("sigma", SIGMAS_SAMPLED)
def test_box_sizes_sampled(ctx, sigma, path):
    rng = np.random.default_rng(200 + sigma)
    n = 2 * O.box_nsize(sigma, 4096, 4096)
    nx, ny = int(1.6 * n) + 1, int(1.7 * n) + 2
    a, b = smooth_pair(rng, nx, ny, shift=(1.3, 0.6))
    ctx.set_path("fp64", n, path)
    try:
        g = ctx.track(a, b, 2.0, 0.5, sigma)
        assert ctx.last_path() == path
    finally:
        ctx.set_path("fp64", n, None)
    i, j = np.mgrid[0:nx, 0:ny]
    edge = (i < 2) | (i >= nx - 2) | (j < 2) | (j >= ny - 2)
    pix = np.concatenate([np.flatnonzero(edge)[::7], rng.choice(nx * ny, 120, replace=False)])
    o = O.flct_pixels(a, b, 2.0, 0.5, sigma, pix)
    assert_parity((g[0].ravel()[pix], g[1].ravel()[pix], g[2].ravel()[pix]), o,
                  where=f"sigma {sigma} N {n}")


def test_box_56_clamped_by_image_size(ctx):
    """Box 56 has no integer sigma (nt jumps from 27 to 30 between sigma 13
    and 14); sigma 14 on an image 85 rows high clamps nsize to 28
    (flct.c:570-578)."""
    rng = np.random.default_rng(56)
    nx, ny = 85, 131
    assert ctx.box_size(nx, ny, 14.0) == 56
    a, b = smooth_pair(rng, nx, ny, shift=(0.9, -0.4), width=3.0)
    g = ctx.track(a, b, 1.0, 1.0, 14.0)
    i, j = np.mgrid[0:nx, 0:ny]
    edge = (i < 2) | (i >= nx - 2) | (j < 2) | (j >= ny - 2)
    pix = np.concatenate([np.flatnonzero(edge)[::5], rng.choice(nx * ny, 150, replace=False)])
    o = O.flct_pixels(a, b, 1.0, 1.0, 14.0, pix)
    assert_parity((g[0].ravel()[pix], g[1].ravel()[pix], g[2].ravel()[pix]), o, where="N 56")
    g = ctx.track(a, b, 1.0, 1.0, 14.0, hires=True)
    o = O.flct_pixels(a, b, 1.0, 1.0, 14.0, pix, hires=True)
    assert_parity((g[0].ravel()[pix], g[1].ravel()[pix], g[2].ravel()[pix]), o, where="N 56 -h")


@pytest.mark.parametrize("thresh,absolute", [(0.3, False), (0.3, True), (0.05, False),
                                             (1.2, False), (-1.0, False)])
def test_threshold_sets(ctx, thresh, absolute):
    rng = np.random.default_rng(7)
    a, b = smooth_pair(rng, 150, 161)
    g = ctx.track(a, b, 1.0, 1.0, 6.0, thresh=thresh, thresh_abs=absolute)
    o = O.flct(a, b, 1.0, 1.0, 6.0, thresh=thresh, thresh_abs=absolute)
    assert_parity(g, o, where=f"thresh {thresh}{'a' if absolute else ''}")


@pytest.mark.parametrize("skip,p,q,interp", [(3, 0, 0, True), (4, 1, 2, True), (5, -2, -1, True),
                                             (4, 3, -3, False), (2, 0, 1, True), (7, 0, 0, True)])
def test_skip_and_fill(ctx, skip, p, q, interp):
    rng = np.random.default_rng(8)
    a, b = smooth_pair(rng, 113, 127)
    g = ctx.track(a, b, 1.0, 1.0, 5.0, skip=skip, poff=p, qoff=q, interp=interp)
    o = O.flct(a, b, 1.0, 1.0, 5.0, skip=skip, poff=p, qoff=q, interp=interp)
    assert_parity(g, o, where=f"-s {skip}p{p}q{q}{'i' if interp else ''}")


def test_skip_fill_with_threshold_holes(ctx):
    """Thresholded lattice points are xmiss and poison their 4x4 stencils
    (flct.c:2430-2442), including the extrapolated border cells."""
    rng = np.random.default_rng(9)
    a, b = smooth_pair(rng, 121, 133)
    for t in (0.2, 0.45):
        g = ctx.track(a, b, 1.0, 1.0, 7.0, skip=4, poff=-1, qoff=2, interp=True, thresh=t)
        o = O.flct(a, b, 1.0, 1.0, 7.0, skip=4, poff=-1, qoff=2, interp=True, thresh=t)
        assert_parity(g, o, where=f"skip+thresh {t}")


@pytest.mark.parametrize("sigma,kr", [(5, 0.5), (6, 0.25), (10, 0.8), (15, 0.4)])
def test_kspace_filter(ctx, sigma, kr):
    rng = np.random.default_rng(10)
    n = 2 * O.box_nsize(sigma, 4096, 4096)
    a, b = smooth_pair(rng, int(1.6 * n) + 2, int(1.6 * n) + 7)
    g = ctx.track(a, b, 1.0, 1.0, sigma, kr=kr)
    nx, ny = a.shape
    pix = rng.choice(nx * ny, 150, replace=False)
    o = O.flct_pixels(a, b, 1.0, 1.0, sigma, pix, kr=kr)
    assert_parity((g[0].ravel()[pix], g[1].ravel()[pix], g[2].ravel()[pix]), o,
                  where=f"-k {kr}")


def test_bias_correction_parity_and_effect(ctx):
    from scipy.ndimage import gaussian_filter
    from pyflct_amd import synth
    rng = np.random.default_rng(11)
    a = gaussian_filter(rng.standard_normal((96, 96)), 2.0, mode="wrap")
    b = synth.shift_uniform(a, 1.5, -0.75)
    g = ctx.track(a, b, 1.0, 1.0, 6.0, bias=True)
    o = O.flct(a, b, 1.0, 1.0, 6.0, bias=True)
    assert_parity(g, o, where="-b")
    g0 = ctx.track(a, b, 1.0, 1.0, 6.0)
    c = (slice(24, 72), slice(24, 72))
    e0 = abs(np.median(g0[0][c]) - 1.5) + abs(np.median(g0[1][c]) + 0.75)
    e1 = abs(np.median(g[0][c]) - 1.5) + abs(np.median(g[1][c]) + 0.75)
    assert e1 < 0.75 * e0


def test_invariants(ctx):
    rng = np.random.default_rng(12)
    a, _ = smooth_pair(rng, 70, 66)
    for b in (a, -a):
        vx, vy, vm = ctx.track(a, b, 1.0, 1.0, 6.0)
        assert np.all(vm == 1) and np.max(np.abs(vx)) < 1e-10 and np.max(np.abs(vy)) < 1e-10
    z = np.zeros((70, 66))
    vx, vy, vm = ctx.track(z, z, 1.0, 1.0, 6.0)  # max == 0 branch everywhere
    assert np.all(vx == 0) and np.all(vy == 0) and np.all(vm == 1)
    c = np.full((70, 66), 3.25)
    g = ctx.track(c, c, 1.0, 1.0, 6.0)
    o = O.flct(c, c, 1.0, 1.0, 6.0)
    assert_parity(g, o, where="constant")


def test_cfg1_uniform_shift(ctx):
    """128^2 blobs, image 2 shifted by (0.3, -0.2), "1 1 8"."""
    from pyflct_amd import synth
    f1, f2 = synth.cfg1()
    g = ctx.track(f1, f2, 1.0, 1.0, 8.0)
    o = O.flct(f1, f2, 1.0, 1.0, 8.0)
    assert_parity(g, o, where="cfg1")


def test_cfg2_threshold_filter_bias(ctx):
    """512^2 magnetogram rotated by 0.01 deg (GPU warp), "1 1 6 -t 0.1 -k 0.5"
    with and without -b."""
    from pyflct_amd import synth
    f1, dx, dy = synth.cfg2_inputs()
    f2 = ctx.warp(f1, dx, dy)
    for bias in (False, True):
        g = ctx.track(f1, f2, 1.0, 1.0, 6.0, thresh=0.1, kr=0.5, bias=bias)
        o = O.flct(f1, f2, 1.0, 1.0, 6.0, thresh=0.1, kr=0.5, bias=bias)
        assert_parity(g, o, where=f"cfg2 bias={bias}")
        assert 0 < np.sum(g[2] == 1) < f1.size


def test_cfg3_skip_fill(ctx):
    """1024^2 magnetogram under a smooth random field, "1 1 10 -s 4i"."""
    from pyflct_amd import synth
    f1, dx, dy = synth.cfg3_inputs()
    f2 = ctx.warp(f1, dx, dy)
    g = ctx.track(f1, f2, 1.0, 1.0, 10.0, skip=4, interp=True)
    o = O.flct(f1, f2, 1.0, 1.0, 10.0, skip=4, interp=True)
    assert_parity(g, o, where="cfg3")
    assert ctx.last_stats()["tracked"] == 256 * 256
    assert np.sum(g[2] == 0.5) == 1024 * 1024 - 256 * 256 - np.sum(g[2] == 0)


def test_cfg3_fp32_form_in_one_graph_launch(ctx):
    """The same case with float32 images in the FP32 form, tracked through
    flct_hip_track_batch (one graph launch), against the FP64 form: the vm
    mask, the lattice and the filled set are bit-exact; velocities of tracked
    and filled pixels within RMS 1e-4 px and max 5e-3 px (observed ~1e-6 and
    ~3e-4 at 4096^2)."""
    from pyflct_amd import synth
    f1, dx, dy = synth.cfg3_inputs()
    f1 = f1.astype(np.float32).astype(np.float64)
    f2 = ctx.warp(f1, dx, dy).astype(np.float32).astype(np.float64)
    ref = ctx.track(f1, f2, 1.0, 1.0, 10.0, skip=4, interp=True)
    (got,) = ctx.track_batch([(f1, f2, 1.0, 1.0, 10.0)], skip=4, interp=True, precision="fp32")
    assert np.array_equal(got[2], ref[2])
    e = np.hypot(got[0] - ref[0], got[1] - ref[1])
    for sel in (ref[2] == 1.0, ref[2] == 0.5):
        assert sel.any()
        rms, mx = float(np.sqrt(np.mean(e[sel] ** 2))), float(e[sel].max())
        print(f"1024^2 f32 -s 4i, FP32 graph vs FP64: rms {rms:.2e} max {mx:.2e} px over {sel.sum()} px")
        assert rms <= 1e-4 and mx <= 5e-3, (rms, mx)
    z = ref[2] == 0.0
    assert np.array_equal(got[0][z], ref[0][z]) and np.array_equal(got[1][z], ref[1][z])


def test_cfg4_full_disk_sampled(ctx):
    """4096^2 full-disk pair, "1 1 15": sampled parity (interior, limb with
    the max == 0 branch, frame edges) and recovery of the rotation law."""
    from pyflct_amd import synth
    f1, f2, vtrue = synth.cfg4()
    g = ctx.track(f1, f2, 1.0, 1.0, 15.0)
    assert ctx.last_stats()["tracked"] == 4096 * 4096
    rng = np.random.default_rng(13)
    n = 4096
    pix = np.concatenate([rng.choice(n * n, 250, replace=False),
                          np.array([0, n - 1, n * n - 1, n * (n - 1), n * 2048 + 2048]),
                          np.arange(n * 2048, n * 2048 + n, 97)])
    o = O.flct_pixels(f1, f2, 1.0, 1.0, 15.0, pix)
    assert_parity((g[0].ravel()[pix], g[1].ravel()[pix], g[2].ravel()[pix]), o, where="cfg4")
    assert np.all(g[2] == 1)
    core = (slice(1200, 2900), slice(1200, 2900))
    assert abs(np.median(g[0][core]) - np.median(vtrue[core])) < 0.15
    assert np.all(g[0][:10, :10] == 0) and np.all(g[1][:10, :10] == 0)  # limb: zero boxes


def test_fp32_form_error_bound(ctx):
    from pyflct_amd import synth
    f1, f2 = synth.cfg1()
    g = ctx.track(f1, f2, 1.0, 1.0, 8.0, precision="fp32")
    o = O.flct(f1, f2, 1.0, 1.0, 8.0)
    assert np.array_equal(g[2], o[2])
    rms = np.sqrt(np.mean((g[0] - o[0]) ** 2 + (g[1] - o[1]) ** 2))
    assert rms <= 1e-4, rms


def test_error_codes(ctx):
    from pyflct_amd import _lib as L
    from pyflct_amd.api import FlctHipError
    a = np.zeros((40, 40))
    cases = [({"sigma": 4.0, "skip": 50}, L.FLCT_HIP_E_SKIP_TOO_BIG),
             ({"sigma": 4.0, "skip": 4, "poff": 5}, L.FLCT_HIP_E_PQ_RANGE),
             ({"sigma": 4.0, "skip": -1}, L.FLCT_HIP_E_SKIP_NONPOS),
             ({"sigma": 0.5}, L.FLCT_HIP_E_BOX),  # (int)sigma = 0: box 0 (flct.c:570-579)
             ({"sigma": 0.0, "skip": 40}, L.FLCT_HIP_E_SKIP_TOO_BIG)]
    for kw, code in cases:
        sig = kw.pop("sigma")
        with pytest.raises(FlctHipError) as e:
            ctx.track(a, a, 1.0, 1.0, sig, **kw)
        assert e.value.status == code
    with pytest.raises(FlctHipError) as e:
        ctx.track(np.zeros((4, 4)), np.zeros((4, 4)), 1.0, 1.0, 5.0)
    assert e.value.status == L.FLCT_HIP_E_BOX


# ------------------------------------------------------------- sigma = 0
def _whole_pair(nx, ny, seed, shift=(0.35, -0.6)):
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    g = gaussian_filter(rng.standard_normal((nx, ny)), 2.5, mode="wrap") + 3.0
    kx = np.fft.fftfreq(nx)[:, None]
    ky = np.fft.fftfreq(ny)[None, :]
    h = np.fft.ifft2(np.fft.fft2(g) * np.exp(-2j * np.pi * (kx * shift[0] + ky * shift[1]))).real
    return g, h + 0.02 * rng.standard_normal((nx, ny))


@pytest.mark.parametrize("nx,ny,kw", [(64, 64, {}), (101, 201, {}), (127, 96, {"kr": 0.5}),
                                      (250, 333, {"deltat": 2.0, "deltas": 0.7}),
                                      (512, 384, {"kr": 0.3})])
def test_sigma0_whole_image(ctx, nx, ny, kw):
    """sigma = 0: one correlation of the whole images (flct.c:567-586,
    802-808, 920-947), any nx, ny: the single velocity against the oracle."""
    kw = dict(kw)
    dt, ds = kw.pop("deltat", 1.0), kw.pop("deltas", 1.0)
    f1, f2 = _whole_pair(nx, ny, seed=nx + ny)
    g = ctx.track(f1, f2, dt, ds, 0.0, **kw)
    o = O.flct(f1, f2, dt, ds, 0.0, **kw)
    assert g[0].shape == (1, 1)
    assert_parity(g, o, where=f"sigma=0 {nx}x{ny} {kw}")
    assert abs(g[0][0, 0] * dt / ds + 0.6) < 0.1 and abs(g[1][0, 0] * dt / ds - 0.35) < 0.1


def test_sigma0_ignores_threshold_skip_and_takes_device_arrays(ctx):
    f1, f2 = _whole_pair(90, 110, seed=7, shift=(-1.3, 2.2))
    ref = ctx.track(f1, f2, 1.0, 1.0, 0.0)
    for kw in ({"thresh": 0.9}, {"thresh": 1e6, "thresh_abs": True}, {"skip": 6, "qoff": 2}):
        g = ctx.track(f1, f2, 1.0, 1.0, 0.0, **kw)
        for a, b in zip(ref, g):
            assert np.array_equal(a, b), kw
    # a negative sigma is the same whole-image case (flct.c:357-365)
    for neg in (-3.0, -0.25):
        g = ctx.track(f1, f2, 1.0, 1.0, neg)
        for a, b in zip(ref, g):
            assert a.shape == (1, 1) and np.array_equal(a, b), neg
    # device entry point: the velocity goes to element 0 of each output
    d = [ctx.malloc(f1.nbytes) for _ in range(2)] + [ctx.malloc(8) for _ in range(3)]
    try:
        ctx.upload(d[0], f1)
        ctx.upload(d[1], f2)
        ctx.track_device(d[0], d[1], 90, 110, d[2], d[3], d[4], ctx.make_opts(1.0, 1.0, 0.0))
        got = [ctx.download(np.empty(1), d[k]) for k in (2, 3, 4)]
        for a, b in zip(ref, got):
            assert np.array_equal(a.ravel(), b)
    finally:
        for p in d:
            ctx.free(p)
    # zero images: the maximum-zero branch puts the peak at the centre
    z = np.zeros((33, 48))
    assert [float(v[0, 0]) for v in ctx.track(z, z, 1.0, 1.0, 0.0)] == [0.0, 0.0, 1.0]
    assert_parity(ctx.track(z, z, 1.0, 1.0, 0.0), O.flct(z, z, 1.0, 1.0, 0.0))


# ------------------------------------------------------------------ -h
@pytest.mark.parametrize("sigma,kw", [(4.0, {}), (5.0, {"kr": 0.5}), (6.0, {"bias": True}),
                                      (8.0, {"thresh": 0.3}), (10.0, {"skip": 3, "poff": 1, "interp": True}),
                                      (15.0, {})])
def test_hires_peak(ctx, sigma, kw):
    """-h: the cubic-convolution peak on the 0.1 px grid (flct.c:1357-1368,
    1564-1645) in the FP64 kernels, against the oracle on every pixel."""
    rng = np.random.default_rng(int(sigma * 10))
    nx, ny = (100, 110) if sigma < 15 else (150, 140)
    f1, f2 = smooth_pair(rng, nx, ny, shift=(0.45, -0.7), width=2.5)
    g = ctx.track(f1, f2, 1.0, 1.0, sigma, hires=True, **kw)
    assert ctx.last_path() == "fp64-fft"
    o = O.flct(f1, f2, 1.0, 1.0, sigma, hires=True, **kw)
    assert_parity(g, o, where=f"hires sigma={sigma} {kw}")
    plain = ctx.track(f1, f2, 1.0, 1.0, sigma, **kw)
    assert not np.array_equal(g[0], plain[0])


HIRES_SIGMAS = [1, 2, 3, 7, 9, 11, 12, 13, 14, 16, 17, 18, 19, 20, 21, 22, 23]


@pytest.mark.parametrize("sigma", HIRES_SIGMAS)
def test_hires_every_box_size(ctx, sigma):
    """-h in each FP64 box-size instantiation the cases above leave out
    (boxes 4..96 but 56, which test_box_56_clamped_by_image_size covers), on
    sampled pixels including the frame edges.  Boxes up to 16 have fewer
    threads per slot (2N) than the 36 padded-block entries and the 9
    neighbours the -h step fills."""
    rng = np.random.default_rng(300 + sigma)
    n = 2 * O.box_nsize(sigma, 4096, 4096)
    nx, ny = int(1.6 * n) + 3, int(1.7 * n) + 4
    a, b = smooth_pair(rng, nx, ny, shift=(0.35, -0.6), width=max(1.5, n / 16))
    g = ctx.track(a, b, 1.0, 1.0, sigma, hires=True)
    i, j = np.mgrid[0:nx, 0:ny]
    edge = (i < 2) | (i >= nx - 2) | (j < 2) | (j >= ny - 2)
    pix = np.concatenate([np.flatnonzero(edge)[::9], rng.choice(nx * ny, 60, replace=False)])
    o = O.flct_pixels(a, b, 1.0, 1.0, sigma, pix, hires=True)
    assert_parity((g[0].ravel()[pix], g[1].ravel()[pix], g[2].ravel()[pix]), o,
                  where=f"hires sigma {sigma} N {n}")


def test_hires_zero_images_and_other_forms(ctx):
    """All-zero boxes give the grid-corner shift (-1, -1) under -h, as in the
    reference's rules; FP32 and split-BF16 requests with -h run the FP64
    kernels and return the FP64 result."""
    z = np.zeros((40, 44))
    g = ctx.track(z, z, 2.0, 3.0, 5.0, hires=True)
    assert np.all(g[0] == -1.5) and np.all(g[1] == -1.5) and np.all(g[2] == 1.0)
    assert_parity(g, O.flct(z, z, 2.0, 3.0, 5.0, hires=True))
    rng = np.random.default_rng(77)
    f1, f2 = smooth_pair(rng, 128, 128, shift=(-0.3, 0.55))
    ref = ctx.track(f1, f2, 1.0, 1.0, 12.0, hires=True)
    for form in ("fp32", "split_bf16"):
        got = ctx.track(f1, f2, 1.0, 1.0, 12.0, hires=True, precision=form)
        assert ctx.last_path() == "fp64-fft"
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), form


@pytest.mark.parametrize("nx,ny,kw", [(64, 80, {}), (101, 97, {"kr": 0.5})])
def test_hires_sigma0(ctx, nx, ny, kw):
    f1, f2 = _whole_pair(nx, ny, seed=nx * ny)
    g = ctx.track(f1, f2, 1.0, 1.0, 0.0, hires=True, **kw)
    o = O.flct(f1, f2, 1.0, 1.0, 0.0, hires=True, **kw)
    assert_parity(g, o, where=f"hires sigma=0 {nx}x{ny}")


@pytest.mark.parametrize("path", ["fp32-fft", "fp32-rowseg"])
def test_fp32_form_refines_unsettled_pixels(ctx, path):
    """The FP32 form lists pixels whose FP32 surface cannot settle the peak
    (a second cell within 1e-5 of the maximum, or an ill-conditioned fit)
    and tracks them again in FP64, which bounds its worst error on a
    full-disk pair with a zero limb (boxes grazing the limb are the hard
    ones).  Both FP32 kernels: the quad-FFT tracker and the row-shared one."""
    from pyflct_amd import synth
    f1, f2, _ = synth.cfg4(n=512)
    ref = ctx.track(f1, f2, 1.0, 1.0, 8.0)
    ctx.set_path("fp32", 32, path)
    try:
        g = ctx.track(f1, f2, 1.0, 1.0, 8.0, precision="fp32")
        st = ctx.last_stats()
        assert ctx.last_path() == path and st["refined"] > 0
    finally:
        ctx.set_path("fp32", 32, None)
    assert np.array_equal(g[2], ref[2])
    err = np.hypot(g[0] - ref[0], g[1] - ref[1])
    assert np.sqrt(np.mean(err ** 2)) <= 1e-4 and err.max() <= 1e-2, (err.max(), st["refined"])


OPTION_SIGMAS = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23]


@pytest.mark.parametrize("sigma", OPTION_SIGMAS)
def test_filter_and_bias_every_box_size(ctx, sigma):
    """-k 0.6 -b at boxes 4..96 (but 56): FP64 against the oracle on sampled
    pixels; FP32 and split-BF16 against FP64 (RMS <= 1e-4 px, max <= 5e-3
    px, vm bit-exact)."""
    rng = np.random.default_rng(400 + sigma)
    n = 2 * O.box_nsize(sigma, 4096, 4096)
    nx, ny = int(1.6 * n) + 5, int(1.7 * n) + 2
    a, b = smooth_pair(rng, nx, ny, shift=(0.8, 0.45), width=max(1.5, n / 16))
    kw = dict(kr=0.6, bias=True)
    g = ctx.track(a, b, 1.0, 1.0, sigma, **kw)
    i, j = np.mgrid[0:nx, 0:ny]
    edge = (i < 2) | (i >= nx - 2) | (j < 2) | (j >= ny - 2)
    pix = np.concatenate([np.flatnonzero(edge)[::11], rng.choice(nx * ny, 50, replace=False)])
    o = O.flct_pixels(a, b, 1.0, 1.0, sigma, pix, **kw)
    assert_parity((g[0].ravel()[pix], g[1].ravel()[pix], g[2].ravel()[pix]), o,
                  where=f"-k -b sigma {sigma} N {n}")
    for form in ("fp32", "split_bf16"):
        f = ctx.track(a, b, 1.0, 1.0, sigma, precision=form, **kw)
        assert np.array_equal(f[2], g[2]), form
        e = np.hypot(f[0] - g[0], f[1] - g[1])
        rms, mx = float(np.sqrt(np.mean(e ** 2))), float(e.max())
        assert rms <= 1e-4 and mx <= 5e-3, (form, ctx.last_path(), rms, mx)


@pytest.mark.parametrize("sigma", OPTION_SIGMAS)
def test_threshold_and_skip_every_box_size(ctx, sigma):
    """-t 0.25 -s 3p1q2i at boxes 4..96 (but 56): FP64 against the oracle
    (whole image, fill included); FP32 and split-BF16 against FP64 with the
    vm mask, the thresholded and the filled sets bit-exact."""
    rng = np.random.default_rng(600 + sigma)
    n = 2 * O.box_nsize(sigma, 4096, 4096)
    nx, ny = int(1.6 * n) + 4, int(1.6 * n) + 9
    a, b = smooth_pair(rng, nx, ny, shift=(0.3, -0.55), width=max(1.5, n / 16))
    kw = dict(thresh=0.25, skip=3, poff=1, qoff=2, interp=True)
    g = ctx.track(a, b, 1.0, 1.0, sigma, **kw)
    o = O.flct(a, b, 1.0, 1.0, sigma, **kw)
    assert_parity(g, o, where=f"-t -s sigma {sigma} N {n}")
    for form in ("fp32", "split_bf16"):
        f = ctx.track(a, b, 1.0, 1.0, sigma, precision=form, **kw)
        assert np.array_equal(f[2], g[2]), form
        z = g[2] == 0.0
        assert np.array_equal(f[0][z], g[0][z]) and np.array_equal(f[1][z], g[1][z])
        m = g[2] > 0.0
        e = np.hypot(f[0][m] - g[0][m], f[1][m] - g[1][m])
        if e.size:
            rms, mx = float(np.sqrt(np.mean(e ** 2))), float(e.max())
            assert rms <= 1e-4 and mx <= 5e-3, (form, ctx.last_path(), rms, mx)


def test_large_image_8192(ctx):
    """8192 x 8192 (67 M pixels, box 64): FP64 against the oracle on sampled
    pixels including the four frame edges, split-BF16 against FP64."""
    rng = np.random.default_rng(8192)
    n = 8192
    a, b = smooth_pair(rng, n, n, shift=(0.45, -0.3), width=3.0)
    g = ctx.track(a, b, 1.0, 1.0, 15.0)
    assert ctx.last_stats()["tracked"] == n * n
    pix = np.concatenate([rng.choice(n * n, 40, replace=False),
                          [0, n - 1, n * (n - 1), n * n - 1, n * (n // 2), n * (n // 2) + n - 1]])
    o = O.flct_pixels(a, b, 1.0, 1.0, 15.0, pix)
    assert_parity((g[0].ravel()[pix], g[1].ravel()[pix], g[2].ravel()[pix]), o, where="8192^2")
    s = ctx.track(a, b, 1.0, 1.0, 15.0, precision="split_bf16")
    assert np.array_equal(s[2], g[2])
    e = np.hypot(s[0] - g[0], s[1] - g[1])
    rms = float(np.sqrt(np.mean(e ** 2)))
    assert rms <= 1e-4 and float(e.max()) <= 5e-3, (rms, float(e.max()))
