"""Split-BF16 MFMA form against the FP64 plain form (GPU).

The precision form replaces the per-pixel FFTs by DFT-matrix contractions in
BF16 hi/lo operands with FP32 accumulation; window, means, peak fit and bias
stay FP64.  Its contract: the vm mask and the thresholded / skipped pixel set
are bit-exact (they never depend on the correlation surface), and the RMS
velocity error against the FP64 form is <= 1e-4 px.  Pixels whose BF16
surface has a near tie or an ill-conditioned fit are recomputed in FP64, so
the largest error is bounded too (5e-3 px here, observed ~1.2e-3).
"""
import numpy as np
import pytest

from pyflct_amd import synth

pytestmark = pytest.mark.gpu

RMS_TOL = 1e-4   # px, RMS over all tracked pixels
MAX_TOL = 5e-3   # px, any pixel

# (n, sigma, options): boxes up to 64 use the row-segment kernel for interior
# columns and a per-pixel kernel (radix-2 for 36..64, dense DFT below) for
# the clamped edge columns and the bias pass; box 84 runs the FP32 kernels
# by default (test_bf16_dense_kernel_large_boxes covers the BF16 kernel there)
CASES = [
    (192, 5.0, {}),                         # N = 20
    (256, 9.0, {}),                         # N = 36
    (256, 12.0, {"kr": 0.5}),               # N = 48, filter
    (384, 15.0, {}),                        # N = 64
    (256, 12.0, {"bias": True}),            # bias pass
    (384, 15.0, {"thresh": 0.2, "skip": 3}),  # thresholded lattice
    (256, 20.0, {}),                        # N = 84
]


def _errs(a, b):
    e = np.hypot(b[0] - a[0], b[1] - a[1])
    return float(np.sqrt(np.mean(e ** 2))), float(e.max())


@pytest.mark.parametrize("n,sigma,kw", CASES, ids=[f"n{c[0]}-s{c[1]:g}-{'-'.join(c[2]) or 'plain'}"
                                                    for c in CASES])
def test_bf16_matches_fp64(ctx, n, sigma, kw):
    f1, f2, _ = synth.cfg4(n=n)
    ref = ctx.track(f1, f2, 1.0, 1.0, sigma, **kw)
    st_ref = ctx.last_stats()
    got = ctx.track(f1, f2, 1.0, 1.0, sigma, precision="split_bf16", **kw)
    st = ctx.last_stats()
    assert st["tracked"] == st_ref["tracked"]
    assert np.array_equal(got[2], ref[2])
    assert not np.isnan(got[0]).any() and not np.isnan(got[1]).any()
    rms, mx = _errs(ref, got)
    assert rms <= RMS_TOL, (rms, mx)
    assert mx <= MAX_TOL, (rms, mx)


@pytest.mark.parametrize("n,sigma", [(256, 16.0), (256, 17.0), (192, 5.0)])
def test_bf16_dense_kernel_large_boxes(monkeypatch, n, sigma):
    """The per-pixel dense-DFT BF16 kernel at boxes 68, 72 and 20 (it is
    eligible up to 72), which the untuned path rule does not pick above 64,
    against FP64."""
    from pyflct_amd import api
    monkeypatch.setenv("FLCT_HIP_BF16_KERNEL", "pixel")
    c = api.Context(0)
    try:
        f1, f2, _ = synth.cfg4(n=n)
        ref = c.track(f1, f2, 1.0, 1.0, sigma)
        got = c.track(f1, f2, 1.0, 1.0, sigma, precision="split_bf16")
        assert c.last_path() == "bf16-pixel"
        assert np.array_equal(got[2], ref[2])
        rms, mx = _errs(ref, got)
        assert rms <= RMS_TOL and mx <= MAX_TOL, (rms, mx)
    finally:
        c.close()


def test_bf16_refinement_settles_outliers(ctx):
    """On the full-disk field at 1024^2 the BF16 surfaces of about 0.1 % of
    the pixels have near ties (e.g. peaks half-way between two cells) whose
    fit flips between neighbours; those are recomputed in FP64."""
    f1, f2, _ = synth.cfg4(n=1024)
    ref = ctx.track(f1, f2, 1.0, 1.0, 15.0)
    got = ctx.track(f1, f2, 1.0, 1.0, 15.0, precision="split_bf16")
    st = ctx.last_stats()
    assert 0 < st["refined"] < 0.01 * st["tracked"]
    rms, mx = _errs(ref, got)
    assert rms <= RMS_TOL and mx <= MAX_TOL, (rms, mx)


@pytest.mark.parametrize("sigma,shift", [(15.0, 31), (15.0, -32), (9.0, 17)])
def test_segment_peak_at_surface_edge(ctx, sigma, shift):
    """Shifts of Nh - 1 and -Nh cells (box 64; Nh - 1 at box 36) put the
    true peak on the border of the correlation surface, where no 3x3 fit runs
    (flct.c:1536-1556) and the velocity is the peak cell itself: the cell the
    row-segment kernel locates from its summed ballots must be FP64's."""
    from conftest import smooth_pair
    rng = np.random.default_rng(400 + shift)
    a, _ = smooth_pair(rng, 200, 232, width=2.0)
    b = np.roll(a, (shift, -shift), axis=(0, 1))
    ref = ctx.track(a, b, 1.0, 1.0, sigma)
    got = ctx.track(a, b, 1.0, 1.0, sigma, precision="split_bf16")
    assert ctx.last_path() == "bf16-segment"
    assert np.array_equal(got[2], ref[2])
    rms, mx = _errs(ref, got)
    assert rms <= RMS_TOL and mx <= MAX_TOL, (rms, mx)


@pytest.mark.parametrize("nx,ny,sigma,kw", [
    (150, 190, 15.0, {}),                                  # box 64, clamped rows and columns
    (64, 100, 15.0, {}),                                   # nx = N: one unclamped row
    (97, 120, 9.0, {"thresh": 0.3}),                       # box 36, relative threshold
    (180, 170, 7.0, {"skip": 3, "poff": 1, "qoff": 2}),    # box 28 (single tile), skip lattice
])
def test_segment_edge_columns_through_strips(ctx, nx, ny, sigma, kw):
    """The clamped edge columns of the split-BF16 form run through the
    row-segment kernel on transposed column strips (their boxes and windows,
    flct.c:773-870, are those of interior columns there), the four corners
    through the per-pixel kernel.  Against FP64 on the edge columns alone:
    the mask bit-exact, the velocities within the form's bounds."""
    from conftest import smooth_pair
    rng = np.random.default_rng(nx * 1000 + ny)
    a, b = smooth_pair(rng, nx, ny, shift=(0.7, -0.4), width=2.0)
    ref = ctx.track(a, b, 1.0, 1.0, sigma, **kw)
    got = ctx.track(a, b, 1.0, 1.0, sigma, precision="split_bf16", **kw)
    assert ctx.last_path() == "bf16-segment"
    assert np.array_equal(got[2], ref[2])
    nbox = ctx.box_size(nx, ny, sigma)
    edge = np.zeros((nx, ny), bool)
    edge[:, :nbox // 2 - 1] = True
    edge[:, ny - nbox // 2:] = True
    tracked = edge & (ref[2] == 1.0)
    assert tracked.sum() > 0
    e = [ref[k][tracked] - got[k][tracked] for k in (0, 1)]
    rms = float(np.sqrt(np.mean(e[0] ** 2 + e[1] ** 2)))
    mx = float(max(np.max(np.abs(e[0])), np.max(np.abs(e[1]))))
    assert rms <= RMS_TOL and mx <= MAX_TOL, (rms, mx)


def test_bf16_scale_invariance(ctx):
    """Inputs are scaled by powers of two before the split; results must not
    depend on the overall amplitude (tiny and huge images)."""
    rng = np.random.default_rng(11)
    from conftest import smooth_pair
    a, b = smooth_pair(rng, 160, 176)
    base = ctx.track(a, b, 1.0, 1.0, 12.0, precision="split_bf16")
    for s in (2.0 ** -60, 2.0 ** 70):
        got = ctx.track(a * s, b * s, 1.0, 1.0, 12.0, precision="split_bf16")
        assert np.array_equal(got[0], base[0]) and np.array_equal(got[1], base[1])
        assert np.array_equal(got[2], base[2])


def test_bf16_other_forms_report_no_refinement(ctx):
    f1, f2, _ = synth.cfg4(n=192)
    ctx.track(f1, f2, 1.0, 1.0, 9.0)
    assert ctx.last_stats()["refined"] == 0


# non-square and odd shapes for the row-segment kernel: narrow interior
# widths (ny - N + 1), partial last bands, edge rows and columns
SEG_SHAPES = [
    (200, 70, 15.0, {}),                          # N = 44, 27 interior columns
    (97, 301, 15.0, {}),                          # N = 64, bands of 60 and 61 columns
    (150, 131, 10.0, {"kr": 0.4}),                # N = 40 with the filter
    (180, 210, 15.0, {"thresh": 0.3}),            # relative threshold
    (170, 190, 11.0, {"skip": 4, "poff": 1, "qoff": -2, "interp": True}),  # N = 44 lattice
    # boxes up to 32 use the single-tile variant
    (60, 77, 2.0, {}),                            # N = 8
    (70, 81, 3.0, {"kr": 0.6}),                   # N = 12 with the filter
    (100, 120, 4.0, {}),                          # N = 16
    (90, 101, 5.0, {"kr": 0.45}),                 # N = 20 with the filter
    (110, 97, 7.0, {"thresh": 0.25}),             # N = 28, relative threshold
    (140, 150, 8.0, {}),                          # N = 32
]


@pytest.mark.parametrize("nx,ny,sigma,kw", SEG_SHAPES,
                         ids=[f"{c[0]}x{c[1]}-s{c[2]:g}" for c in SEG_SHAPES])
def test_bf16_segment_shapes(ctx, nx, ny, sigma, kw):
    rng = np.random.default_rng(nx * 1000 + ny)
    from conftest import smooth_pair
    a, b = smooth_pair(rng, nx, ny, shift=(0.7, -1.2), width=2.5)
    ref = ctx.track(a, b, 1.0, 1.0, sigma, **kw)
    nbox = ctx.box_size(nx, ny, sigma)
    ctx.set_path("split_bf16", nbox, "bf16-segment")  # also below box 20, where FP32 is the default
    try:
        got = ctx.track(a, b, 1.0, 1.0, sigma, precision="split_bf16", **kw)
        assert ctx.last_path() == "bf16-segment"
    finally:
        ctx.set_path("split_bf16", nbox, None)
    assert np.array_equal(got[2], ref[2])
    m = ref[2] == 1.0
    rms, mx = _errs((ref[0][m], ref[1][m]), (got[0][m], got[1][m]))
    assert rms <= RMS_TOL and mx <= MAX_TOL, (rms, mx)
    # thresholded and unfilled pixels carry the reference's values exactly;
    # filled pixels (vm = 0.5) interpolate the tracked velocities
    z = ref[2] == 0.0
    assert np.array_equal(got[0][z], ref[0][z]) and np.array_equal(got[1][z], ref[1][z])
    h = ref[2] == 0.5
    if h.any():
        rms, mx = _errs((ref[0][h], ref[1][h]), (got[0][h], got[1][h]))
        assert rms <= RMS_TOL and mx <= MAX_TOL, (rms, mx)


def test_bf16_segment_matches_pixel_kernel(monkeypatch):
    """The row-segment kernel and the per-pixel radix-2 kernel compute the
    same surfaces in a different order: their velocities agree to the split
    precision."""
    from pyflct_amd import api
    f1, f2, _ = synth.cfg4(n=384)
    seg = api.Context(0)
    monkeypatch.setenv("FLCT_HIP_BF16_KERNEL", "pixel")
    pix = api.Context(0)
    try:
        g = seg.track(f1, f2, 1.0, 1.0, 15.0, precision="split_bf16")
        p = pix.track(f1, f2, 1.0, 1.0, 15.0, precision="split_bf16")
    finally:
        seg.close()
        pix.close()
    assert np.array_equal(g[2], p[2])
    rms, mx = _errs(g, p)
    assert rms <= RMS_TOL and mx <= MAX_TOL, (rms, mx)


def test_set_path_rules(ctx):
    from pyflct_amd._lib import FlctHipError
    for args in (("fp64", 64, "bf16-segment"), ("fp32", 64, "bf16-pixel"), ("split_bf16", 66, "fp64-fft"),
                 ("split_bf16", 64, "no-such-path"), ("split_bf16", 84, "bf16-segment"),
                 ("split_bf16", 76, "bf16-pixel")):
        with pytest.raises(FlctHipError):
            ctx.set_path(*args)
    f1, f2, _ = synth.cfg4(n=192)
    ctx.set_path("split_bf16", 64, "fp64-fft")
    ctx.set_path("fp64", 64, "fp64-fft")
    try:
        got = ctx.track(f1, f2, 1.0, 1.0, 15.0, precision="split_bf16")
        assert ctx.last_path() == "fp64-fft"
        ref = ctx.track(f1, f2, 1.0, 1.0, 15.0)
        for x, y in zip(got, ref):
            assert np.array_equal(x, y)
    finally:
        ctx.set_path("split_bf16", 64, None)
        ctx.set_path("fp64", 64, None)
    ctx.track(f1, f2, 1.0, 1.0, 15.0, precision="split_bf16")
    assert ctx.last_path() == "bf16-segment"


def test_paths_and_autotune():
    """Untuned paths; the autotuner picks among paths that meet the form's
    contract and later calls of that form and box size use its choice."""
    from pyflct_amd import api
    c = api.Context(0)
    try:
        f1, f2, _ = synth.cfg4(n=384)
        ref = c.track(f1, f2, 1.0, 1.0, 15.0)
        assert c.last_path() == "fp64-rowseg"  # FP64 default at box 64
        c.track(f1, f2, 1.0, 1.0, 15.0, precision="split_bf16")
        assert c.last_path() == "bf16-segment"
        b84 = c.track(f1, f2, 1.0, 1.0, 20.0, precision="split_bf16")  # box 84
        assert c.last_path() == "fp32-rowseg"
        r84 = c.track(f1, f2, 1.0, 1.0, 20.0)
        assert np.array_equal(b84[2], r84[2])
        rms, mx = _errs(r84, b84)
        assert rms <= RMS_TOL and mx <= MAX_TOL, (rms, mx)
        d = [c.malloc(f1.nbytes) for _ in range(2)]
        try:
            c.upload(d[0], f1)
            c.upload(d[1], f2)
            for form, sigma, allowed in (("split_bf16", 15.0, {"bf16-segment", "bf16-pixel", "fp32-fft",
                                                               "fp32-rowseg", "fp64-fft", "fp64-rowseg"}),
                                         ("fp32", 9.0, {"fp32-fft", "fp32-rowseg", "fp64-fft", "fp64-rowseg"}),
                                         ("fp64", 12.0, {"fp64-fft", "fp64-rowseg"})):
                best = c.autotune(d[0], d[1], 384, 384, c.make_opts(1.0, 1.0, sigma, precision=form))
                assert best in allowed, (form, best)
                got = c.track(f1, f2, 1.0, 1.0, sigma, precision=form)
                assert c.last_path() == best
                r = ref if sigma == 15.0 else c.track(f1, f2, 1.0, 1.0, sigma)
                assert np.array_equal(got[2], r[2])
                rms, _ = _errs(r, got)
                assert rms <= (RMS_TOL if form == "split_bf16" else 2e-3), (form, best, rms)
            # tuning with -h (FP64 kernels only) or -b stores nothing for the
            # plain calls of that form and box size
            c.set_path("split_bf16", 40, None)
            c.autotune(d[0], d[1], 384, 384, c.make_opts(1.0, 1.0, 10.0, precision="split_bf16",
                                                        hires=True))
            c.autotune(d[0], d[1], 384, 384, c.make_opts(1.0, 1.0, 10.0, precision="split_bf16",
                                                        bias=True))
            c.track(f1, f2, 1.0, 1.0, 10.0, precision="split_bf16")
            assert c.last_path() == "bf16-segment"
        finally:
            for p_ in d:
                c.free(p_)
    finally:
        c.close()


# every box size 4..96 in the FP32 and split-BF16 forms (their default kernel
# paths: FP32 kernels for every box in FP32; in split-BF16 the FP32 kernels up
# to box 16 and above 64, the row-segment kernel from 20 to 64)
BOXES = list(range(4, 100, 4))
# boxes whose untuned FP32 kernel (FP32 form; split-BF16 outside 24..64) is
# the row-shared one (fp32-rowseg) when no threshold is given
FP32_ROWSEG_BOXES = {8, 12, 16, 28, 32, 36, 40, 44, 48, 52, 56, 60, 64, 68, 72, 76, 80, 84, 88, 92, 96}


def _box_case(nbox):
    """(sigma, nx, ny) giving box nbox: an integer sigma whose nt (flct.c:567-580,
    sigma truncated to an integer) gives it, else sigma 14 with the image's
    size clamping the box (only box 56 has no integer sigma)."""
    import math
    for s in range(1, 24):
        nt = int(s * math.sqrt(math.log(100.0)))
        if 2 * ((nt // 2) * 2) == nbox:
            n = max(64, 3 * nbox + 6)
            return float(s), n, n + 5
    return 14.0, 3 * (nbox // 2) + 1, 3 * (nbox // 2) + 6


@pytest.mark.parametrize("form", ["fp32", "split_bf16"])
@pytest.mark.parametrize("nbox", BOXES, ids=[f"N{b}" for b in BOXES])
def test_forms_every_box_size(ctx, form, nbox):
    sigma, nx, ny = _box_case(nbox)
    assert ctx.box_size(nx, ny, sigma) == nbox
    rng = np.random.default_rng(nbox)
    from conftest import smooth_pair
    a, b = smooth_pair(rng, nx, ny, shift=(0.6, -0.9), width=max(1.5, nbox / 16))
    ref = ctx.track(a, b, 1.0, 1.0, sigma)
    got = ctx.track(a, b, 1.0, 1.0, sigma, precision=form)
    path = ctx.last_path()
    assert np.array_equal(got[2], ref[2])
    rms, mx = _errs(ref, got)
    assert rms <= RMS_TOL and mx <= MAX_TOL, (path, rms, mx)


def test_pixel_kernels_every_box_size(monkeypatch):
    """The per-pixel split-BF16 kernels (FLCT_HIP_BF16_KERNEL=pixel: the
    dense-DFT kernel and the radix-2 kernel for boxes 36..64) at every box
    size up to 72 against FP64; above 72, where their error can exceed the
    form's contract, the request runs the FP32 kernels."""
    from pyflct_amd import api
    from conftest import smooth_pair
    monkeypatch.setenv("FLCT_HIP_BF16_KERNEL", "pixel")
    c = api.Context(0)
    try:
        for nbox in BOXES:
            sigma, nx, ny = _box_case(nbox)
            rng = np.random.default_rng(900 + nbox)
            a, b = smooth_pair(rng, nx, ny, shift=(-0.7, 0.4), width=max(1.5, nbox / 16))
            ref = c.track(a, b, 1.0, 1.0, sigma)
            got = c.track(a, b, 1.0, 1.0, sigma, precision="split_bf16")
            fp32_path = "fp32-rowseg" if nbox in FP32_ROWSEG_BOXES else "fp32-fft"
            assert c.last_path() == ("bf16-pixel" if nbox <= 72 else fp32_path), (nbox, c.last_path())
            assert np.array_equal(got[2], ref[2]), nbox
            rms, mx = _errs(ref, got)
            assert rms <= RMS_TOL and mx <= MAX_TOL, (nbox, rms, mx)
    finally:
        c.close()


# ------------------------------------------------ against the CPU oracle
def _oracle_sample(nx, ny, rng, nrand=300):
    """Frame corners, every 5th frame-edge pixel and random pixels."""
    i, j = np.mgrid[0:nx, 0:ny]
    edge = np.flatnonzero((i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1))
    return np.unique(np.concatenate([edge[::5], [0, ny - 1, (nx - 1) * ny, nx * ny - 1],
                                     rng.choice(nx * ny, nrand, replace=False)]))


def _vs_oracle(g, o, pix, where):
    vx, vy, vm = (a.ravel()[pix] for a in g)
    assert np.array_equal(vm, o[2]), where
    e = np.hypot(vx - o[0], vy - o[1])
    rms = float(np.sqrt(np.mean(e ** 2)))
    assert rms <= RMS_TOL and e.max() <= MAX_TOL, (where, rms, float(e.max()))
    return rms


# one box per kernel family: 20 (single-tile row segment; dense kernel
# NP = 32), 32 (single-tile row segment at its limit), 48 (row segment;
# dense kernel NP = 64) and 64 (row segment; dense kernel NP = 64, radix-2
# edge kernel); every BF16 path the box allows, and the default path
@pytest.mark.parametrize("sigma", [5.0, 8.0, 12.0, 15.0])
def test_bf16_paths_against_oracle(ctx, sigma):
    import oracle_lib as O
    f1, f2, _ = synth.cfg4(n=320)
    nx, ny = f1.shape
    nbox = ctx.box_size(nx, ny, sigma)
    pix = _oracle_sample(nx, ny, np.random.default_rng(int(sigma)))
    o = O.flct_pixels(f1, f2, 1.0, 1.0, sigma, pix)
    for path in (None, "bf16-segment", "bf16-pixel"):
        ctx.set_path("split_bf16", nbox, path)
        try:
            g = ctx.track(f1, f2, 1.0, 1.0, sigma, precision="split_bf16")
            if path is not None:
                assert ctx.last_path() == path
            _vs_oracle(g, o, pix, f"box {nbox} {ctx.last_path()}")
        finally:
            ctx.set_path("split_bf16", nbox, None)


def test_bf16_full_disk_4096_against_oracle(ctx):
    """The 4096^2, sigma = 15 full-disk pair in split-BF16: sampled pixels
    (interior, limb, frame corners and edges) against the oracle."""
    import oracle_lib as O
    f1, f2, _ = synth.cfg4()
    n = f1.shape[0]
    g = ctx.track(f1, f2, 1.0, 1.0, 15.0, precision="split_bf16")
    assert ctx.last_path() == "bf16-segment"
    rng = np.random.default_rng(14)
    pix = np.unique(np.concatenate([_oracle_sample(n, n, rng, 250)[::3],
                                    np.arange(n * 2048, n * 2048 + n, 61),  # across the disk
                                    n * np.arange(0, n, 61) + 2048]))
    o = O.flct_pixels(f1, f2, 1.0, 1.0, 15.0, pix)
    rms = _vs_oracle(g, o, pix, "4096^2 sigma 15")
    assert rms <= 5e-5, rms


@pytest.mark.parametrize("form", ["fp32", "split_bf16"])
def test_pedestal_image_against_oracle(ctx, form):
    """An intensity-like pair with a large constant level (1e4 plus 1 %
    structure): the box means remove the level in FP64 before any
    reduced-precision arithmetic, so the error contract holds."""
    import oracle_lib as O
    from conftest import smooth_pair
    rng = np.random.default_rng(15)
    a, b = smooth_pair(rng, 240, 260, shift=(0.7, -0.4), width=2.5)
    s = 100.0 / np.std(a)
    a, b = 1e4 + s * a, 1e4 + s * b
    for sigma in (6.0, 12.0, 15.0):
        ref = ctx.track(a, b, 1.0, 1.0, sigma)
        g = ctx.track(a, b, 1.0, 1.0, sigma, precision=form)
        assert np.array_equal(g[2], ref[2])
        rms, mx = _errs(ref, g)
        assert rms <= RMS_TOL and mx <= MAX_TOL, (form, sigma, rms, mx)
        pix = _oracle_sample(*a.shape, rng, 150)
        _vs_oracle(g, O.flct_pixels(a, b, 1.0, 1.0, sigma, pix), pix, f"{form} sigma {sigma}")
