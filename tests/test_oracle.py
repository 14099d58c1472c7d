"""CPU checks that pin the oracle (oracle/flct_oracle.c) before it is
trusted as the parity reference for the GPU path."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import f32, smooth_pair


# ---------------------------------------------------------------- fixtures
def test_flct_reproduces_reference_fixture(golden):
    """flct hashgauss.dat testvels.dat 1 1 5 -k 0.5 (README.md:207-208) ->
    tests/golden/testgaussvel.dat, at the file's float32 precision."""
    vx, vy, vm = O.flct(golden["f1"], golden["f2"], 1.0, 1.0, 5.0, kr=0.5)
    assert np.array_equal(vm, golden["vm"])
    for a, r in ((vx, golden["vx"]), (vy, golden["vy"])):
        same = np.sum(f32(a) == r)
        # values that differ are ~1e-8 in size and differ by a float32 ulp
        assert same >= 0.998 * r.size
        assert np.max(np.abs(f32(a) - r)) < 1e-11


def test_warp_reproduces_reference_fixture(golden):
    """hashgauss f2 = shift_frac2d.pro(f1, deltax, deltay) (README.md:257-261),
    stored as float32: warp_frac2d reproduces it to the fixture precision."""
    w = O.warp_frac2d(golden["f1"], golden["dx"], golden["dy"])
    assert np.max(np.abs(w - golden["f2"])) < 5e-7
    assert np.sqrt(np.mean((w - golden["f2"]) ** 2)) < 1e-7


# ---------------------------------------------------------------- closed forms
def test_make_freq_closed_form():
    # flct.c:1651-1692: 0..n/2 (Nyquist positive) then negatives; odd n has +-n/2
    assert list(O.make_freq(8)) == [0, 1, 2, 3, 4, -3, -2, -1]
    assert list(O.make_freq(5)) == [0, 1, 2, -2, -1]
    assert list(O.make_freq(1)) == [0]
    for n in range(2, 40):
        k = O.make_freq(n)
        ref = np.fft.fftfreq(n, 1.0 / n)
        if n % 2 == 0:
            ref[n // 2] = n // 2
        assert np.array_equal(k, ref)


APPENDIX_B = {4: 16, 5: 20, 6: 24, 7: 28, 8: 32, 9: 36, 10: 40, 11: 44, 12: 48, 13: 52,
              14: 60, 15: 64, 16: 68, 17: 72, 18: 76, 19: 80, 20: 84}


def test_box_size_table():
    # flct.c:567-580: N = 2 * min((nt/2)*2, ndmin), nt = (int)((int)sigma * sqrt(ln 100))
    for sig, n in APPENDIX_B.items():
        assert 2 * O.box_nsize(sig, 4096, 4096) == n
        assert 2 * O.box_nsize(sig + 0.7, 4096, 4096) == n  # sigma truncated first
    assert 2 * O.box_nsize(15, 90, 4096) == 2 * ((90 // 3) // 2) * 2  # ndmin clamp
    assert O.box_nsize(15, 5, 5) == -1  # "illegal box size"


def test_fft_against_numpy_and_brute_force():
    rng = np.random.default_rng(0)
    for shape in [(16, 16), (20, 28), (64, 64), (84, 76), (9, 7), (101, 201)]:
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        ref = np.fft.fft2(a)
        got = O.fft2(a, -1)
        assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))
        back = O.fft2(got, +1) / a.size
        assert np.max(np.abs(back - a)) < 1e-13
    # brute-force DFT on a tiny odd/prime size
    a = rng.standard_normal((5, 7)) + 1j * rng.standard_normal((5, 7))
    i, k = np.arange(5), np.arange(7)
    D1 = np.exp(-2j * np.pi * np.outer(i, i) / 5)
    D2 = np.exp(-2j * np.pi * np.outer(k, k) / 7)
    assert np.max(np.abs(O.fft2(a, -1) - D1 @ a @ D2.T)) < 1e-12


def test_peak_fit_on_sampled_quadratic():
    """The 3x3 Taylor fit (flct.c:1536-1556) returns the exact vertex of a
    sampled 2-D quadratic with a cross term."""
    n = 16
    x0, y0 = 8.3, 7.6
    i, j = np.mgrid[0:n, 0:n].astype(float)
    dx, dy = i - x0, j - y0
    surf = 100.0 - (1.3 * dx * dx + 0.4 * dx * dy + 0.9 * dy * dy)
    sx, sy = O.peak_fit(surf)
    assert abs(sx - (x0 - n // 2)) < 1e-12 and abs(sy - (y0 - n // 2)) < 1e-12


def test_peak_edge_and_zero_rules():
    n = 12
    s = np.zeros((n, n))
    assert O.peak_fit(s) == (0.0, 0.0)  # max == 0 -> centre, no fit
    s[0, 5] = 1.0  # edge peak: integer shift, no fit
    assert O.peak_fit(s) == (0.0 - n // 2, 5.0 - n // 2)
    s[:] = 0.0
    s[3, 4] = s[7, 2] = 2.0  # tie: the first in row-major order wins
    sx, sy = O.peak_fit(s)
    assert np.floor(sx + n // 2 + 0.5) == 3 and np.floor(sy + n // 2 + 0.5) == 4


def test_hires_peak_on_sampled_quadratic():
    """-h (flct.c:1564-1645): cubic convolution reproduces a quadratic
    exactly away from the border, so the 0.1 px grid holds the quadratic and
    the 3 x 3 fit there, scaled to pixels, returns its exact vertex; the
    vertex is placed between grid points so that the fit's scale matters."""
    n = 16
    for x0, y0 in ((8.33, 7.57), (7.96, 8.04), (6.71, 9.38)):
        i, j = np.mgrid[0:n, 0:n].astype(float)
        dx, dy = i - x0, j - y0
        surf = 100.0 - (1.3 * dx * dx + 0.4 * dx * dy + 0.9 * dy * dy)
        sx, sy = O.peak_hires(surf)
        assert abs(sx - (x0 - n // 2)) < 1e-10 and abs(sy - (y0 - n // 2)) < 1e-10, (x0, y0)
        # the coarse fit agrees on this surface (both exact for quadratics)
        cx, cy = O.peak_fit(surf)
        assert abs(cx - sx) < 1e-10 and abs(cy - sy) < 1e-10


def test_hires_zero_and_edge_rules():
    """The grid is interpolated with xmiss = 0 (flct.c:1589-1590): grid points
    outside the surface, and stencils touching a zero, read 0; an all-zero
    grid has its first point as maximum, so the shift is the grid corner
    (-1, -1) about the integer peak, which is the centre when the surface is
    zero (flct.c:1518-1530)."""
    n = 12
    assert O.peak_hires(np.zeros((n, n))) == (-1.0, -1.0)
    s = np.zeros((n, n))
    s[0, 5] = 1.0  # a lone edge peak: every stencil touches a zero
    assert O.peak_hires(s) == (-1.0 - n // 2, 5.0 - 1.0 - n // 2)
    # a smooth peak: the grid maximum is interior and near the true centre
    i, j = np.mgrid[0:n, 0:n].astype(float)
    g = np.exp(-((i - 5.37) ** 2 + (j - 6.81) ** 2) / 8.0) + 0.1
    sx, sy = O.peak_hires(g)
    assert abs(sx - (5.37 - n // 2)) < 0.05 and abs(sy - (6.81 - n // 2)) < 0.05


def test_flct_hires_tracks_a_uniform_shift():
    """flct with -h on a uniformly shifted image: the velocities follow the
    shift like the plain fit's (both carry the windowed correlation's
    underestimate at sigma = 6), and differ from it (the peak step is really
    replaced)."""
    rng = np.random.default_rng(21)
    from scipy.ndimage import gaussian_filter
    g = gaussian_filter(rng.standard_normal((48, 52)), 2.0, mode="wrap")
    kx = np.fft.fftfreq(48)[:, None]
    ky = np.fft.fftfreq(52)[None, :]
    h = np.fft.ifft2(np.fft.fft2(g) * np.exp(-2j * np.pi * (kx * 0.4 + ky * -0.25))).real
    hv = O.flct(g, h, 1.0, 1.0, 6.0, hires=True)
    pv = O.flct(g, h, 1.0, 1.0, 6.0)
    sl = (slice(12, 36), slice(12, 40))
    for k, true in ((0, -0.25), (1, 0.4)):
        mh, mp = np.median(hv[k][sl]), np.median(pv[k][sl])
        assert abs(mh - mp) < 0.05 and 0.6 * abs(true) < abs(mh) <= abs(true) and mh * true > 0
        assert np.max(np.abs(hv[k][sl] - pv[k][sl])) < 0.1
    assert not np.array_equal(hv[0], pv[0])


def test_interpcc2d_reproduces_quadratics_and_border_rule():
    """Keys cubic convolution (a = -0.5) is exact for quadratics in the
    interior; the one-cell border uses f[2] - 3 f[1] + 3 f[0] (flct.c:2294)."""
    nx, ny = 9, 11
    i, j = np.mgrid[0:nx, 0:ny].astype(float)
    q = 0.5 + 0.3 * i - 0.2 * j + 0.05 * i * i + 0.02 * i * j - 0.03 * j * j
    xw = np.linspace(1.0, nx - 2.0, 13)
    yw = np.linspace(1.0, ny - 2.0, 17)
    out = O.interpcc2d(q, -999999.0, xw, yw)
    X, Y = np.meshgrid(xw, yw, indexing="ij")
    exact = 0.5 + 0.3 * X - 0.2 * Y + 0.05 * X * X + 0.02 * X * Y - 0.03 * Y * Y
    assert np.max(np.abs(out - exact)) < 1e-12
    # out of range -> xmiss
    o2 = O.interpcc2d(q, -7.0, np.array([-0.1, nx - 1 + 0.1]), np.array([1.0]))
    assert np.all(o2 == -7.0)
    # border extrapolation by hand at x = 0.5 (needs cdata row 0)
    f = np.arange(nx * ny, dtype=float).reshape(nx, ny) ** 1.5
    x, y = 0.5, 4.0
    got = O.interpcc2d(f, -999999.0, np.array([x]), np.array([y]))[0, 0]

    def keys(t):
        t = abs(t)
        return (((-0.5 * t + 2.5) * t) - 4.0) * t + 2.0 if t >= 1 else (1.5 * t - 2.5) * t * t + 1.0

    col = lambda r: f[r, :] if r >= 0 else f[2, :] - 3.0 * f[1, :] + 3.0 * f[0, :]
    exp = sum(col(ii)[jj] * keys(x - ii) * keys(y - jj) for ii in (-1, 0, 1, 2) for jj in (3, 4, 5, 6))
    assert abs(got - exp) < 1e-9 * abs(exp)
    # a missing value in the 4x4 stencil poisons the result
    g = q.copy()
    g[4, 5] = -999999.0
    o3 = O.interpcc2d(g, -999999.0, np.array([4.2, 1.0]), np.array([5.3, 1.0]))
    assert o3[0, 0] == -999999.0 and o3[1, 1] != -999999.0


# ---------------------------------------------------------------- invariants
def test_identical_and_negated_images_give_zero_velocity():
    rng = np.random.default_rng(3)
    a, _ = smooth_pair(rng, 40, 44)
    for b in (a, -a):
        vx, vy, vm = O.flct(a, b, 1.0, 1.0, 4.0)
        assert np.all(vm == 1.0)
        assert np.max(np.abs(vx)) < 1e-10 and np.max(np.abs(vy)) < 1e-10


def test_swapped_images_reverse_velocity():
    rng = np.random.default_rng(4)
    a, b = smooth_pair(rng, 48, 48)
    v1 = O.flct(a, b, 1.0, 1.0, 5.0)
    v2 = O.flct(b, a, 1.0, 1.0, 5.0)
    # interior peaks: the fit is symmetric under f1 <-> f2
    sel = (np.abs(v1[0]) < 3) & (np.abs(v1[1]) < 3)
    assert np.max(np.abs(v1[0][sel] + v2[0][sel])) < 1e-9
    assert np.max(np.abs(v1[1][sel] + v2[1][sel])) < 1e-9


def test_integer_circular_shift_is_recovered():
    rng = np.random.default_rng(5)
    from scipy.ndimage import gaussian_filter
    g = gaussian_filter(rng.standard_normal((24, 24)), 1.5, mode="wrap")
    h = np.roll(np.roll(g, 2, axis=0), -3, axis=1)
    sx, sy, _ = O.cross_cor(g, h)
    assert abs(sx - 2.0) < 0.3 and abs(sy + 3.0) < 0.3


def test_zero_images_hit_the_max_zero_branch():
    z = np.zeros((40, 40))
    vx, vy, vm = O.flct(z, z, 1.0, 1.0, 4.0)
    assert np.all(vx == 0) and np.all(vy == 0) and np.all(vm == 1)


def test_uniform_shift_recovers_known_velocity():
    """Image 2 moved by (+0.3, -0.2) px in IDL (x, y): vx > 0, vy < 0 with
    FLCT's known bias toward zero (Fisher & Welsch 2008, p. 379-380)."""
    from pyflct_amd import synth
    f1, f2 = synth.cfg1()
    vx, vy, vm = O.flct(f1, f2, 1.0, 1.0, 8.0)
    core = (slice(32, 96), slice(32, 96))
    assert 0.1 < np.median(vx[core]) < 0.3
    assert -0.2 < np.median(vy[core]) < -0.05


def test_bias_correction_reduces_bias():
    """-b (window offset by -+d/2, no reference counterpart): on uniform
    shifts of a smooth periodic field the median error shrinks."""
    from scipy.ndimage import gaussian_filter
    from pyflct_amd import synth
    rng = np.random.default_rng(1)
    a = gaussian_filter(rng.standard_normal((96, 96)), 2.0, mode="wrap")
    c = (slice(24, 72), slice(24, 72))
    for d in (0.5, 1.5, 2.5):
        b = synth.shift_uniform(a, d, -0.5 * d)
        e = []
        for bias in (False, True):
            vx, vy, _ = O.flct(a, b, 1.0, 1.0, 6.0, bias=bias)
            e.append(abs(np.median(vx[c]) - d) + abs(np.median(vy[c]) + 0.5 * d))
        assert e[1] < 0.75 * e[0]
    # zero motion stays zero
    vx, vy, _ = O.flct(a, a, 1.0, 1.0, 6.0, bias=True)
    assert np.max(np.abs(vx)) < 1e-10 and np.max(np.abs(vy)) < 1e-10


def _direct_estimate(g1, g2):
    """One cross_cor estimate by direct sums, no FFT: c(m) = sum_x g1(x)
    g2(x + m) over the box taken periodically (the c2r of conj(A) B), |c|
    moved by N/2 in both axes, the first maximum in row-major order, the
    3x3 Taylor fit for interior peaks (paper eqs. 5-8) and the centre for
    an all-zero surface.  Returns (rows, columns) shifts."""
    n = g1.shape[0]
    c = np.empty((n, n))
    for mx in range(n):
        for my in range(n):
            c[mx, my] = np.sum(g1 * np.roll(g2, (-mx, -my), axis=(0, 1)))
    s = np.roll(np.abs(c), (n // 2, n // 2), axis=(0, 1))
    px, py = divmod(int(np.argmax(s)), n)
    if s[px, py] == 0:
        return 0.0, 0.0
    dx = dy = 0.0
    if 0 < px < n - 1 and 0 < py < n - 1:
        a = s[px - 1:px + 2, py - 1:py + 2]
        fx, fy = 0.5 * (a[2, 1] - a[0, 1]), 0.5 * (a[1, 2] - a[1, 0])
        fxx, fyy = a[2, 1] + a[0, 1] - 2 * a[1, 1], a[1, 2] + a[1, 0] - 2 * a[1, 1]
        fxy = 0.25 * (a[2, 2] + a[0, 0] - a[2, 0] - a[0, 2])
        den = fxy * fxy - fxx * fyy
        dx, dy = (fyy * fx - fy * fxy) / den, (fxx * fy - fx * fxy) / den
    return px + dx - n // 2, py + dy - n // 2


def _direct_bias_pixel(f1, f2, sigma, i, j, n):
    """-b at pixel (i, j) from its definition: the plain estimate d with
    the Gaussian centred on the pixel, then image 1 windowed about
    pixel - d/2 and image 2 about pixel + d/2 over the same clamped box with
    the same box means, and the estimate again.  Returns (vx, vy)."""
    nx, ny = f1.shape
    ns = n // 2
    i0 = nx - n if i + ns >= nx - 1 else max(0, i - ns + 1)
    j0 = ny - n if j + ns >= ny - 1 else max(0, j - ns + 1)
    b1, b2 = f1[i0:i0 + n, j0:j0 + n], f2[i0:i0 + n, j0:j0 + n]
    b1, b2 = b1 - b1.mean(), b2 - b2.mean()
    r = np.arange(i0, i0 + n, dtype=float)[:, None]
    q = np.arange(j0, j0 + n, dtype=float)[None, :]

    def win(ci, cj):
        return np.exp(-((r - ci) / sigma) ** 2 - ((q - cj) / sigma) ** 2)

    d = _direct_estimate(win(i, j) * b1, win(i, j) * b2)
    e = _direct_estimate(win(i - d[0] / 2, j - d[1] / 2) * b1, win(i + d[0] / 2, j + d[1] / 2) * b2)
    return e[1], e[0]  # transp = 1 (flct.c:901-905)


def test_bias_matches_direct_sum_definition():
    """-b on boxes 8 and 12 (sigma 2, 3), every pixel of the frame (clamped
    edge boxes included): the oracle against its definition evaluated by
    direct O(N^4) correlation sums, at 1e-10 px."""
    rng = np.random.default_rng(31)
    for sigma, (nx, ny) in ((2.0, (16, 18)), (3.0, (20, 22))):
        n = 2 * O.box_nsize(sigma, nx, ny)
        assert n == 4 * sigma
        a, b = smooth_pair(rng, nx, ny, shift=(0.8, -1.3), width=1.5)
        pix = np.arange(nx * ny)
        vx, vy, vm = O.flct_pixels(a, b, 1.0, 1.0, sigma, pix, bias=True)
        px, py, _ = O.flct_pixels(a, b, 1.0, 1.0, sigma, pix)
        moved = 0
        for p in pix:
            i, j = divmod(int(p), ny)
            ex, ey = _direct_bias_pixel(a, b, sigma, i, j, n)
            assert abs(vx[p] - ex) < 1e-10 and abs(vy[p] - ey) < 1e-10, (sigma, i, j)
            moved += abs(vx[p] - px[p]) + abs(vy[p] - py[p]) > 1e-3
        assert moved > 0.5 * pix.size  # the second pass changes most estimates


def test_bias_zero_offset_is_the_plain_pass():
    """The second pass with d = 0 is the plain pass bit for bit."""
    rng = np.random.default_rng(32)
    a, b = smooth_pair(rng, 40, 44, shift=(0.6, 0.9))
    pix = np.array([0, 43, 5 * 44 + 7, 20 * 44 + 21, 39 * 44 + 43, 33 * 44 + 2])
    vx, vy, _ = O.flct_pixels(a, b, 1.0, 1.0, 4.0, pix)
    for k, p in enumerate(pix):
        sx, sy = O.bias_pass(a, b, 4.0, *divmod(int(p), 44), (0.0, 0.0))
        assert sy == vx[k] and sx == vy[k]


def test_bias_exact_on_a_moved_integer_blob():
    """A zero-sum integer pattern moved by an integer s, wholly inside the
    box: with d = s the two windows follow the pattern, image 2's windowed
    box is image 1's moved by s, and the pass returns s (to FFT rounding);
    the plain pass does not."""
    nx = ny = 48
    pat = np.array([[1, -2, 1, 0], [-2, 5, -3, 1], [0, -2, 1, 0]], float)
    assert pat.sum() == 0
    i, j = 24, 23
    for s in ((2, -1), (-3, 2), (1, 1)):
        f1, f2 = np.zeros((nx, ny)), np.zeros((nx, ny))
        f1[i - 1:i + 2, j - 1:j + 3] = pat
        f2[i - 1 + s[0]:i + 2 + s[0], j - 1 + s[1]:j + 3 + s[1]] = pat
        sx, sy = O.bias_pass(f1, f2, 4.0, i, j, s)
        assert abs(sx - s[0]) < 1e-10 and abs(sy - s[1]) < 1e-10, s
        px, py = O.bias_pass(f1, f2, 4.0, i, j, (0.0, 0.0))
        assert abs(px - s[0]) + abs(py - s[1]) > 1e-3, s


def test_bias_swap_negates_velocity():
    """Swapping the two images under -b negates every velocity whose peaks
    are interior (the windows of the swapped pair are the original ones
    exchanged), on a non-uniform flow."""
    from pyflct_amd import synth
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(33)
    nx, ny = 64, 64
    a = gaussian_filter(rng.standard_normal((nx, ny)), 2.0, mode="wrap")
    u = 0.9 + 0.5 * np.sin(2 * np.pi * np.arange(nx) / nx)
    b = synth.shift_rows(a, u)
    v1 = O.flct(a, b, 1.0, 1.0, 5.0, bias=True)
    v2 = O.flct(b, a, 1.0, 1.0, 5.0, bias=True)
    assert np.max(np.abs(v1[0] + v2[0])) < 1e-9 and np.max(np.abs(v1[1] + v2[1])) < 1e-9
    assert np.median(np.abs(v1[0])) > 0.3


def test_deltat_deltas_scaling():
    rng = np.random.default_rng(6)
    a, b = smooth_pair(rng, 40, 40)
    v1 = O.flct(a, b, 1.0, 1.0, 4.0)
    v2 = O.flct(a, b, 2.0, 3.0, 4.0)
    assert np.allclose(v2[0], v1[0] * 1.5, rtol=0, atol=1e-12)
    assert np.allclose(v2[1], v1[1] * 1.5, rtol=0, atol=1e-12)


def test_threshold_relative_and_absolute_sets():
    rng = np.random.default_rng(7)
    a, b = smooth_pair(rng, 36, 36)
    fmax = max(np.max(np.abs(a)), np.max(np.abs(b)))
    below = 0.5 * np.abs(a + b) < 0.3 * fmax
    _, _, vm = O.flct(a, b, 1.0, 1.0, 4.0, thresh=0.3)
    assert np.array_equal(vm == 0, below)
    _, _, vm2 = O.flct(a, b, 1.0, 1.0, 4.0, thresh=0.3, thresh_abs=True)
    assert np.array_equal(vm2 == 0, 0.5 * np.abs(a + b) < 0.3)
    assert O.resolve_thresh(a, b, 0.3) == 0.3 * fmax
    assert O.resolve_thresh(a, b, 1.5) == 1.5  # >= 1 is absolute


def test_skip_lattice_and_fill_mask():
    rng = np.random.default_rng(8)
    a, b = smooth_pair(rng, 41, 37)
    for skip, p, q in [(3, 0, 0), (4, 1, 2), (5, -2, -1)]:
        _, _, vm = O.flct(a, b, 1.0, 1.0, 4.0, skip=skip, poff=p, qoff=q)
        pp, qq = (skip - abs(p) if p < 0 else p), (skip - abs(q) if q < 0 else q)
        i, j = np.mgrid[0:41, 0:37]
        lattice = ((i - qq) % skip == 0) & ((j - pp) % skip == 0)
        assert np.array_equal(vm == 1, lattice)
        _, _, vmi = O.flct(a, b, 1.0, 1.0, 4.0, skip=skip, poff=p, qoff=q, interp=True)
        assert np.all(vmi[lattice] == 1)
        assert set(np.unique(vmi[~lattice])) <= {0.0, 0.5}


def _fourier_shift(a, dx, dy):
    """Band-limited circular shift of ``a`` by (dx, dy) pixels along (axis 0,
    axis 1), with numpy's FFT."""
    kx = np.fft.fftfreq(a.shape[0])[:, None]
    ky = np.fft.fftfreq(a.shape[1])[None, :]
    return np.fft.ifft2(np.fft.fft2(a) * np.exp(-2j * np.pi * (kx * dx + ky * dy))).real


@pytest.mark.parametrize("nx,ny", [(48, 64), (51, 37), (101, 60)])
def test_sigma0_whole_image_integer_shift(nx, ny):
    """sigma = 0 correlates the two whole images once (flct.c:567-586,
    802-808): a circular integer shift gives a correlation surface symmetric
    about the peak, so the 3 x 3 fit adds nothing and the velocity is the
    shift times deltas / deltat, with x and y swapped (transp = 1)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(nx * 1000 + ny)
    g = gaussian_filter(rng.standard_normal((nx, ny)), 2.0, mode="wrap") + 5.0
    h = np.roll(g, (2, -3), axis=(0, 1)) + 1.25  # an offset is removed by the mean
    for kw in ({}, {"kr": 0.6}):
        vx, vy, vm = O.flct(g, h, 2.0, 0.5, 0.0, **kw)
        assert vx.shape == (1, 1) and vm[0, 0] == 1.0
        assert abs(vx[0, 0] - (-3) * 0.25) < 1e-9, kw
        assert abs(vy[0, 0] - 2 * 0.25) < 1e-9, kw


def test_sigma0_ignores_threshold_and_skip():
    """sigma = 0 computes the one velocity whatever the threshold and the
    skip lattice say (flct.c:594, 920)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(11)
    g = gaussian_filter(rng.standard_normal((40, 44)), 1.5, mode="wrap")
    h = np.roll(g, (-1, 2), axis=(0, 1))
    ref = O.flct(g, h, 1.0, 1.0, 0.0)
    for kw in ({"thresh": 0.9}, {"thresh": 100.0, "thresh_abs": True}, {"skip": 5, "poff": 2}):
        got = O.flct(g, h, 1.0, 1.0, 0.0, **kw)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), kw


def test_sigma0_subpixel_shift_and_negative_sigma():
    """A band-limited sub-pixel shift is recovered to the accuracy of the
    quadratic peak fit; a negative sigma takes the same whole-image branch as
    sigma = 0 (flct.c:357-365 sets sigmaeq0 for any sigma not > 0, and the
    box rules of flct.c:567-580 run only without it)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(12)
    g = gaussian_filter(rng.standard_normal((96, 80)), 3.0, mode="wrap")
    h = _fourier_shift(g, 0.3, -0.45)
    vx, vy, _ = O.flct(g, h, 1.0, 1.0, 0.0)
    assert abs(vx[0, 0] + 0.45) < 0.05 and abs(vy[0, 0] - 0.3) < 0.05
    h2 = _fourier_shift(g, -0.3, 0.45)
    vx2, vy2, _ = O.flct(g, h2, 1.0, 1.0, 0.0)
    assert abs(vx2[0, 0] + vx[0, 0]) < 1e-6 and abs(vy2[0, 0] + vy[0, 0]) < 1e-6
    for neg in (-4.0, -0.5):
        got = O.flct(g, h, 1.0, 1.0, neg)
        for a, b in zip((vx, vy, _), got):
            assert a.shape == (1, 1) and np.array_equal(a, b)


def test_warp_fixture_was_made_in_double_precision(golden):
    """The reference's warp fixture (hashgauss.dat f2 = shift_frac2d(f1,
    deltax, deltay), README.md:250-262) matches warp_frac2d in FP64 to a few
    float32 steps of the stored file, while a strict single-precision
    evaluation of shift_frac2d.pro's non-uniform branch (float phases with
    the float pi, float products and sums, IDL-io-procedures/
    shift_frac2d.pro:200-251) misses it by ~1e-5 on the sampled pixels: the
    fixture pins the FP64 form, and an FP32 warp form would move away from
    it.  Sampled pixels (numpy float32 restatement)."""
    f1, f2, dx, dy = golden["f1"], golden["f2"], golden["dx"], golden["dy"]
    nx, ny = f1.shape
    rng = np.random.default_rng(0)
    pix = rng.choice(nx * ny, 200, replace=False)
    # strict float32 restatement, IDL layout: x = C column j, y = C row i
    ff = (np.fft.fft2(f1.T) / (nx * ny)).astype(np.complex64)  # [kx_idl, ky_idl]
    kx = np.asarray(O.make_freq(ny), np.float32)
    ky = np.asarray(O.make_freq(nx), np.float32)
    pi2 = np.float32(np.float32(np.pi) * np.float32(2))
    single = np.empty(pix.size)
    for q, p in enumerate(pix):
        i, j = divmod(int(p), ny)  # C row i = IDL y, C column j = IDL x
        xarg = pi2 * (np.float32(j) - np.float32(dx[i, j])) / np.float32(ny)
        yarg = pi2 * (np.float32(i) - np.float32(dy[i, j])) / np.float32(nx)
        ax, ay = kx * xarg, ky * yarg
        ex = (np.cos(ax) + 1j * np.sin(ax)).astype(np.complex64)
        ey = (np.cos(ay) + 1j * np.sin(ay)).astype(np.complex64)
        single[q] = np.sum((np.outer(ex, ey) * ff).astype(np.complex64), dtype=np.complex64).real
    wo = O.warp_frac2d(f1, dx, dy).ravel()[pix]
    ref = f2.ravel()[pix]
    assert np.max(np.abs(wo - ref)) < 5e-7
    assert np.max(np.abs(single - ref)) > 5e-6
