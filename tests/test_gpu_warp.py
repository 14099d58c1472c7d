"""GPU parity of flct_hip_warp (warp_frac2d) and flct_hip_shift
(shift_frac2d) against the oracle and the reference fixture."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def test_warp_fixture(ctx, golden):
    w = ctx.warp(golden["f1"], golden["dx"], golden["dy"])
    wo = O.warp_frac2d(golden["f1"], golden["dx"], golden["dy"])
    assert np.max(np.abs(w - wo)) < 1e-12 * np.max(np.abs(wo)) * 100
    assert np.max(np.abs(w - golden["f2"])) < 5e-7


@pytest.mark.parametrize("shape", [(37, 50), (64, 64), (33, 17), (48, 31), (96, 80)])
def test_warp_random_fields(ctx, shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape) * 10.0
    dx = rng.standard_normal(shape) * 1.5
    dy = rng.standard_normal(shape) * 1.5
    for transp in (1, 0):
        w = ctx.warp(a, dx, dy, transp=bool(transp))
        wo = O.warp_frac2d(a, dx, dy, transp=transp)
        assert np.max(np.abs(w - wo)) < 1e-10 * np.max(np.abs(a))


def test_warp_integer_shift_is_roll(ctx):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((60, 72))
    dx = np.full(a.shape, 3.0)   # IDL x = columns j
    dy = np.full(a.shape, -2.0)  # IDL y = rows i
    w = ctx.warp(a, dx, dy)
    assert np.max(np.abs(w - np.roll(np.roll(a, -2, axis=0), 3, axis=1))) < 1e-11


def test_uniform_warp_equals_shift(ctx):
    rng = np.random.default_rng(2)
    a = rng.standard_normal((45, 64))
    w = ctx.warp(a, np.full(a.shape, 0.37), np.full(a.shape, -1.21))
    s = ctx.shift(a, 0.37, -1.21)
    assert np.max(np.abs(w - s)) < 1e-11


@pytest.mark.parametrize("shape", [(64, 48), (101, 201), (31, 31)])
def test_shift_matches_oracle(ctx, shape):
    rng = np.random.default_rng(3)
    a = rng.standard_normal(shape)
    s = ctx.shift(a, 0.3, -0.2)
    so = O.shift_frac2d(a, 0.3, -0.2)
    assert np.max(np.abs(s - so)) < 1e-12


def test_warp_large_properties(ctx):
    """512^2: the zero field is the identity, an integer field is a roll."""
    rng = np.random.default_rng(4)
    a = rng.laplace(0, 50, (512, 512))
    z = np.zeros_like(a)
    w = ctx.warp(a, z, z)
    assert np.max(np.abs(w - a)) < 1e-9 * np.max(np.abs(a))
    w = ctx.warp(a, z + 5.0, z + 1.0)
    assert np.max(np.abs(w - np.roll(np.roll(a, 1, axis=0), 5, axis=1))) < 1e-9 * np.max(np.abs(a))


# ------------------------------------------------ direct sum and NUFFT
@pytest.mark.parametrize("shape", [(37, 50), (64, 64), (48, 31), (96, 80), (129, 100)])
def test_warp_nufft_matches_direct_sum(ctx, shape):
    """The NUFFT evaluation of warp_frac2d's Fourier series against the
    reference's direct per-pixel sum (GPU) and the oracle, both transposes."""
    rng = np.random.default_rng(100 + sum(shape))
    a = rng.standard_normal(shape) * 10.0
    dx = rng.standard_normal(shape) * 2.0
    dy = rng.standard_normal(shape) * 2.0
    try:
        for transp in (1, 0):
            ctx.set_warp_method("direct")
            wd = ctx.warp(a, dx, dy, transp=bool(transp))
            assert ctx.last_warp_method() == "direct"
            ctx.set_warp_method("nufft")
            wn = ctx.warp(a, dx, dy, transp=bool(transp))
            assert ctx.last_warp_method() == "nufft"
            scale = np.max(np.abs(a))
            assert np.max(np.abs(wn - wd)) < 1e-11 * scale, (shape, transp)
            wo = O.warp_frac2d(a, dx, dy, transp=transp)
            assert np.max(np.abs(wn - wo)) < 1e-11 * scale
    finally:
        ctx.set_warp_method("auto")


def test_warp_method_selection(ctx):
    from pyflct_amd._lib import FlctHipError
    rng = np.random.default_rng(9)
    small = rng.standard_normal((12, 40))
    z = np.zeros_like(small)
    ctx.warp(small, z, z)
    assert ctx.last_warp_method() == "direct"  # below 16 rows: the direct sum
    big = rng.standard_normal((40, 40))
    ctx.warp(big, np.zeros_like(big), np.zeros_like(big))
    assert ctx.last_warp_method() == "nufft"
    try:
        ctx.set_warp_method("nufft")
        with pytest.raises(FlctHipError):
            ctx.warp(small, z, z)
    finally:
        ctx.set_warp_method("auto")


def test_warp_nufft_large_fields(ctx):
    """2048 x 1536 (NUFFT): an integer field is a roll and a uniform field is
    shift_frac2d; at 1024 x 768 a smooth field agrees with the direct sum."""
    rng = np.random.default_rng(5)
    nx, ny = 2048, 1536
    a = rng.laplace(0, 50, (nx, ny))
    z = np.zeros_like(a)
    scale = np.max(np.abs(a))
    w = ctx.warp(a, z - 7.0, z + 3.0)
    assert ctx.last_warp_method() == "nufft"
    assert np.max(np.abs(w - np.roll(np.roll(a, 3, axis=0), -7, axis=1))) < 1e-10 * scale
    w = ctx.warp(a, z + 0.37, z - 1.21)
    assert np.max(np.abs(w - ctx.shift(a, 0.37, -1.21))) < 1e-10 * scale
    nx, ny = 1024, 768
    a = a[:nx, :ny].copy()
    scale = np.max(np.abs(a))
    i, j = np.mgrid[0:nx, 0:ny]
    dx = 1.2 * np.sin(2 * np.pi * i / 700.0) * np.cos(2 * np.pi * j / 500.0)
    dy = -0.8 * np.cos(2 * np.pi * i / 300.0)
    wn = ctx.warp(a, dx, dy)
    try:
        ctx.set_warp_method("direct")
        wd = ctx.warp(a, dx, dy)
    finally:
        ctx.set_warp_method("auto")
    assert np.max(np.abs(wn - wd)) < 1e-10 * scale


def test_warp_nufft_far_shifts(ctx):
    """Displacements of thousands of pixels: the NUFFT takes the reference's
    phase arguments (its pi literal makes them not exactly periodic in the
    shift) and reduces the grid position to one period; it agrees with the
    direct sum, and with the near shift to the size of the pi literal's
    phase drift."""
    rng = np.random.default_rng(11)
    a = rng.standard_normal((48, 40))
    dx = rng.standard_normal(a.shape) * 0.7 + 40 * 1000.0
    dy = rng.standard_normal(a.shape) * 0.7 - 48 * 300.0
    far = ctx.warp(a, dx, dy)
    assert ctx.last_warp_method() == "nufft"
    try:
        ctx.set_warp_method("direct")
        far_d = ctx.warp(a, dx, dy)
    finally:
        ctx.set_warp_method("auto")
    scale = np.max(np.abs(a))
    assert np.max(np.abs(far - far_d)) < 1e-8 * scale
    near = ctx.warp(a, dx - 40 * 1000.0, dy + 48 * 300.0)
    assert np.max(np.abs(far - near)) < 1e-6 * scale


@pytest.mark.parametrize("shape", [(16, 16), (17, 4093), (4093, 23), (3001, 40), (16, 4096)])
def test_warp_nufft_extreme_shapes(ctx, shape):
    """The NUFFT at its limits (16 rows or columns, 4096 and odd lengths up to
    it, where the spectrum transforms are Bluestein) against the direct sum."""
    rng = np.random.default_rng(700 + shape[0])
    a = rng.standard_normal(shape) * 5.0
    dx = rng.standard_normal(shape) * 1.5
    dy = rng.standard_normal(shape) * 1.5
    wn = ctx.warp(a, dx, dy)
    assert ctx.last_warp_method() == "nufft"
    try:
        ctx.set_warp_method("direct")
        wd = ctx.warp(a, dx, dy)
    finally:
        ctx.set_warp_method("auto")
    assert np.max(np.abs(wn - wd)) < 1e-11 * np.max(np.abs(a)), shape
