"""Synthetic image pairs for the benchmark and test configurations.

Every generator is numpy PCG64 with a recorded seed (default 20261203 plus
the configuration index).  Conventions, fixed here because the workload
descriptions leave them open:

* a blob is A * exp(-r^2 / (2 s^2));
* "N(0, 0.01)" noise has standard deviation 0.01;
* a bipolar blob is a +A / -A pair of blobs of the same width s whose
  centres are 2 s apart along a random direction;
* magnetogram-like values are in gauss (G).

Arrays are C-space (nx rows = IDL y, ny columns = IDL x).  Displacement
fields are returned as (delx, dely) in IDL axes, i.e. delx moves content
along the column index j and dely along the row index i, the convention of
warp's shift files (warp.c:948-952).
"""
from __future__ import annotations

import numpy as np

SEED0 = 20261203


def rng_for(cfg_index, seed=None):
    return np.random.Generator(np.random.PCG64(SEED0 + cfg_index if seed is None else seed))


def make_freq(n):
    """Integer wavenumbers in FFT order, Nyquist = +n/2 (flct.c:1651-1692)."""
    k = np.fft.fftfreq(n, d=1.0 / n)
    if n % 2 == 0:
        k[n // 2] = n // 2
    return k


def add_blobs(img, rng, count, amp_fn, width_lo, width_hi, bipolar=False, mask=None):
    """Adds `count` blobs; with `mask`, a centre the mask rejects is drawn
    again, so all `count` blobs land where the mask accepts."""
    nx, ny = img.shape
    for _ in range(count):
        s = rng.uniform(width_lo, width_hi)
        ci, cj = rng.uniform(0, nx), rng.uniform(0, ny)
        while mask is not None and not mask(ci, cj):
            ci, cj = rng.uniform(0, nx), rng.uniform(0, ny)
        a = amp_fn(rng)
        parts = [(ci, cj, a)]
        if bipolar:
            th = rng.uniform(0, 2 * np.pi)
            di, dj = s * np.cos(th), s * np.sin(th)
            parts = [(ci - di, cj - dj, a), (ci + di, cj + dj, -a)]
        for (pi, pj, pa) in parts:
            r = int(np.ceil(5 * s))
            i0, i1 = max(0, int(pi) - r), min(nx, int(pi) + r + 1)
            j0, j1 = max(0, int(pj) - r), min(ny, int(pj) + r + 1)
            if i0 >= i1 or j0 >= j1:
                continue
            ii = np.arange(i0, i1)[:, None] - pi
            jj = np.arange(j0, j1)[None, :] - pj
            img[i0:i1, j0:j1] += pa * np.exp(-(ii * ii + jj * jj) / (2 * s * s))
    return img


def shift_uniform(arr, delx, dely):
    """Uniform Fourier shift with shift_frac2d's conventions (warp.c:757-881,
    transp = 1): content moves by delx along j and dely along i."""
    nx, ny = arr.shape
    kx, ky = make_freq(nx), make_freq(ny)
    dxarg, dyarg = -dely / nx, -delx / ny
    ph = np.exp(2j * np.pi * (kx[:, None] * dxarg + ky[None, :] * dyarg))
    return np.real(np.fft.ifft2(np.fft.fft2(arr) * ph))


def shift_rows(arr, vx_of_row):
    """Shift every row i by vx_of_row[i] pixels along j (exact 1-D Fourier
    shift of a periodic row; equals warp for a j-only displacement that is
    constant along each row)."""
    nx, ny = arr.shape
    k = make_freq(ny)
    F = np.fft.fft(arr, axis=1)
    ph = np.exp(-2j * np.pi * k[None, :] * np.asarray(vx_of_row)[:, None] / ny)
    return np.real(np.fft.ifft(F * ph, axis=1))


def magnetogram(n, rng, scale=50.0, nblobs=300, amp=(200.0, 2000.0), width=(2.0, 8.0)):
    """Laplace(0, scale) background plus bipolar blobs of +-U[amp] G."""
    img = rng.laplace(0.0, scale, size=(n, n))
    return add_blobs(img, rng, nblobs, lambda r: r.uniform(*amp), width[0], width[1],
                     bipolar=True)


def rotation_field(n, degrees):
    """Displacement of a solid-body rotation by `degrees` about the frame
    centre ((n-1)/2, (n-1)/2): returns (delx, dely) in IDL axes."""
    c = (n - 1) / 2.0
    th = np.deg2rad(degrees)
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)  # y = row i, x = column j
    xr, yr = x - c, y - c
    delx = (np.cos(th) * xr - np.sin(th) * yr) - xr
    dely = (np.sin(th) * xr + np.cos(th) * yr) - yr
    return delx, dely


def fourier_mode_field(n, rng, nmodes=8, lam=(64.0, 512.0), vmax=1.5):
    """Smooth random velocity: a sum of random Fourier modes with
    wavelengths in lam, scaled to max |v| = vmax."""
    t = np.arange(n, dtype=np.float64)  # x = column j, y = row i
    vx = np.zeros((n, n))
    vy = np.zeros((n, n))
    for _ in range(nmodes):
        L = rng.uniform(*lam)
        th = rng.uniform(0, 2 * np.pi)
        kx, ky = np.cos(th) / L, np.sin(th) / L
        a = 2 * np.pi * kx * t                                 # phase along x
        b = 2 * np.pi * ky * t + rng.uniform(0, 2 * np.pi)    # phase along y
        # sin(a + b) as two outer products (row index y, column index x)
        mode = np.outer(np.cos(b), np.sin(a)) + np.outer(np.sin(b), np.cos(a))
        ax, ay = rng.standard_normal(2)
        vx += ax * mode
        vy += ay * mode
    m = np.max(np.hypot(vx, vy))
    return vx * (vmax / m), vy * (vmax / m)


def cfg1(seed=None):
    """128^2: 40 Gaussian blobs + N(0, 0.01); image 2 = image 1 shifted by
    (0.3, -0.2) px.  Run: deltat 1, deltas 1, sigma 8."""
    rng = rng_for(0, seed)
    n = 128
    img = add_blobs(np.zeros((n, n)), rng, 40, lambda r: r.standard_normal(), 2.0, 6.0)
    img += rng.normal(0.0, 0.01, size=(n, n))
    return img, shift_uniform(img, 0.3, -0.2)


def cfg2_inputs(seed=None):
    """512^2 magnetogram and the 0.01 degree rotation field; image 2 is
    warp(image 1, field) (computed by the caller's warp)."""
    rng = rng_for(1, seed)
    n = 512
    img = magnetogram(n, rng)
    delx, dely = rotation_field(n, 0.01)
    return img, delx, dely


def cfg3_inputs(seed=None):
    """1024^2 magnetogram and a smooth random velocity field (8 modes,
    wavelengths 64-512 px, max |v| 1.5 px)."""
    rng = rng_for(2, seed)
    n = 1024
    img = magnetogram(n, rng)
    delx, dely = fourier_mode_field(n, rng)
    return img, delx, dely


def uneven_inputs(n, seed):
    """One pair of the uneven batch: an n^2 magnetogram (generator of the
    512^2 case) and a smooth random velocity field (8 modes, wavelengths
    64-512 px, max |v| 1.5 px); image 2 is warp(image 1, field), computed by
    the caller's warp."""
    rng = np.random.default_rng(seed)
    img = magnetogram(n, rng)
    delx, dely = fourier_mode_field(n, rng)
    return img, delx, dely


def cfg4(n=4096, seed=None, radius=None):
    """Full-disk-like pair: disk of radius 1900 px (scaled with n), Laplace
    scale 30 G plus 2000 bipolar active-region blobs up to +-3000 G inside,
    zero outside.  Image 2: differential rotation v = A + B sin^2(lat) px
    with A = 0.5, B = -0.1, along j, applied as an exact per-row Fourier
    shift of the unmasked field, then the same disk mask.
    Returns (image1, image2, vx_true)."""
    rng = rng_for(3, seed)
    r = radius if radius is not None else 1900.0 * n / 4096.0
    c = (n - 1) / 2.0
    base = rng.laplace(0.0, 30.0, size=(n, n))

    def inside(ci, cj):
        return (ci - c) ** 2 + (cj - c) ** 2 < (0.95 * r) ** 2

    nbl = max(1, int(round(2000 * (n / 4096.0) ** 2)))
    add_blobs(base, rng, nbl, lambda g: g.uniform(100.0, 3000.0), 3.0, 12.0, bipolar=True,
              mask=inside)
    i = np.arange(n, dtype=np.float64)
    sinlat = np.clip((i - c) / r, -1.0, 1.0)
    v_row = 0.5 - 0.1 * sinlat ** 2
    moved = shift_rows(base, v_row)
    y, x = np.mgrid[0:n, 0:n]
    disk = ((y - c) ** 2 + (x - c) ** 2) <= r * r
    f1 = np.where(disk, base, 0.0)
    f2 = np.where(disk, moved, 0.0)
    vx_true = np.where(disk, v_row[:, None], 0.0)
    return f1, f2, vx_true
