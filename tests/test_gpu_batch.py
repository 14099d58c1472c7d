"""flct_hip_track_batch: many pairs of uneven size, sigma, options and form in
one captured HIP graph (GPU).

Each pair runs the same kernels as a single flct_hip_track call, so the
results must be bit-identical to the single-pair path for every form.  The
graph is relaunched while the pairs are the same; replays are checked
against the eager results with eager work and synchronisations in between,
also after new data is written into the same device arrays.
"""
import numpy as np
import pytest

from conftest import smooth_pair
from pyflct_amd import synth
from pyflct_amd._lib import FlctHipError

pytestmark = pytest.mark.gpu

# (nx, ny, sigma, options)
PAIRS = [
    (200, 230, 6.0, {}),                                         # FP64, N = 12
    (160, 150, 5.0, {"kr": 0.5, "precision": "fp32"}),           # FP32 with the filter
    (256, 256, 12.0, {"precision": "split_bf16"}),               # row-segment kernel, N = 48
    (180, 190, 11.0, {"thresh": 0.2, "precision": "split_bf16"}),  # relative threshold list
    (130, 140, 4.0, {"skip": 3, "poff": 1, "interp": True}),     # skip lattice and fill
    (256, 256, 20.0, {"precision": "split_bf16"}),               # N = 84: the FP32 kernels
    (220, 200, 9.0, {"thresh": 0.4, "thresh_abs": True, "bias": True}),  # absolute threshold, bias
]


def _inputs(seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for nx, ny, sigma, kw in PAIRS:
        a, b = smooth_pair(rng, nx, ny, shift=(0.6, -0.9), width=2.5)
        out.append((a, b, sigma, kw))
    return out


def test_batch_matches_single_pairs(ctx):
    ins = _inputs()
    got = ctx.track_batch([(a, b, 1.0, 1.0, s, kw) for a, b, s, kw in ins])
    for (a, b, s, kw), g in zip(ins, got):
        ref = ctx.track(a, b, 1.0, 1.0, s, **kw)
        for k in range(3):
            assert np.array_equal(g[k], ref[k]), (a.shape, s, kw, k)


def test_batch_graph_replays(ctx):
    ins = _inputs(seed=1)
    dev, pairs = [], []
    for a, b, s, kw in ins:
        d = [ctx.malloc(a.nbytes) for _ in range(5)]
        dev.append(d)
        pairs.append((d[0], d[1], d[2], d[3], d[4], a.shape[0], a.shape[1],
                      ctx.make_opts(1.0, 1.0, s, **kw)))
    try:
        for rnd, data in enumerate((ins, ins, _inputs(seed=2), _inputs(seed=2))):
            for d, (a, b, _, _) in zip(dev, data):
                ctx.upload(d[0], a)
                ctx.upload(d[1], b)
            stats = ctx.track_batch_device(pairs)
            for d, (a, b, s, kw), st in zip(dev, data, stats):
                g = [ctx.download(np.empty(a.shape), d[k]) for k in (2, 3, 4)]
                # eager single-pair work between the replays
                ref = ctx.track(a, b, 1.0, 1.0, s, **kw)
                assert st["tracked"] == ctx.last_stats()["tracked"]
                for k in range(3):
                    assert np.array_equal(g[k], ref[k]), (rnd, a.shape, s, kw, k)
    finally:
        for d in dev:
            for p in d:
                ctx.free(p)


def test_batch_rejects_invalid_pair_and_runs_nothing(ctx):
    a, b = smooth_pair(np.random.default_rng(3), 96, 100)
    d = [ctx.malloc(a.nbytes) for _ in range(5)]
    try:
        ctx.upload(d[0], a)
        ctx.upload(d[1], b)
        sentinel = np.full(a.shape, 7.0)
        for k in (2, 3, 4):
            ctx.upload(d[k], sentinel)
        good = (d[0], d[1], d[2], d[3], d[4], 96, 100, ctx.make_opts(1.0, 1.0, 5.0))
        bad = (d[0], d[1], d[2], d[3], d[4], 96, 100, ctx.make_opts(1.0, 1.0, 5.0, skip=100))
        with pytest.raises(FlctHipError) as e:
            ctx.track_batch_device([good, bad])
        assert e.value.status == 4  # FLCT_HIP_E_SKIP_TOO_BIG (flct.c:559-563)
        for k in (2, 3, 4):
            assert np.array_equal(ctx.download(np.empty(a.shape), d[k]), sentinel)
    finally:
        for p in d:
            ctx.free(p)


def test_uneven_batch_forms_and_warp_residual(ctx):
    """A reduced uneven batch (sizes 256 / 512, sigma 4..20, smooth random
    velocity fields): FP32 and split-BF16 against FP64, then image 1 warped
    by the recovered velocities against image 2."""
    rng = np.random.default_rng(20261207)
    ins = []
    for k in range(6):
        n = int(rng.choice([256, 512]))
        sigma = float(rng.uniform(4, 20))
        f1, dx, dy = synth.uneven_inputs(n, seed=int(rng.integers(1 << 30)))
        ins.append((f1, ctx.warp(f1, dx, dy), sigma))
    ref = ctx.track_batch([(a, b, 1.0, 1.0, s) for a, b, s in ins])
    for form, tol_rms in (("fp32", 2e-3), ("split_bf16", 1e-4)):
        got = ctx.track_batch([(a, b, 1.0, 1.0, s) for a, b, s in ins], precision=form)
        for r, g in zip(ref, got):
            assert np.array_equal(r[2], g[2])
            e = np.hypot(g[0] - r[0], g[1] - r[1])
            assert np.sqrt(np.mean(e ** 2)) <= tol_rms, (form, np.sqrt(np.mean(e ** 2)))
    for (a, b, s), (vx, vy, vm) in zip(ins, ref):
        w = ctx.warp(a, vx, vy)
        n = a.shape[0]
        m = slice(n // 8, n - n // 8)
        res = np.sqrt(np.mean((w[m, m] - b[m, m]) ** 2)) / np.sqrt(np.mean(b[m, m] ** 2))
        base = np.sqrt(np.mean((a[m, m] - b[m, m]) ** 2)) / np.sqrt(np.mean(b[m, m] ** 2))
        assert res < 0.5 * base, (a.shape, s, res, base)


@pytest.mark.parametrize("form", ["fp64", "fp32", "split_bf16"])
def test_batch_of_every_box_size(ctx, form):
    """One graph launch holding a pair at every box size 4..96 (but 56), with
    a relative threshold and a skip lattice on every other pair: each result
    equals its single-pair call bit for bit."""
    import oracle_lib as O
    rng = np.random.default_rng(500)
    items = []
    for k, sigma in enumerate(range(1, 24)):
        n = 2 * O.box_nsize(sigma, 4096, 4096)
        a, b = smooth_pair(rng, int(1.6 * n) + 3 + k % 3, int(1.7 * n) + 1, shift=(-0.5, 0.7))
        kw = {"precision": form}
        if k % 2:
            kw.update(thresh=0.2, skip=3, poff=1, interp=True)
        items.append((a, b, 1.0, 1.0, float(sigma), kw))
    got = ctx.track_batch(items)
    for (a, b, dt, ds, s, kw), g in zip(items, got):
        ref = ctx.track(a, b, dt, ds, s, **kw)
        for k in range(3):
            assert np.array_equal(g[k], ref[k]), (a.shape, s, kw, k)


def test_batch_mixed_hires_bias_filter_forms(ctx):
    """-h, -b and -k pairs of different forms in one graph launch equal their
    single-pair calls bit for bit (-h runs the FP64 kernels in every form)."""
    rng = np.random.default_rng(501)
    specs = [(96, 110, 5.0, {"hires": True}),
             (150, 140, 15.0, {"hires": True, "precision": "split_bf16"}),
             (120, 128, 6.0, {"bias": True, "precision": "fp32"}),
             (200, 210, 12.0, {"bias": True, "kr": 0.5, "precision": "split_bf16"}),
             (64, 70, 2.0, {"hires": True, "kr": 0.7}),
             (256, 240, 20.0, {"kr": 0.4, "precision": "split_bf16"})]
    items = []
    for nx, ny, sigma, kw in specs:
        a, b = smooth_pair(rng, nx, ny, shift=(0.35, 0.8), width=2.0)
        items.append((a, b, 1.0, 1.0, sigma, kw))
    got = ctx.track_batch(items)
    for (a, b, dt, ds, s, kw), g in zip(items, got):
        ref = ctx.track(a, b, dt, ds, s, **kw)
        for k in range(3):
            assert np.array_equal(g[k], ref[k]), (a.shape, s, kw, k)


def test_batch_of_300_pairs(ctx):
    """A graph of 300 small pairs (sizes 40..71, sigma 2..6, every form):
    results equal single-pair calls on every tenth pair, the graph replays."""
    rng = np.random.default_rng(502)
    items = []
    forms = ("fp64", "fp32", "split_bf16")
    for k in range(300):
        nx, ny = int(rng.integers(40, 72)), int(rng.integers(40, 72))
        a, b = smooth_pair(rng, nx, ny, shift=(0.2, -0.4), width=1.8)
        items.append((a, b, 1.0, 1.0, float(rng.integers(2, 7)), {"precision": forms[k % 3]}))
    got = ctx.track_batch(items)
    again = ctx.track_batch(items)
    for k in range(0, 300, 10):
        a, b, dt, ds, s, kw = items[k]
        ref = ctx.track(a, b, dt, ds, s, **kw)
        for c in range(3):
            assert np.array_equal(got[k][c], ref[c]), (k, a.shape, s, kw, c)
            assert np.array_equal(again[k][c], ref[c]), (k, a.shape, s, kw, c)
