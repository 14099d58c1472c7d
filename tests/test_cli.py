"""The flct / warp command-line programs: argument grammar and exit paths
(CPU), and end-to-end runs on vcimage files (GPU)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, f32

BIN = os.path.join(ROOT, "pyflct_amd", "bin")
FLCT = os.path.join(BIN, "flct")
WARP = os.path.join(BIN, "warp")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(*args):
    p = subprocess.run(list(args), capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout


@pytest.fixture(scope="module", autouse=True)
def built():
    if not (os.path.exists(FLCT) and os.path.exists(WARP)):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "pyflct_amd", "cli")], check=True)


def test_usage_and_fatal_option_paths(tmp_path):
    rc, out = run(FLCT)
    assert rc == 1 and "usage" in out
    rc, out = run(FLCT, "a", "b", "1", "1", "5", "-s", "0")
    assert rc == 1 and "zero or negative" in out  # flct.c:459-464
    rc, out = run(FLCT, "a", "b", "1", "1", "5", "-s", "4p4")
    assert rc == 1 and "abs(p) or abs(q)" in out  # flct.c:466-472
    rc, out = run(FLCT, "a", "b", "1", "1", "5", "-s", "4q-5i")
    assert rc == 1 and "abs(p) or abs(q)" in out
    rc, out = run(FLCT, "/nonexistent.dat", "b", "1", "1", "5")
    assert rc == 1 and "cannot open" in out  # flct.c:1175-1179
    rc, out = run(FLCT, os.path.join(GOLD, "hashgauss.dat"), "/tmp/x.dat", "1", "1", "5", "-s", "200")
    assert rc == 1 and "too big" in out  # flct.c:559-563
    from pyflct_amd import vcimage
    tiny = tmp_path / "tiny.dat"
    vcimage.write(tiny, np.ones((5, 5)), np.ones((5, 5)))
    rc, out = run(FLCT, str(tiny), "/tmp/x.dat", "1", "1", "5")
    assert rc == 1 and "illegal box size" in out  # flct.c:575-579
    rc, out = run(WARP, "a")
    assert rc == 1 and "usage" in out


def test_warp_rejects_bad_shift_arrays(tmp_path):
    from pyflct_amd import vcimage
    img = tmp_path / "img.dat"
    sh = tmp_path / "sh.dat"
    vcimage.write(img, np.zeros((10, 12)))
    vcimage.write(sh, np.zeros((3, 4)), np.zeros((3, 4)))
    rc, out = run(WARP, str(img), str(sh), str(tmp_path / "o.dat"))
    # the device is opened after the shapes are known only on a machine with a GPU;
    # either way the run must fail
    assert rc == 1


@pytest.mark.gpu
def test_flct_cli_reproduces_fixture(tmp_path, golden):
    out = tmp_path / "testvels.dat"
    rc, text = run(FLCT, os.path.join(GOLD, "hashgauss.dat"), str(out), "1", "1", "5", "-k", "0.5", "-q")
    assert rc == 0, text
    from pyflct_amd import vcimage
    vx, vy, vm = vcimage.read3(out)
    assert np.array_equal(vm, golden["vm"])
    assert np.max(np.abs(vx - golden["vx"])) < 1e-11
    assert np.max(np.abs(vy - golden["vy"])) < 1e-11


@pytest.mark.gpu
def test_flct_cli_options_match_api(tmp_path, ctx):
    from pyflct_amd import vcimage, synth
    f1, f2 = synth.cfg1()
    inp = tmp_path / "in.dat"
    vcimage.write(inp, f1, f2)
    g1, g2 = vcimage.read2(inp)  # float32-quantised inputs, as flct reads them
    out = tmp_path / "out.dat"
    rc, text = run(FLCT, str(inp), str(out), "2", "0.5", "8", "-t", "0.05", "-s", "3p-1q1i", "-k", "0.6")
    assert rc == 0, text
    r = vcimage.read3(out)
    g = ctx.track(g1, g2, 2.0, 0.5, 8.0, thresh=0.05, skip=3, poff=-1, qoff=1, interp=True, kr=0.6)
    for a, b in zip(r, g):
        assert np.array_equal(a, f32(b))


@pytest.mark.gpu
def test_flct_cli_hires_matches_api(tmp_path, ctx, golden):
    """flct hashgauss.dat out.dat 1 1 5 -k 0.5 -h: the -h peak step through
    the program equals the API's."""
    from pyflct_amd import vcimage
    out = tmp_path / "h.dat"
    rc, text = run(FLCT, os.path.join(GOLD, "hashgauss.dat"), str(out), "1", "1", "5", "-k", "0.5", "-h")
    assert rc == 0, text
    assert "hires option turned on" in text
    r = vcimage.read3(out)
    g = ctx.track(golden["f1"], golden["f2"], 1.0, 1.0, 5.0, kr=0.5, hires=True)
    for a, b in zip(r, g):
        assert np.array_equal(a, f32(b))


@pytest.mark.gpu
def test_flct_cli_sigma0_writes_one_velocity(tmp_path, ctx):
    """flct in.dat out.dat 1 1 0 -k 0.5 (flct.1:212-219): the net shift of
    the whole image, written as 1 x 1 arrays."""
    from pyflct_amd import vcimage
    rng = np.random.default_rng(4)
    from scipy.ndimage import gaussian_filter
    f1 = gaussian_filter(rng.standard_normal((70, 90)), 2.0, mode="wrap")
    f2 = np.roll(f1, (1, -2), axis=(0, 1))
    inp = tmp_path / "in.dat"
    vcimage.write(inp, f1, f2)
    g1, g2 = vcimage.read2(inp)
    out = tmp_path / "out.dat"
    rc, text = run(FLCT, str(inp), str(out), "1", "1", "0", "-k", "0.5")
    assert rc == 0, text
    assert "vx = " in text and "box size" not in text
    r = vcimage.read3(out)
    g = ctx.track(g1, g2, 1.0, 1.0, 0.0, kr=0.5)
    for a, b in zip(r, g):
        assert a.shape == (1, 1) and np.array_equal(a, f32(b))
    assert abs(r[0][0, 0] + 2) < 1e-5 and abs(r[1][0, 0] - 1) < 1e-5
    # a negative sigma takes the same branch (flct.c:357-365)
    out2 = tmp_path / "out2.dat"
    rc, text = run(FLCT, str(inp), str(out2), "1", "1", "-5", "-k", "0.5")
    assert rc == 0, text
    for a, b in zip(vcimage.read3(out2), r):
        assert a.shape == (1, 1) and np.array_equal(a, b)


@pytest.mark.gpu
def test_warp_cli_warp_and_uniform_shift(tmp_path, golden):
    from pyflct_amd import vcimage
    out = tmp_path / "w.dat"
    rc, text = run(WARP, os.path.join(GOLD, "hashgauss.dat"), os.path.join(GOLD, "deltaxygauss.dat"),
                   str(out), "-q")
    assert rc == 0, text
    w = vcimage.read1(out)
    assert np.max(np.abs(w - golden["f2"])) < 1e-6
    sh = tmp_path / "s.dat"
    vcimage.write(sh, np.full((1, 1), 0.3), np.full((1, 1), -0.2))
    rc, text = run(WARP, os.path.join(GOLD, "hashgauss.dat"), str(sh), str(out), "-q")
    assert rc == 0, text
    import oracle_lib as O
    s = vcimage.read1(out)
    assert np.max(np.abs(s - f32(O.shift_frac2d(golden["f1"], 0.3, -0.2)))) < 1e-6
