import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def golden():
    from pyflct_amd import vcimage
    f1, f2 = vcimage.read2(os.path.join(GOLDEN, "hashgauss.dat"))
    dx, dy = vcimage.read2(os.path.join(GOLDEN, "deltaxygauss.dat"))
    vx, vy, vm = vcimage.read3(os.path.join(GOLDEN, "testgaussvel.dat"))
    return {"f1": f1, "f2": f2, "dx": dx, "dy": dy, "vx": vx, "vy": vy, "vm": vm}


@pytest.fixture(scope="session")
def ctx():
    from pyflct_amd import api
    c = api.Context(0)
    yield c
    c.close()


def smooth_pair(rng, nx, ny, shift=(0.4, -0.3), width=2.0):
    """Smoothed random field and a shifted, slightly noisy copy."""
    from scipy.ndimage import gaussian_filter
    from pyflct_amd import synth
    a = gaussian_filter(rng.standard_normal((nx, ny)), width, mode="wrap")
    b = synth.shift_uniform(a, shift[0], shift[1])
    b = b + 0.05 * gaussian_filter(rng.standard_normal((nx, ny)), width, mode="wrap")
    return a, b


def f32(a):
    return np.asarray(a).astype(np.float32).astype(np.float64)
