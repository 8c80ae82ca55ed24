import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def fixtures():
    """Decoded regression images of the reference (tests/golden/make_golden.py)."""
    return np.load(os.path.join(GOLDEN, "regression_fixtures.npz"), allow_pickle=False)


def gray_image(rows, dtype=np.uint8):
    """The reference's gray_image! literal (utils.rs:83-121): rows of pixels."""
    return np.array(rows, dtype=dtype).reshape(len(rows), -1) if rows else np.zeros((0, 0), dtype)


def gray_bench_image(width, height):
    """utils.rs:671-680: intensity (x % 7 + y % 6)."""
    y, x = np.mgrid[0:height, 0:width]
    return ((x % 7) + (y % 6)).astype(np.uint8)


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
