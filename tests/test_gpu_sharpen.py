"""sharpen3x3 and sharpen_gaussian (filter/sharpen.rs:13-38) on the GPU:
the reference's truth images (tests/regression.rs:529-539) and the oracle's
composition on random images, bit for bit."""
import numpy as np
import pytest

import oracle as O
from imageproc_amd import filter as flt
from imageproc_amd._lib import Panic

pytestmark = pytest.mark.gpu
F = np.float32


def _oracle_sharpen_gaussian(img, sigma, amount):
    r = img.astype(F)
    smooth = O.gaussian_blur_f32(r, sigma)
    a = F(amount)
    v = (F(1.0) + a) * r - a * smooth
    return np.where(v < 255, np.where(v > 0, v, 0), 255).astype(np.uint8)


def test_robin_truths(fixtures):
    robin = fixtures["robin"]
    assert np.array_equal(flt.sharpen3x3(robin), fixtures["robin_sharpen3x3"])
    assert np.array_equal(flt.sharpen3x3_parallel(robin), fixtures["robin_sharpen3x3"])
    assert np.array_equal(flt.sharpen_gaussian(robin, 0.7, 7.0), fixtures["robin_sharpen_gaussian"])


# sigma 3.5 gives 15 taps, 6.0 gives 25 and 11.5 gives 47 (the fused
# kernel's loop over any tap count, up to its 48-tap limit, on shapes whose
# last 64 x 32 tile is partial), sigma 13 gives 53 taps (the unfused path
# through f32 images); the others take the fused kernel's
# 4-outputs-per-thread form
@pytest.mark.parametrize("shape,sigma,amount", [((37, 53), 1.5, 2.0), ((300, 411), 0.7, 7.0),
                                                ((128, 96), 4.0, 0.5), ((5, 200), 2.0, -1.0),
                                                ((70, 1), 1.0, 3.0), ((90, 140), 13.0, 1.0),
                                                ((65, 129), 0.4, 2.5), ((33, 70), 3.5, 1.0),
                                                ((50, 77), 2.5, 1.0), ((41, 90), 3.0, 2.0),
                                                ((97, 131), 6.0, 1.5), ((75, 150), 11.5, 0.8)])
def test_sharpen_gaussian_random(shape, sigma, amount):
    img = np.random.default_rng(shape[1]).integers(0, 256, shape, dtype=np.uint8)
    assert np.array_equal(flt.sharpen_gaussian(img, sigma, amount),
                          _oracle_sharpen_gaussian(img, sigma, amount))


def test_sharpen3x3_random_and_errors():
    img = np.random.default_rng(1).integers(0, 256, (61, 77), dtype=np.uint8)
    k = np.array([0, -1, 0, -1, 5, -1, 0, -1, 0], np.int32)
    want = O.filter_clamped(img, k, 3, 3, np.uint8)
    assert np.array_equal(flt.sharpen3x3(img), want)
    with pytest.raises(Panic, match="sigma must be > 0.0"):
        flt.sharpen_gaussian(img, 0.0, 1.0)


def test_device_entry_matches_host_entry():
    import torch

    from imageproc_amd import device

    from imageproc_amd._lib import lib

    # blurs of up to 48 taps run fused and need no workspace
    assert lib().iph_dev_sharpen_gaussian_workspace(389, 257, 11.5) == 0
    assert lib().iph_dev_sharpen_gaussian_workspace(389, 257, 13.0) >= 2 * 4 * 389 * 257
    img = np.random.default_rng(23).integers(0, 256, (257, 389), dtype=np.uint8)
    for sigma, amount in [(0.7, 7.0), (2.0, 1.5), (13.0, 0.5)]:
        got = device.sharpen_gaussian(torch.from_numpy(img).cuda(), sigma, amount).cpu().numpy()
        assert np.array_equal(got, flt.sharpen_gaussian(img, sigma, amount))
        assert np.array_equal(got, _oracle_sharpen_gaussian(img, sigma, amount))
