"""integral_image / integral_squared_image (integral_image.rs:50-152) and
adaptive_threshold (contrast.rs:19-47) on the GPU, bit for bit: the
reference's doc and unit-test examples, its zebra truth image
(tests/regression.rs:576-582), the oracle on ragged random shapes, and numpy's
wrapping uint32 cumulative sums at sizes where the u32 sums wrap."""
import numpy as np
import pytest
import torch

import oracle as O
from imageproc_amd import contrast, device, integral_image as ii
from imageproc_amd._lib import Panic

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (2, 3), (64, 64), (65, 63), (63, 129), (130, 257), (300, 1), (1, 300),
          (777, 1023), (3, 16), (65, 128), (130, 256), (1000, 1024), (70, 4112)]


def _cumsum_u32(img, squared):
    v = img.astype(np.uint32)
    if squared:
        v = v * v
    out = np.zeros((img.shape[0] + 1, img.shape[1] + 1), np.uint32)
    out[1:, 1:] = v.cumsum(0, dtype=np.uint32).cumsum(1, dtype=np.uint32)
    return out


def test_doc_examples():
    img = np.array([[1, 2, 3], [4, 5, 6]], np.uint8)
    # integral_image.rs:32-41 and 66-80
    assert np.array_equal(ii.integral_image(img),
                          np.array([[0, 0, 0, 0], [0, 1, 3, 6], [0, 5, 12, 21]], np.uint32))
    assert np.array_equal(ii.integral_squared_image(img),
                          np.array([[0, 0, 0, 0], [0, 1, 5, 14], [0, 17, 46, 91]], np.uint32))


@pytest.mark.parametrize("shape", SHAPES)
def test_integral_vs_oracle(shape):
    print("integral", shape, flush=True)
    img = np.random.default_rng(shape[0] * 7 + shape[1]).integers(0, 256, shape, dtype=np.uint8)
    assert np.array_equal(ii.integral_image(img), O.integral_image(img))
    assert np.array_equal(ii.integral_squared_image(img), O.integral_image(img, True))


def test_integral_empty():
    for shape in [(0, 5), (5, 0), (0, 0)]:
        img = np.zeros(shape, np.uint8)
        want = np.zeros((shape[0] + 1, shape[1] + 1), np.uint32)
        assert np.array_equal(ii.integral_image(img), want)
        assert np.array_equal(ii.integral_squared_image(img), want)


def test_integral_wraps_and_large():
    # 255^2 * 300 * 400 is above 2^32: the squared table wraps like the
    # reference's release-mode u32 additions.
    img = np.full((300, 400), 255, np.uint8)
    assert np.array_equal(ii.integral_squared_image(img), _cumsum_u32(img, True))
    img = np.random.default_rng(5).integers(0, 256, (2051, 3001), dtype=np.uint8)
    assert np.array_equal(ii.integral_image(img), _cumsum_u32(img, False))
    assert np.array_equal(ii.integral_squared_image(img), _cumsum_u32(img, True))


def test_integral_device_entry_matches_host_entry():
    img = np.random.default_rng(9).integers(0, 256, (517, 389), dtype=np.uint8)
    d = torch.from_numpy(img).cuda()
    for sq in (False, True):
        got = device.integral_image(d, sq).cpu().numpy().view(np.uint32)
        assert np.array_equal(got, _cumsum_u32(img, sq))
    # a second, smaller call through the same cached workspace
    small = torch.from_numpy(img[:70, :90].copy()).cuda()
    got = device.integral_image(small).cpu().numpy().view(np.uint32)
    assert np.array_equal(got, _cumsum_u32(img[:70, :90], False))


@pytest.mark.parametrize("shape", [(300, 512), (129, 400), (64, 48)])
def test_integral_unaligned_buffers(shape):
    """The device entry on an image and a table that are not 16-byte aligned
    (the kernels then load bytes and store single elements)."""
    h, w = shape
    img = np.random.default_rng(w).integers(0, 256, shape, dtype=np.uint8)
    flat = torch.zeros(h * w + 16, dtype=torch.uint8, device="cuda")
    for src_off in (0, 1):
        src = flat[src_off:src_off + h * w].view(h, w)
        src.copy_(torch.from_numpy(img))
        for out_off in (0, 1, 2, 3):
            buf = torch.zeros((h + 1) * (w + 1) + 4, dtype=torch.int32, device="cuda")
            out = buf[out_off:out_off + (h + 1) * (w + 1)].view(h + 1, w + 1)
            for sq in (False, True):
                device.integral_image(src, sq, out=out)
                got = out.cpu().numpy().view(np.uint32)
                assert np.array_equal(got, _cumsum_u32(img, sq)), (src_off, out_off, sq)


def test_adaptive_zebra_truth(fixtures):
    assert np.array_equal(contrast.adaptive_threshold(fixtures["zebra"], 41, 0),
                          fixtures["zebra_adaptive_threshold"])


def test_adaptive_reference_unit_tests():
    # contrast.rs:471-540
    assert (contrast.adaptive_threshold(np.full((3, 3), 100, np.uint8), 1, 0) == 255).all()
    for y in range(3):
        for x in range(3):
            img = np.full((3, 3), 200, np.uint8)
            img[y, x] = 100
            exp = np.full((3, 3), 255, np.uint8)
            exp[y, x] = 0
            assert np.array_equal(contrast.adaptive_threshold(img, 1, 0), exp)
    for y in range(5):
        for x in range(5):
            img = np.full((5, 5), 100, np.uint8)
            img[y, x] = 200
            assert np.array_equal(contrast.adaptive_threshold(img, 1, 0),
                                  O.adaptive_threshold(img, 1, 0))
    img = np.full((3, 3), 100, np.uint8)
    img[2, 2] = 0
    assert (contrast.adaptive_threshold(img, 1, 100) == 255).all()
    exp = np.full((3, 3), 255, np.uint8)
    exp[2, 2] = 0
    assert np.array_equal(contrast.adaptive_threshold(img, 1, 50), exp)


@pytest.mark.parametrize("shape,r,delta", [((37, 53), 1, 0), ((130, 257), 7, -5), ((300, 1), 3, 2),
                                           ((1, 300), 9, 0), ((64, 65), 100, 10),
                                           ((513, 700), 41, 0), ((20, 30), 2 ** 31 - 1, -300),
                                           ((90, 70), 5, 2 ** 31 - 1), ((90, 70), 5, -2 ** 31)])
def test_adaptive_vs_oracle(shape, r, delta):
    print("adaptive", shape, r, delta, flush=True)
    img = np.random.default_rng(shape[1] + r % 1000).integers(0, 256, shape, dtype=np.uint8)
    want = O.adaptive_threshold(img, r, delta)
    assert np.array_equal(contrast.adaptive_threshold(img, r, delta), want)
    got = device.adaptive_threshold(torch.from_numpy(img).cuda(), r, delta).cpu().numpy()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape,r", [((200, 256), 2), ((131, 300), 4), ((300, 517), 63),
                                     ((129, 130), 64), ((150, 160), 65), ((64, 64), 64),
                                     ((70, 1), 3), ((1, 70), 3), ((1024, 1024), 41),
                                     ((100, 208), 5), ((77, 48), 30)])
def test_adaptive_radii_around_the_block_path(shape, r):
    """Radii up to 64 take the block-sum kernel, larger ones the integral
    table; both against the oracle, through the host and the device entry."""
    print("adaptive radius", shape, r, flush=True)
    img = np.random.default_rng(r * 31 + shape[1]).integers(0, 256, shape, dtype=np.uint8)
    for delta in (0, -3, 7):
        want = O.adaptive_threshold(img, r, delta)
        assert np.array_equal(contrast.adaptive_threshold(img, r, delta), want), delta
    got = device.adaptive_threshold(torch.from_numpy(img).cuda(), r, 0).cpu().numpy()
    assert np.array_equal(got, O.adaptive_threshold(img, r, 0))


def test_adaptive_unaligned_buffers():
    h, w = 97, 256
    img = np.random.default_rng(3).integers(0, 256, (h, w), dtype=np.uint8)
    want = O.adaptive_threshold(img, 9, 2)
    flat = torch.zeros(h * w + 32, dtype=torch.uint8, device="cuda")
    for off in (0, 1, 5):
        src = flat[off:off + h * w].view(h, w)
        src.copy_(torch.from_numpy(img))
        obuf = torch.zeros(h * w + 32, dtype=torch.uint8, device="cuda")
        out = obuf[3:3 + h * w].view(h, w)
        device.adaptive_threshold(src, 9, 2, out=out)
        assert np.array_equal(out.cpu().numpy(), want), off


def test_adaptive_errors_and_empty():
    img = np.zeros((4, 4), np.uint8)
    with pytest.raises(Panic, match="block_radius > 0"):
        contrast.adaptive_threshold(img, 0, 0)
    with pytest.raises(Panic, match="2\\^31"):
        contrast.adaptive_threshold(img, 2 ** 31, 0)
    assert contrast.adaptive_threshold(np.zeros((0, 7), np.uint8), 3, 0).shape == (0, 7)
