"""GPU parity of the filter kernels against the oracle and the reference's
own known answers and truth images. Integer and u8 outputs must be
bit-identical; f32 outputs too, since every kernel keeps the reference's
operation order."""
import numpy as np
import pytest

import oracle as O
from conftest import gray_bench_image, gray_image

import imageproc_amd as ip
from imageproc_amd import filter as flt

pytestmark = pytest.mark.gpu
F = np.float32


def _same(a, b):
    assert a.dtype == b.dtype and a.shape == b.shape
    if a.dtype == np.float32:
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    else:
        assert np.array_equal(a, b)


# ---- gaussian blur -------------------------------------------------------------

def test_blur_zebra_truths(fixtures):
    z = fixtures["zebra"]
    _same(flt.gaussian_blur_f32(z, 3.0), fixtures["zebra_gaussian_3"])
    _same(flt.gaussian_blur_f32(z, 10.0), fixtures["zebra_gaussian_10"])


def test_blur_f32_robin_sharpen_gaussian(fixtures):
    r = fixtures["robin"].astype(F)
    smooth = flt.gaussian_blur_f32(r, 0.7)
    _same(smooth, O.gaussian_blur_f32(r, 0.7))
    amount = F(7.0)
    v = (F(1.0) + amount) * r - amount * smooth
    out = np.where(v < 255, np.where(v > 0, v, 0), 255).astype(np.uint8)
    _same(out, fixtures["robin_sharpen_gaussian"])


@pytest.mark.parametrize("sigma", [0.5, 1.0, 2.0, 4.0, 8.0, 12.0, 30.0])
def test_blur_u8_random_bit_exact(rng, sigma):
    img = rng.integers(0, 256, (97, 131), dtype=np.uint8)
    _same(flt.gaussian_blur_f32(img, sigma), O.gaussian_blur_f32(img, sigma))


@pytest.mark.parametrize("sigma", [0.7, 3.0, 16.0])
def test_blur_f32_random_bit_exact(rng, sigma):
    img = (rng.standard_normal((70, 45)) * 40).astype(F)
    _same(flt.gaussian_blur_f32(img, sigma), O.gaussian_blur_f32(img, sigma))


def test_blur_config_1024x768_sigma2_bit_exact():
    # 1024x768 i.i.d. uniform u8 from a seed, sigma 2.0 (9 taps)
    img = np.random.default_rng(1).integers(0, 256, (768, 1024), dtype=np.uint8)
    _same(flt.gaussian_blur_f32(img, 2.0), O.gaussian_blur_f32(img, 2.0))


def test_blur_white_idempotent():
    img = np.full((12, 12), 255, np.uint8)
    _same(flt.gaussian_blur_f32(img, 6.0), img)
    imgf = np.ones((12, 12), F)
    assert np.max(np.abs(flt.gaussian_blur_f32(imgf, 6.0) - 1.0)) <= 1e-6


def test_blur_tiny_images_large_sigma():
    # proptest_gaussian_blur_f32 (filter/mod.rs:855-862): sizes 0..20, sigma up to 150
    rng = np.random.default_rng(7)
    for h, w, s in [(1, 1, 0.1), (3, 2, 5.0), (19, 7, 40.0), (5, 17, 150.0), (20, 20, 0.3)]:
        img = rng.integers(0, 256, (h, w), dtype=np.uint8)
        _same(flt.gaussian_blur_f32(img, s), O.gaussian_blur_f32(img, s))


# ---- separable / 1-D filters ------------------------------------------------------

def test_separable_kats():
    img = gray_image([[1, 2, 3], [4, 5, 6], [7, 8, 9]])
    k = np.full(3, F(1.0) / F(3.0), F)
    _same(flt.separable_filter_equal(img, k), gray_image([[2, 3, 3], [4, 5, 5], [6, 7, 7]]))
    ki = np.ones(3, np.int32)
    _same(flt.separable_filter_equal(img, ki),
          gray_image([[21, 27, 33], [39, 45, 51], [57, 63, 69]]))
    img2 = gray_image([[1, 4, 1], [4, 7, 4], [1, 4, 1]])
    _same(flt.horizontal_filter(img2, k), gray_image([[2, 2, 2], [5, 5, 5], [2, 2, 2]]))
    _same(flt.vertical_filter(img2, k), gray_image([[2, 5, 2], [2, 5, 2], [2, 5, 2]]))
    k10 = np.full(10, F(0.1), F)
    _same(flt.horizontal_filter(img2, k10), O.horizontal_filter(img2, k10))
    _same(flt.vertical_filter(img2, k10), O.vertical_filter(img2, k10))


def test_1d_filters_sweep_small_sizes():
    # filter/mod.rs:576-613: every H, W in 0..5 and kernel length in 0..15
    for h in range(5):
        for w in range(5):
            img = gray_bench_image(w, h)
            for L in range(15):
                k = (np.arange(L, dtype=F) % F(1.35)).astype(F)
                _same(flt.horizontal_filter(img, k), O.horizontal_filter(img, k))
                _same(flt.vertical_filter(img, k), O.vertical_filter(img, k))
                _same(flt.separable_filter(img, k, k[::-1].copy()),
                      O.separable_filter(img, k, k[::-1].copy()))


@pytest.mark.parametrize("L", [1, 2, 5, 8, 13, 15, 17, 21, 25, 47, 48, 49, 64, 101])
def test_separable_random_lengths(rng, L):
    img = rng.integers(0, 256, (83, 150), dtype=np.uint8)
    hk = rng.uniform(-0.3, 0.6, L).astype(F)
    vk = rng.uniform(-0.3, 0.6, L).astype(F)
    _same(flt.separable_filter(img, hk, vk), O.separable_filter(img, hk, vk))
    hi = rng.integers(-3, 5, L).astype(np.int32)
    vi = rng.integers(-3, 5, L).astype(np.int32)
    _same(flt.separable_filter(img, hi, vi), O.separable_filter(img, hi, vi))
    imgf = rng.standard_normal((40, 77)).astype(F)
    _same(flt.separable_filter(imgf, hk, vk), O.separable_filter(imgf, hk, vk))


# ---- filter_clamped ----------------------------------------------------------------

def test_filter_clamped_kats():
    img = gray_image([[3, 2, 1], [6, 5, 4], [9, 8, 7]])
    out = flt.filter_clamped(img, ip.kernel.SOBEL_HORIZONTAL_3X3, np.int16)
    _same(out, np.array([[-4, -8, -4]] * 3, np.int16))
    _same(flt.filter_clamped_parallel(img, ip.kernel.SOBEL_HORIZONTAL_3X3, np.int16), out)
    img2 = gray_image([[9, 4], [8, 1]])
    k = ip.Kernel(np.array([0.1, 0.2, 0.1, 0.2, 0.4, 0.2, 0.1, 0.2, 0.1], F), 3, 3)
    _same(flt.filter_clamped(img2, k, np.uint8), gray_image([[11, 7], [10, 5]]))
    img3 = gray_image([[3, 2], [4, 1]])
    _same(flt.filter_clamped(img3, ip.Kernel([1, 2], 2, 1), np.uint8), gray_image([[9, 7], [12, 6]]))
    img4 = gray_image([[1, 2, 3], [4, 5, 6], [7, 8, 9]])
    _same(flt.laplacian_filter(img4),
          O.filter_clamped(img4, ip.kernel.LAPLACIAN_3X3.data, 3, 3, np.int16))


def test_filter_clamped_regression_truths(fixtures):
    sharpen = ip.Kernel([0, -1, 0, -1, 5, -1, 0, -1, 0], 3, 3)
    _same(flt.filter_clamped(fixtures["robin"], sharpen, np.uint8), fixtures["robin_sharpen3x3"])
    e = fixtures["elephant"]
    gh = flt.filter_clamped(e, ip.kernel.SOBEL_HORIZONTAL_3X3, np.int16).astype(F)
    gv = flt.filter_clamped(e, ip.kernel.SOBEL_VERTICAL_3X3, np.int16).astype(F)
    mag = np.sqrt(gh * gh + gv * gv).astype(np.uint16)
    _same(np.minimum(mag, 255).astype(np.uint8), fixtures["elephant_gradients"])


@pytest.mark.parametrize("kw,kh", [(1, 1), (3, 3), (5, 5), (3, 5), (5, 3), (2, 7), (16, 16),
                                   (17, 17), (40, 3)])
def test_filter_clamped_random(rng, kw, kh):
    img = rng.integers(0, 256, (61, 90), dtype=np.uint8)
    kf = rng.standard_normal(kw * kh).astype(F)
    kf /= kf.sum() if abs(kf.sum()) > 0.25 else 1.0
    kf = kf.astype(F)
    for st in (np.uint8, np.uint16, np.float32):
        _same(flt.filter_clamped(img, ip.Kernel(kf, kw, kh), st),
              O.filter_clamped(img, kf, kw, kh, st))
    ki = rng.integers(-9, 10, kw * kh).astype(np.int32)
    for st in (np.uint8, np.uint16, np.int16, np.int32):
        _same(flt.filter_clamped(img, ip.Kernel(ki, kw, kh), st),
              O.filter_clamped(img, ki, kw, kh, st))
    imgf = (rng.standard_normal((33, 29)) * 100).astype(F)
    for st in (np.uint8, np.float32):
        _same(flt.filter_clamped(imgf, ip.Kernel(kf, kw, kh), st),
              O.filter_clamped(imgf, kf, kw, kh, st))


def test_filter_clamped_small_kernels_tile_edges():
    """u8 images, f32 3/5-sided kernels, u8 outputs (the compile-time kernel):
    widths that take its 8-byte stores and its byte stores, partial last
    tiles, single rows and images narrower than the kernel."""
    rng = np.random.default_rng(23)
    for h, w in [(70, 256), (97, 131), (1, 8), (40, 3), (33, 64), (65, 1000)]:
        img = rng.integers(0, 256, (h, w), dtype=np.uint8)
        for kw, kh in [(3, 3), (3, 5), (5, 3), (5, 5)]:
            kf = (rng.standard_normal(kw * kh) * 0.4).astype(F)
            _same(flt.filter_clamped(img, ip.Kernel(kf, kw, kh), np.uint8),
                  O.filter_clamped(img, kf, kw, kh, np.uint8))


def test_filter_clamped_config_5x5_normalised():
    # 5x5 f32 kernel of N(0,1) weights normalised to sum 1 (redrawn if |sum| < 0.25)
    rng = np.random.default_rng(4)
    while True:
        k = rng.standard_normal(25)
        if abs(k.sum()) >= 0.25:
            break
    k = (k / k.sum()).astype(F)
    img = rng.integers(0, 256, (777, 1029), dtype=np.uint8)
    _same(flt.filter_clamped(img, ip.Kernel(k, 5, 5), np.uint8),
          O.filter_clamped(img, k, 5, 5, np.uint8, nthreads=8))


# ---- box filter ----------------------------------------------------------------------

def test_box_filter_kat():
    img = gray_image([[1, 2, 3], [4, 5, 6], [7, 8, 9]])
    _same(flt.box_filter(img, 1, 1), gray_image([[2, 3, 3], [4, 5, 5], [6, 7, 7]]))


@pytest.mark.parametrize("h,w,rx,ry", [(1, 1, 0, 0), (5, 7, 2, 3), (17, 3, 9, 1),
                                       (200, 199, 7, 7), (3, 50, 60, 4), (130, 70, 0, 99),
                                       (500, 500, 7, 7),
                                       # rows staged in LDS: one row of 256 segments per
                                       # workgroup, 16 rows of 1024, windows wider than the
                                       # row, window sums that wrap u32; 257 segments take
                                       # the per-thread rows pass
                                       (9, 16384, 3, 2), (64, 1024, 5, 5), (33, 260, 100, 1),
                                       (5, 100, (1 << 24) - 1, 1), (3, 16448, 2, 2),
                                       # one-kernel tiles (radii up to 32): partial tiles,
                                       # rows that are not 4-aligned, the largest radii,
                                       # a zero radius on either axis
                                       (130, 1001, 32, 32), (257, 300, 0, 32),
                                       (66, 129, 32, 0), (1000, 1003, 5, 5), (65, 4097, 31, 2)])
def test_box_filter_random(rng, h, w, rx, ry):
    img = rng.integers(0, 256, (h, w), dtype=np.uint8)
    _same(flt.box_filter(img, rx, ry), O.box_filter(img, rx, ry))


def test_box_filter_exact_on_saturated_windows():
    img = np.full((64, 70), 255, np.uint8)
    _same(flt.box_filter(img, 30, 31), img)


# ---- MFMA-estimated u8 blur: exactness stress ------------------------------------------

def test_mfma_blur_many_images_bit_exact():
    rng = np.random.default_rng(99)
    for i in range(24):
        h, w = int(rng.integers(1, 400)), int(rng.integers(1, 400))
        sigma = float(rng.uniform(0.3, 8.0))
        img = rng.integers(0, 256, (h, w), dtype=np.uint8)
        _same(flt.gaussian_blur_f32(img, sigma), O.gaussian_blur_f32(img, sigma))


def test_mfma_blur_flat_and_structured_images():
    # flat regions put every sum next to an integer: all outputs take the exact path
    for v in (0, 1, 37, 128, 254, 255):
        img = np.full((130, 260), v, np.uint8)
        _same(flt.gaussian_blur_f32(img, 2.0), O.gaussian_blur_f32(img, 2.0))
    y, x = np.mgrid[0:300, 0:500]
    img = (((x // 7) + (y // 5)) % 2 * 255).astype(np.uint8)
    for s in (1.0, 2.0, 5.5):
        _same(flt.gaussian_blur_f32(img, s), O.gaussian_blur_f32(img, s))


def test_mfma_separable_signed_and_unnormalised_taps():
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (257, 333), dtype=np.uint8)
    for L in (1, 2, 3, 9, 16, 33):
        hk = rng.uniform(-0.5, 1.0, L).astype(F)
        vk = rng.uniform(-0.5, 1.0, L).astype(F)
        _same(flt.separable_filter(img, hk, vk), O.separable_filter(img, hk, vk))
        big = (hk * 37.0).astype(F)
        _same(flt.separable_filter(img, big, big), O.separable_filter(img, big, big))


def test_mfma_blur_batch_device_api():
    import torch
    from imageproc_amd import device as dev
    rng = np.random.default_rng(3)
    imgs = rng.integers(0, 256, (5, 200, 333), dtype=np.uint8)
    out = dev.gaussian_blur_f32(torch.from_numpy(imgs).cuda(), 2.0).cpu().numpy()
    for i in range(5):
        _same(out[i], O.gaussian_blur_f32(imgs[i], 2.0))


# ---- streaming u8 blur: strips, bands, ragged edges, unequal taps ---------------------

def test_stream_blur_bands_strips_and_ragged_edges():
    """Shapes that cut the streaming kernel into several row bands and column
    strips, with widths and heights that are not multiples of 4, 16, 32 or
    128, and single-row / single-column images."""
    rng = np.random.default_rng(21)
    for h, w, sigma in [(1000, 37, 2.0), (2050, 131, 1.0), (77, 1029, 3.3), (4099, 260, 8.0),
                        (33, 4, 0.5), (300, 1, 2.0), (1, 300, 2.0), (31, 129, 7.6),
                        (768, 1024, 2.0), (65, 2049, 4.0), (4099, 1024, 2.0),
                        (1001, 528, 3.0)]:
        img = rng.integers(0, 256, (h, w), dtype=np.uint8)
        _same(flt.gaussian_blur_f32(img, sigma), O.gaussian_blur_f32(img, sigma))


@pytest.mark.parametrize("L", [1, 2, 3, 4, 7, 9, 16, 17, 24, 27, 29, 31, 32, 33])
def test_stream_separable_unequal_positive_taps(L):
    """Distinct non-negative horizontal and vertical taps (separate Toeplitz
    operands for the two passes), odd and even lengths up to 33."""
    rng = np.random.default_rng(100 + L)
    img = rng.integers(0, 256, (203, 301), dtype=np.uint8)
    hk = rng.uniform(0.0, 1.0, L).astype(F)
    vk = rng.uniform(0.0, 1.0, L).astype(F)
    hk = (hk / np.float32(hk.sum() * 1.001)).astype(F)
    vk = (vk / np.float32(vk.sum() * 1.001)).astype(F)
    _same(flt.separable_filter(img, hk, vk), O.separable_filter(img, hk, vk))
    _same(flt.separable_filter(img, hk, hk), O.separable_filter(img, hk, hk))


def test_wave_blur_device_batches_aligned_widths():
    """Device batches whose widths take each store form of the per-wave
    kernel (16-byte rows, 4-byte rows, byte rows), strips at both image
    edges, odd and even step counts and several row bands per image."""
    import torch
    from imageproc_amd import device as dev
    rng = np.random.default_rng(31)
    for n, h, w, sigma in [(3, 70, 1024, 2.0), (2, 96, 100, 1.2), (4, 33, 64, 4.0),
                           (2, 257, 36, 2.0), (1, 1000, 4, 0.8), (6, 45, 1040, 6.5)]:
        imgs = rng.integers(0, 256, (n, h, w), dtype=np.uint8)
        out = dev.gaussian_blur_f32(torch.from_numpy(imgs).cuda(), sigma).cpu().numpy()
        for i in range(n):
            _same(out[i], O.gaussian_blur_f32(imgs[i], sigma))
