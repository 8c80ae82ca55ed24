"""GPU parity of match_template and find_extremes.

* direct path: the reference's per-window f32 loop, bit-identical to the oracle;
* MFMA path, u8 images: exact integer sums, so CrossCorrelation and
  SumOfSquaredErrors equal the exact value rounded once to f32, and the
  normalised methods are within a few f32 ulps of the FP64 oracle;
* MFMA path, f32 images (BF16x3): within the stated tolerance of FP64.
"""
import numpy as np
import pytest

import oracle as O
from conftest import gray_image

from imageproc_amd.template_matching import MatchTemplateMethod as M
from imageproc_amd.template_matching import find_extremes, match_template

pytestmark = pytest.mark.gpu
F = np.float32

IMG = gray_image([[1, 4, 2], [2, 1, 3], [3, 3, 4]])
TPL = gray_image([[1, 2], [3, 4]])
PATHS = ["direct", "mfma"]


def _bits_equal(a, b):
    assert a.shape == b.shape and a.dtype == b.dtype == np.float32
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("path", PATHS)
def test_match_template_kats(path):
    # template_matching.rs:562-657
    tss = F(30.0)
    nz = lambda s, ii: F(s) / np.sqrt(F(ii) * tss)
    exp = {
        M.SumOfSquaredErrors: [[14, 14], [3, 1]],
        M.SumOfSquaredErrorsNormalized: [[nz(14, 22), nz(14, 30)], [nz(3, 23), nz(1, 35)]],
        M.CrossCorrelation: [[19, 23], [25, 32]],
        M.CrossCorrelationNormalized: [[nz(19, 22), nz(23, 30)], [nz(25, 23), nz(32, 35)]],
    }
    for m, e in exp.items():
        _bits_equal(match_template(IMG, TPL, m, path), np.array(e, F))


@pytest.mark.parametrize("path", PATHS)
def test_match_template_edge_cases(path):
    z = np.zeros((5, 5), np.uint8)
    _bits_equal(match_template(z, z, M.SumOfSquaredErrors, path), np.array([[0.0]], F))
    one = np.zeros((1, 1), np.uint8)
    _bits_equal(match_template(one, one, M.SumOfSquaredErrorsNormalized, path),
                np.array([[0.0]], F))
    # an all-zero image gives zero norms everywhere: the raw sum is returned
    img = np.zeros((9, 12), np.uint8)
    tpl = np.arange(12, dtype=np.uint8).reshape(3, 4)
    for m in M:
        _bits_equal(match_template(img, tpl, m, path), O.match_template(img, tpl, int(m)))


@pytest.mark.parametrize("shape,tshape", [((23, 31), (3, 4)), ((64, 70), (8, 8)),
                                          ((90, 80), (16, 17)), ((150, 140), (32, 32)),
                                          ((75, 300), (40, 9)), ((128, 128), (1, 1))])
def test_direct_path_is_bit_identical(rng, shape, tshape):
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    tpl = rng.integers(0, 256, tshape, dtype=np.uint8)
    for m in M:
        _bits_equal(match_template(img, tpl, m, "direct"), O.match_template(img, tpl, int(m)))


def test_direct_path_f32_images_bit_identical(rng):
    img = rng.standard_normal((60, 70)).astype(F)
    tpl = rng.standard_normal((9, 13)).astype(F)
    for m in M:
        _bits_equal(match_template(img, tpl, m, "direct"), O.match_template(img, tpl, int(m)))


# ---- find_extremes -----------------------------------------------------------------------

def test_find_extremes_kat():
    img = gray_image([[10, 7, 8, 1], [9, 15, 4, 2]])
    e = find_extremes(img)
    assert (e.max_value, e.min_value, e.max_value_location, e.min_value_location) == \
        (15, 1, (1, 1), (3, 0))


@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_find_extremes_ties_and_sizes(rng, dtype):
    for h, w in [(1, 1), (1, 1000), (1000, 1), (37, 29), (1024, 1023), (2017, 2017)]:
        img = rng.integers(0, 7, (h, w)).astype(dtype)
        mx, mn, mxl, mnl = O.find_extremes(img)
        e = find_extremes(img)
        assert (e.max_value, e.min_value) == (mx, mn)
        assert (e.max_value_location, e.min_value_location) == (mxl, mnl), (h, w)


def test_find_extremes_float_specials():
    img = np.array([[1.0, -0.0, 0.0, 5.0], [np.nan, -3.0, 5.0, -3.0]], F)
    e = find_extremes(img)
    assert (e.max_value, e.max_value_location) == (5.0, (3, 0))
    assert (e.min_value, e.min_value_location) == (-3.0, (1, 1))
    img2 = np.array([[0.0, -0.0], [-0.0, 0.0]], F)
    e2 = find_extremes(img2)
    r = O.find_extremes(img2)
    assert e2.min_value_location == r[3] and e2.max_value_location == r[2]
    assert np.signbit(e2.min_value) == np.signbit(r[1])
    nan0 = np.array([[np.nan, 1.0], [2.0, -1.0]], F)
    e3 = find_extremes(nan0)
    assert np.isnan(e3.max_value) and np.isnan(e3.min_value)
    assert e3.max_value_location == (0, 0) and e3.min_value_location == (0, 0)


# ---- int8 MFMA path (u8 images) ------------------------------------------------------

def _check_mfma_u8(img, tpl):
    n = tpl.size
    for m in M:
        got = match_template(img, tpl, m, "mfma")
        exact = O.match_template_f64(img, tpl, int(m))
        if m in (M.SumOfSquaredErrors, M.CrossCorrelation):
            # exact integer sums rounded once to f32
            _bits_equal(got, exact.astype(F))
        else:
            rel = np.abs(got - exact) / np.maximum(np.abs(exact), 1e-30)
            # f32 normalisation below 64x64 windows (~4 ulp), f64 at or above (<= 0.5 ulp)
            assert rel.max() <= (4.0 if n < 4096 else 0.5) * 2.0 ** -23 + 1e-30, (m, rel.max())
        if n * 65025 < 2 ** 24:
            # the reference's f32 sums are exact here: bit-identical to it
            _bits_equal(got, O.match_template(img, tpl, int(m)))


@pytest.mark.parametrize("shape,tshape", [((40, 40), (1, 1)), ((50, 61), (3, 5)),
                                          ((100, 130), (16, 16)), ((161, 200), (32, 32)),
                                          ((97, 97), (33, 31)), ((150, 300), (8, 96)),
                                          ((260, 190), (96, 64)), ((300, 310), (128, 128)),
                                          ((70, 70), (64, 70))])
def test_mfma_u8_random(rng, shape, tshape):
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    tpl = rng.integers(0, 256, tshape, dtype=np.uint8)
    _check_mfma_u8(img, tpl)


def test_mfma_u8_saturated_and_constant():
    # extreme operands for the -128 offset algebra: all 255 / all 0 / mixed
    for iv, tv in [(255, 255), (0, 255), (255, 0), (0, 0), (128, 127)]:
        img = np.full((70, 90), iv, np.uint8)
        tpl = np.full((32, 32), tv, np.uint8)
        _check_mfma_u8(img, tpl)


def test_mfma_u8_unaligned_widths(rng):
    for w in (33, 129, 130, 131):
        img = rng.integers(0, 256, (45, w), dtype=np.uint8)
        tpl = rng.integers(0, 256, (7, 9), dtype=np.uint8)
        _check_mfma_u8(img, tpl)


def test_mfma_u8_config_2048_argmax_matches_direct():
    import bench
    img, tpl, loc = bench.match_workload()
    for m in M:
        a = match_template(img, tpl, m, "mfma")
        d = match_template(img, tpl, m, "direct")
        pick = np.argmin if m in (M.SumOfSquaredErrors, M.SumOfSquaredErrorsNormalized) else np.argmax
        assert pick(a) == pick(d), m
    cc = match_template(img, tpl, M.CrossCorrelationNormalized, "mfma")
    assert tuple(np.unravel_index(np.argmax(cc), cc.shape)[::-1]) == loc


# ---- BF16x3 MFMA path (f32 images) -----------------------------------------------------

def _check_bf16x3(img, tpl, ccn_tol=1e-4):
    """BF16x3 keeps ~16 significant bits per product: |C - C_exact| is bounded
    by 2^-14 * sum|I||T|, so the normalised correlation is within ~1e-4 of
    FP64 (Cauchy-Schwarz), and SSE within twice the correlation bound."""
    from numpy.lib.stride_tricks import sliding_window_view
    win = sliding_window_view(np.abs(img.astype(np.float64)), tpl.shape)
    bound = np.einsum("yxij,ij->yx", win, np.abs(tpl.astype(np.float64))) * 2.0 ** -14 + 1e-6
    for m in M:
        got = match_template(img, tpl, m, "mfma").astype(np.float64)
        exact = O.match_template_f64(img, tpl, int(m))
        err = np.abs(got - exact)
        if m == M.CrossCorrelation:
            assert np.all(err <= bound), (m, (err / bound).max())
        elif m == M.SumOfSquaredErrors:
            assert np.all(err <= 2 * bound + 1e-6 * np.abs(exact)), (m, (err / bound).max())
        elif m == M.CrossCorrelationNormalized:
            assert err.max() <= ccn_tol, (m, err.max())
        else:
            sse_bound = 2 * bound + 1e-6 * np.abs(O.match_template_f64(img, tpl, 0))
            norm = np.abs(O.match_template_f64(img, tpl, 0)) / np.maximum(np.abs(exact), 1e-30)
            assert np.all(err <= sse_bound / np.maximum(norm, 1e-30) + 1e-6), m


@pytest.mark.parametrize("shape,tshape", [((40, 50), (3, 3)), ((90, 100), (16, 16)),
                                          ((130, 120), (33, 65)), ((200, 210), (64, 64)),
                                          ((300, 290), (128, 128)), ((100, 180), (20, 100))])
def test_bf16x3_f32_random(rng, shape, tshape):
    img = rng.standard_normal(shape).astype(F)
    tpl = rng.standard_normal(tshape).astype(F)
    _check_bf16x3(img, tpl)


def test_bf16x3_argmax_matches_direct_on_smooth_field(rng):
    # smoothed N(0,1) field, 64x64 template cropped and perturbed by N(0, 0.05)
    from imageproc_amd import filter as flt
    img = flt.gaussian_blur_f32(rng.standard_normal((512, 512)).astype(F), 3.0)
    y, x = 200, 137
    tpl = (img[y:y + 64, x:x + 64] + 0.05 * rng.standard_normal((64, 64))).astype(F)
    a = match_template(img, tpl, M.CrossCorrelationNormalized, "mfma")
    d = match_template(img, tpl, M.CrossCorrelationNormalized, "direct")
    assert np.argmax(a) == np.argmax(d)
    assert np.unravel_index(np.argmax(a), a.shape) == (y, x)


# ---- default path for every template the reference accepts ----------------------------

@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
@pytest.mark.parametrize("shape,tshape", [((90, 300), (40, 129)), ((230, 70), (200, 40)),
                                          ((220, 240), (200, 200)), ((129, 129), (129, 129))])
def test_auto_path_takes_templates_above_128(rng, dtype, shape, tshape):
    """template_matching.rs:394-404 accepts any template no larger than the
    image: the default path runs templates with a side above 128 on the direct
    kernel, bit-identical to the oracle, and the MFMA path refuses them."""
    if dtype == np.uint8:
        img = rng.integers(0, 256, shape, dtype=np.uint8)
        tpl = rng.integers(0, 256, tshape, dtype=np.uint8)
    else:
        img = rng.standard_normal(shape).astype(F)
        tpl = rng.standard_normal(tshape).astype(F)
    for m in M:
        _bits_equal(match_template(img, tpl, m), O.match_template(img, tpl, int(m), nthreads=8))
    with pytest.raises(TypeError, match="MFMA path takes templates up to 128x128"):
        match_template(img, tpl, M.CrossCorrelation, "mfma")


@pytest.mark.parametrize("tshape", [(128, 128), (128, 5), (3, 128)])
def test_auto_path_keeps_mfma_up_to_128(rng, tshape):
    img = rng.integers(0, 256, (150, 170), dtype=np.uint8)
    tpl = rng.integers(0, 256, tshape, dtype=np.uint8)
    for m in M:
        np.testing.assert_array_equal(match_template(img, tpl, m),
                                      match_template(img, tpl, m, "mfma"))
