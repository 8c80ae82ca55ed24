"""CPU tests of the C ABI and of the host-side mirror: the library loads,
exports every symbol include/imageproc_hip.h declares, computes the Gaussian
taps on the host bit-identically to the oracle, and rejects every argument
the reference panics on before any GPU work is attempted."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from conftest import ROOT

import imageproc_amd as ip
from imageproc_amd import _lib
from imageproc_amd.filter import (box_filter, gaussian_blur_f32, gaussian_kernel_f32,
                                  separable_filter)
from imageproc_amd.template_matching import MatchTemplateMethod, find_extremes, match_template


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "imageproc_hip.h")).read()
    return sorted(set(re.findall(r"IPH_API\s+[\w\s\*]+?\b(iph_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_lib.EXPORTED) == declared


def test_library_holds_gfx950_code_only():
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data
    for other in (b"gfx942", b"gfx90a", b"gfx1100"):
        assert other not in data


@pytest.mark.parametrize("sigma", [0.5, 0.7, 1.0, 1.4, 2.0, 3.0, 4.0, 7.25, 8.0, 10.0, 63.0])
def test_host_gaussian_taps_are_bit_identical_to_the_oracle(sigma):
    a = gaussian_kernel_f32(sigma)
    b = O.gaussian_kernel_f32(sigma)
    assert a.dtype == np.float32 and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_reference_panics_are_raised_before_gpu_work():
    img = np.zeros((4, 4), np.uint8)
    with pytest.raises(ip.Panic, match="sigma must be > 0.0"):
        gaussian_blur_f32(img, 0.0)
    with pytest.raises(ip.Panic, match="sigma must be > 0.0"):
        gaussian_blur_f32(img, -0.5)
    with pytest.raises(ip.Panic, match="sigma must be > 0.0"):
        gaussian_blur_f32(img, float("nan"))
    with pytest.raises(ip.Panic, match="the two 1D kernels must be the same length"):
        separable_filter(img, np.ones(3, np.float32), np.ones(5, np.float32))
    with pytest.raises(ip.Panic, match="image width must be greater than or equal to template width"):
        match_template(np.zeros((5, 5), np.uint8), np.zeros((5, 6), np.uint8),
                       MatchTemplateMethod.SumOfSquaredErrors)
    with pytest.raises(ip.Panic, match="image height must be greater than or equal to template height"):
        match_template(np.zeros((5, 5), np.uint8), np.zeros((6, 5), np.uint8),
                       MatchTemplateMethod.SumOfSquaredErrors)
    with pytest.raises(ip.Panic, match="image must be non-empty"):
        find_extremes(np.zeros((0, 3), np.float32))


def test_kernel_new_panics():
    # kernel.rs:24-33, filter/mod.rs:755-760, kernel.rs:95-106
    with pytest.raises(ip.Panic, match="width and height must be non-zero"):
        ip.Kernel([], 0, 0)
    with pytest.raises(ip.Panic):
        ip.Kernel([1, 2, 3], 2, 2)
    with pytest.raises(ip.Panic):
        ip.Kernel(np.zeros(2, np.float32), 2, 2_147_483_649)


def test_type_combinations_the_reference_rejects():
    img = np.zeros((3, 3), np.float32)
    with pytest.raises(TypeError):
        separable_filter(img, np.ones(3, np.int32), np.ones(3, np.int32))
    with pytest.raises(TypeError):
        ip.filter.filter_clamped(np.zeros((3, 3), np.uint8), ip.Kernel([1.0], 1, 1), np.int16)


def test_empty_images_need_no_device():
    # box_filter.rs:77-82 and filter/mod.rs:779-789: empty in, empty out
    for shape in [(0, 0), (1, 0), (0, 1)]:
        assert box_filter(np.zeros(shape, np.uint8), 3, 3).shape == shape
        assert gaussian_blur_f32(np.zeros(shape, np.uint8), 1.0).shape == shape
        assert ip.filter.filter_clamped(np.zeros(shape, np.uint8),
                                        ip.Kernel([2], 1, 1)).shape == shape


def test_batch_argument_errors_before_device():
    """Batch jobs are validated when added, with the single calls' messages
    (filter/mod.rs:157 sigma assert, kernel.rs:24-35 Kernel::new)."""
    import ctypes as C

    from imageproc_amd._lib import Panic, check, lib

    h = C.c_void_p()
    check(lib().iph_batch_create(C.byref(h)))
    try:
        with pytest.raises(Panic, match="sigma must be > 0.0"):
            check(lib().iph_batch_add_gaussian_blur(h, None, None, 8, 8, 0.0))
        with pytest.raises(TypeError):
            check(lib().iph_batch_add_gaussian_blur(h, None, None, 8, 8, 8.5))  # 35 taps
        k = (C.c_float * 4)()
        with pytest.raises(Panic, match="width and height must be non-zero"):
            check(lib().iph_batch_add_filter_clamped(h, None, None, 8, 8, k, 0, 4))
        with pytest.raises(Panic, match="image width must be greater than or equal"):
            check(lib().iph_batch_add_match_template(h, None, 4, 8, None, 5, 5, 0, None))
        with pytest.raises(Panic, match="not planned"):
            check(lib().iph_batch_run(h, None))
    finally:
        lib().iph_batch_destroy(h)


def test_proposal_search_argument_errors_before_device():
    """The one-call proposal search validates its sizes before any launch."""
    from imageproc_amd._lib import Panic, check, lib

    assert lib().iph_dev_proposal_search_workspace(8192, 8192, 48, 48, 3, 64, 16) > 2048 * 2048
    with pytest.raises(Panic, match="correlation methods"):
        check(lib().iph_dev_proposal_search(None, 64, 64, None, 8, 8, 0, None, None, 64, 16, 4,
                                            None, None, 0, None))
    with pytest.raises(Panic, match="templates of 4 to 128"):
        check(lib().iph_dev_proposal_search(None, 64, 64, None, 3, 8, 3, None, None, 64, 16, 4,
                                            None, None, 0, None))
    with pytest.raises(Panic, match="exact windows"):
        check(lib().iph_dev_proposal_search(None, 64, 64, None, 8, 8, 3, None, None, 8, 16, 4,
                                            None, None, 0, None))
    with pytest.raises(Panic, match="workspace too small"):
        check(lib().iph_dev_proposal_search(None, 64, 64, None, 8, 8, 3, None, None, 8, 4, 4,
                                            None, None, 0, None))


def test_masked_match_argument_errors_before_device():
    """template_matching.rs:430-437: the mask check comes first."""
    from imageproc_amd._lib import Panic, check, lib

    with pytest.raises(Panic, match="the template and mask must be the same size"):
        check(lib().iph_match_template_with_mask(None, 4, 4, None, 2, 2, None, 3, 2, 0, 0, None))
    with pytest.raises(Panic, match="image height must be greater than or equal"):
        check(lib().iph_match_template_with_mask(None, 4, 1, None, 2, 2, None, 2, 2, 0, 0, None))
    assert lib().iph_dev_match_template_with_mask_workspace(64, 64, 8, 8, 3, 2) >= 4
    assert lib().iph_dev_match_template_with_mask_workspace(64, 64, 8, 8, 3, 1) >= 64 * 64 * 8


def test_adaptive_threshold_argument_errors_before_device():
    """contrast.rs:20 asserts block_radius > 0 before any work."""
    from imageproc_amd._lib import Panic
    from imageproc_amd.contrast import adaptive_threshold

    img = np.zeros((3, 3), np.uint8)
    with pytest.raises(Panic, match="assertion failed: block_radius > 0"):
        adaptive_threshold(img, 0, 0)
    with pytest.raises(Panic, match="above 2\\^31 - 1"):
        adaptive_threshold(img, 2 ** 31, 0)
    with pytest.raises(TypeError):
        adaptive_threshold(img.astype(np.uint16), 1, 0)
    assert _lib.lib().iph_dev_integral_image_workspace(100, 70) >= 4 * (2 * 100 + 2 * 70)
    # radii up to 64 (after clipping to the larger side) need no table
    assert _lib.lib().iph_dev_adaptive_threshold_workspace(60, 50, 1000) == 0
    assert _lib.lib().iph_dev_adaptive_threshold_workspace(100, 70, 64) == 0
    assert _lib.lib().iph_dev_adaptive_threshold_workspace(100, 70, 65) >= 4 * 101 * 71


def test_canny_argument_errors_before_device():
    """edges.rs:31 asserts high_threshold >= low_threshold first."""
    from imageproc_amd._lib import Panic
    from imageproc_amd.edges import canny

    img = np.zeros((4, 4), np.uint8)
    with pytest.raises(Panic, match="assertion failed: high_threshold >= low_threshold"):
        canny(img, 5.0, 1.0)
    with pytest.raises(Panic, match="overflow"):
        canny(np.zeros((3, 0), np.uint8), 1.0, 5.0)
    assert _lib.lib().iph_dev_canny_workspace(100, 70) >= 2 * 100 * 70


def test_sum_image_pixels_and_variance_on_tables():
    """integral_image.rs:25-48, 273-301 and 513-535: the host helpers over
    tables (built here by the oracle)."""
    from imageproc_amd.integral_image import sum_image_pixels, variance

    img = np.array([[1, 2, 3], [4, 5, 6]], np.uint8)
    ii, sq = O.integral_image(img), O.integral_image(img, True)
    assert sum_image_pixels(ii, 1, 0, 2, 1) == 2 + 3 + 5 + 6
    assert sum_image_pixels(ii, 0, 0, 2, 0) == 1 + 2 + 3
    assert sum_image_pixels(sq, 1, 0, 2, 1) == 4 + 9 + 25 + 36
    mean = (2.0 + 3.0 + 5.0 + 6.0) / 4.0
    var = ((2.0 - mean) ** 2 + (3.0 - mean) ** 2 + (5.0 - mean) ** 2 + (6.0 - mean) ** 2) / 4.0
    assert variance(ii, sq, 1, 0, 2, 1) == var
    img = np.array([[1, 2], [3, 4]], np.uint8)  # integral_image.rs:513-535
    ii = O.integral_image(img)
    for (l, t, r, b), want in [((0, 0, 0, 0), 1), ((0, 0, 1, 0), 3), ((0, 0, 0, 1), 4),
                               ((0, 0, 1, 1), 10), ((1, 0, 1, 0), 2), ((1, 0, 1, 1), 6),
                               ((0, 1, 0, 1), 3), ((0, 1, 1, 1), 7), ((1, 1, 1, 1), 4)]:
        assert sum_image_pixels(ii, l, t, r, b) == want
    # wrapping: a squared table of 255s past 2^32 still gives exact window sums
    big = np.full((300, 400), 255, np.uint8)
    sq = O.integral_image(big, True)
    assert sum_image_pixels(sq, 10, 20, 41, 51) == 32 * 32 * 255 * 255


def test_device_entries_reject_null_buffers_before_device():
    from imageproc_amd._lib import Panic, check, lib

    with pytest.raises(Panic, match="null"):
        check(lib().iph_dev_integral_image(None, 4, 4, 0, None, None, 0, None))
    with pytest.raises(Panic, match="null"):
        check(lib().iph_dev_adaptive_threshold(None, None, 4, 4, 1, 0, None, 0, None))
    with pytest.raises(Panic, match="null"):
        check(lib().iph_dev_canny(None, None, 4, 4, 1.0, 2.0, None, 0, None))
