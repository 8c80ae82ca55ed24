"""Pins the CPU oracle (oracle/) before it is trusted as the GPU checker.

Each oracle function is checked against what the reference itself holds:
its known-answer tests (quoted with file:line), its regression truth images
(tests/golden/regression_fixtures.npz), or an independent formulation
(exact integer window sums, first-occurrence argmin/argmax, FP64 sums), so
that a dropped term, a wrong clamp index or a transposed operand fails here.
"""
import numpy as np
import pytest
from numpy.lib.stride_tricks import sliding_window_view

import oracle as O
from conftest import gray_bench_image, gray_image

F = np.float32


# ---- Gaussian kernel and blur -------------------------------------------------

def test_blur_matches_zebra_truths(fixtures):
    # tests/regression.rs:562-574, tolerance 0
    z = fixtures["zebra"]
    assert np.array_equal(O.gaussian_blur_f32(z, 3.0), fixtures["zebra_gaussian_3"])
    assert np.array_equal(O.gaussian_blur_f32(z, 10.0), fixtures["zebra_gaussian_10"])


def test_f32_blur_matches_robin_sharpen_gaussian(fixtures):
    # sharpen_gaussian (filter/sharpen.rs:31-38) over the f32 blur path,
    # tests/regression.rs:534-539
    r = fixtures["robin"].astype(F)
    smooth = O.gaussian_blur_f32(r, 0.7)
    amount = F(7.0)
    v = (F(1.0) + amount) * r - amount * smooth
    out = np.where(v < 255, np.where(v > 0, v, 0), 255).astype(np.uint8)
    assert np.array_equal(out, fixtures["robin_sharpen_gaussian"])


@pytest.mark.parametrize("sigma", [0.5, 0.7, 1.0, 2.0, 3.0, 4.0, 6.0, 8.0, 10.0])
def test_gaussian_kernel_shape_and_values(sigma):
    k = O.gaussian_kernel_f32(sigma)
    r = int(np.ceil(2.0 * sigma))
    assert k.size == 2 * r + 1                      # filter/mod.rs:133-134
    assert np.array_equal(k, k[::-1])               # mirrored taps
    x = np.arange(-r, r + 1, dtype=np.float64)
    exact = np.exp(-x * x / (2.0 * sigma * sigma))
    exact /= exact.sum()
    assert np.max(np.abs(k - exact) / exact) < 2e-6  # f32 evaluation of the same Gaussian
    assert abs(float(k.astype(np.float64).sum()) - 1.0) < 1e-6


def test_gaussian_on_u8_white_idempotent():
    # filter/mod.rs:834-840
    img = np.full((12, 12), 255, np.uint8)
    assert np.array_equal(O.gaussian_blur_f32(img, 6.0), img)


def test_gaussian_on_f32_white_idempotent():
    # filter/mod.rs:842-847
    img = np.ones((12, 12), F)
    assert np.max(np.abs(O.gaussian_blur_f32(img, 6.0) - 1.0)) <= 1e-6


# ---- 1-D and separable filters ------------------------------------------------

def test_separable_filter_kat():
    # filter/mod.rs:479-496
    img = gray_image([[1, 2, 3], [4, 5, 6], [7, 8, 9]])
    k = np.full(3, F(1.0) / F(3.0), F)
    assert np.array_equal(O.separable_filter(img, k, k),
                          gray_image([[2, 3, 3], [4, 5, 5], [6, 7, 7]]))


def test_separable_filter_integer_kernel_kat():
    # filter/mod.rs:498-514
    img = gray_image([[1, 2, 3], [4, 5, 6], [7, 8, 9]])
    k = np.ones(3, np.int32)
    assert np.array_equal(O.separable_filter(img, k, k),
                          gray_image([[21, 27, 33], [39, 45, 51], [57, 63, 69]]))


def test_horizontal_and_vertical_filter_kats():
    # filter/mod.rs:615-631 and 643-659
    img = gray_image([[1, 4, 1], [4, 7, 4], [1, 4, 1]])
    k = np.full(3, F(1.0) / F(3.0), F)
    assert np.array_equal(O.horizontal_filter(img, k), gray_image([[2, 2, 2], [5, 5, 5], [2, 2, 2]]))
    assert np.array_equal(O.vertical_filter(img, k), gray_image([[2, 5, 2], [2, 5, 2], [2, 5, 2]]))
    k10 = np.full(10, F(1.0) / F(10.0), F)  # wider/taller than the image, :633-642, :661-670
    O.horizontal_filter(img, k10)
    O.vertical_filter(img, k10)


def _hfilter_padded(img, k):
    """Independent indexing: pad by edge replication, then slide."""
    h, w = img.shape
    L = len(k)
    half = L // 2
    out = np.zeros((h, w), np.uint8)
    if w == 0 or h == 0:
        return out
    pad = np.pad(img, ((0, 0), (half, max(L - 1 - half, 0))), mode="edge").astype(F)
    acc = np.zeros((h, w), F)
    for i in range(L):
        acc = acc + pad[:, i:i + w] * F(k[i])
    return np.where(acc < 255, np.where(acc > 0, np.trunc(acc), 0), 255).astype(np.uint8)


def test_1d_filters_sweep_small_sizes():
    # the enumeration of filter/mod.rs:576-613: H, W in 0..5, len in 0..15
    for h in range(5):
        for w in range(5):
            img = gray_bench_image(w, h)
            for L in range(15):
                k = (np.arange(L, dtype=F) % F(1.35)).astype(F)
                assert np.array_equal(O.horizontal_filter(img, k), _hfilter_padded(img, k))
                assert np.array_equal(O.vertical_filter(img, k), _hfilter_padded(img.T, k).T)


# ---- 2-D filter_clamped ---------------------------------------------------------

def test_filter_clamped_sobel_to_i16_kat():
    # filter/mod.rs:671-693
    img = gray_image([[3, 2, 1], [6, 5, 4], [9, 8, 7]])
    k = np.array([-1, 0, 1, -2, 0, 2, -1, 0, 1], np.int32)
    out = O.filter_clamped(img, k, 3, 3, np.int16)
    assert np.array_equal(out, np.array([[-4, -8, -4]] * 3, np.int16))
    assert np.array_equal(O.filter_clamped(img, k, 3, 3, np.int16, nthreads=3), out)


def test_filter_kernel_larger_than_image_kat():
    # filter/mod.rs:791-811
    img = gray_image([[9, 4], [8, 1]])
    k = np.array([0.1, 0.2, 0.1, 0.2, 0.4, 0.2, 0.1, 0.2, 0.1], F)
    assert np.array_equal(O.filter_clamped(img, k, 3, 3, np.uint8), gray_image([[11, 7], [10, 5]]))


def test_filter_even_kernel_side_kat():
    # filter/mod.rs:762-777 uses K = u8 with an identity map; the sums (<= 12)
    # are the same at K = i32 clamped to u8.
    img = gray_image([[3, 2], [4, 1]])
    out = O.filter_clamped(img, np.array([1, 2], np.int32), 2, 1, np.uint8)
    assert np.array_equal(out, gray_image([[9, 7], [12, 6]]))


def test_laplacian_kat():
    # filter/mod.rs:460-477 (laplacian_filter == filter_clamped to i16)
    img = gray_image([[1, 2, 3], [4, 5, 6], [7, 8, 9]])
    k = np.array([0, 1, 0, 1, -4, 1, 0, 1, 0], np.int32)
    out = O.filter_clamped(img, k, 3, 3, np.int16)
    p = np.pad(img.astype(np.int32), 1, mode="edge")
    exp = p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:] - 4 * p[1:-1, 1:-1]
    assert np.array_equal(out, exp.astype(np.int16))


def test_filter_clamped_matches_regression_truths(fixtures):
    # sharpen3x3, filter/sharpen.rs:13-16, tests/regression.rs:529-531
    k = np.array([0, -1, 0, -1, 5, -1, 0, -1, 0], np.int32)
    assert np.array_equal(O.filter_clamped(fixtures["robin"], k, 3, 3, np.uint8),
                          fixtures["robin_sharpen3x3"])
    # sobel gradients, gradients.rs:99-160, tests/regression.rs:500-514
    e = fixtures["elephant"]
    gh = O.filter_clamped(e, np.array([-1, 0, 1, -2, 0, 2, -1, 0, 1], np.int32), 3, 3, np.int16)
    gv = O.filter_clamped(e, np.array([-1, -2, -1, 0, 0, 0, 1, 2, 1], np.int32), 3, 3, np.int16)
    gh, gv = gh.astype(F), gv.astype(F)
    mag = np.sqrt(gh * gh + gv * gv).astype(np.uint16)
    assert np.array_equal(np.minimum(mag, 255).astype(np.uint8), fixtures["elephant_gradients"])


def test_filter_clamped_rejects_types_the_reference_rejects():
    img = gray_image([[1]])
    with pytest.raises(AssertionError):
        O.filter_clamped(img.astype(F), np.ones(1, np.int32), 1, 1, np.uint8)


# ---- box filter -------------------------------------------------------------------

def test_box_filter_kat():
    # box_filter.rs:84-101
    img = gray_image([[1, 2, 3], [4, 5, 6], [7, 8, 9]])
    assert np.array_equal(O.box_filter(img, 1, 1), gray_image([[2, 3, 3], [4, 5, 5], [6, 7, 7]]))


def _box_direct(img, rx, ry):
    """Independent form: direct window sums over an edge-padded image."""
    def one(a, r):
        h, w = a.shape
        p = np.pad(a.astype(np.int64), ((0, 0), (r, r)), mode="edge")
        s = sliding_window_view(p, 2 * r + 1, axis=1).sum(axis=-1)
        return (s // (2 * r + 1)).astype(np.uint8)
    return one(one(img, rx).T, ry).T


def test_box_filter_matches_direct_window_sums(rng):
    for h, w, rx, ry in [(1, 1, 0, 0), (5, 7, 2, 3), (17, 3, 9, 1), (40, 33, 7, 7), (3, 50, 60, 4)]:
        img = rng.integers(0, 256, (h, w), dtype=np.uint8)
        assert np.array_equal(O.box_filter(img, rx, ry), _box_direct(img, rx, ry)), (h, w, rx, ry)
    for shape in [(0, 0), (1, 0), (0, 1)]:  # box_filter.rs:77-82
        O.box_filter(np.zeros(shape, np.uint8), 3, 3)


# ---- template matching --------------------------------------------------------------

IMG = gray_image([[1, 4, 2], [2, 1, 3], [3, 3, 4]])
TPL = gray_image([[1, 2], [3, 4]])
MASK = gray_image([[0, 1], [2, 3]])


def _f(rows):
    return np.array(rows, F)


def test_match_template_kats():
    # template_matching.rs:562-657
    tss = F(30.0)
    nz = lambda s, ii: F(s) / np.sqrt(F(ii) * tss)
    assert np.array_equal(O.match_template(IMG, TPL, O.SSE), _f([[14, 14], [3, 1]]))
    assert np.array_equal(O.match_template(IMG, TPL, O.SSE_NORMED),
                          _f([[nz(14, 22), nz(14, 30)], [nz(3, 23), nz(1, 35)]]))
    assert np.array_equal(O.match_template(IMG, TPL, O.CCORR), _f([[19, 23], [25, 32]]))
    assert np.array_equal(O.match_template(IMG, TPL, O.CCORR_NORMED),
                          _f([[nz(19, 22), nz(23, 30)], [nz(25, 23), nz(32, 35)]]))


def test_match_template_edge_kats():
    # template of the image's size, :538-548; zero norm, :550-560
    z = np.zeros((5, 5), np.uint8)
    assert np.array_equal(O.match_template(z, z, O.SSE), _f([[0.0]]))
    one = np.zeros((1, 1), np.uint8)
    assert np.array_equal(O.match_template(one, one, O.SSE_NORMED), _f([[0.0]]))


def test_match_template_with_mask_kats():
    # template_matching.rs:659-816
    exp = {
        O.SSE: _f([[89., 25.], [10., 1.]]),
        O.SSE_NORMED: _f([[1.0246822, 0.19536021], [0.067865655, 0.005362412]]),
        O.CCORR: _f([[68., 124.], [146., 186.]]),
        O.CCORR_NORMED: _f([[0.78290325, 0.96898663], [0.9908386, 0.9974086]]),
    }
    for m, e in exp.items():
        assert np.array_equal(O.match_template_with_mask(IMG, TPL, MASK, m), e), m


def _exact_sums(img, tpl):
    win = sliding_window_view(img.astype(np.int64), tpl.shape)
    t = tpl.astype(np.int64)
    c = np.einsum("yxij,ij->yx", win, t)
    ii = np.einsum("yxij,yxij->yx", win, win)
    sse = np.einsum("yxij->yx", (win - t) ** 2)
    return c, ii, sse, int((t * t).sum())


@pytest.mark.parametrize("shape,tshape", [((20, 23), (4, 5)), ((33, 40), (16, 16)),
                                          ((50, 45), (32, 32))])
def test_match_template_f64_is_exact_and_f32_is_close(rng, shape, tshape):
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    tpl = rng.integers(0, 256, tshape, dtype=np.uint8)
    c, ii, sse, tss = _exact_sums(img, tpl)
    norm = np.sqrt(ii.astype(np.float64) * tss)
    exp = {O.SSE: sse.astype(np.float64), O.CCORR: c.astype(np.float64),
           O.SSE_NORMED: sse / norm, O.CCORR_NORMED: c / norm}
    for m, e in exp.items():
        got64 = O.match_template_f64(img, tpl, m)
        assert np.array_equal(got64, e), m
        got32 = O.match_template(img, tpl, m)
        # a sequential f32 sum of n non-negative terms errs by at most ~n ulp
        tol = (tshape[0] * tshape[1] + 4) * 2.0 ** -24
        assert np.max(np.abs(got32 - e) / np.maximum(np.abs(e), 1e-30)) < tol, m
        if tshape[0] * tshape[1] * 65025 < 2 ** 24:
            # every partial sum is an exact f32 integer: the f32 sums are exact
            if m in (O.SSE, O.CCORR):
                assert np.array_equal(got32, e.astype(F)), m


def test_match_template_f32_images_against_float64(rng):
    img = rng.standard_normal((30, 28)).astype(F)
    tpl = rng.standard_normal((6, 5)).astype(F)
    win = sliding_window_view(img.astype(np.float64), tpl.shape)
    t = tpl.astype(np.float64)
    c = np.einsum("yxij,ij->yx", win, t)
    ii = np.einsum("yxij,yxij->yx", win, win)
    sse = np.einsum("yxij->yx", (win - t) ** 2)
    norm = np.sqrt(ii * (t * t).sum())
    exp = {O.SSE: sse, O.CCORR: c, O.SSE_NORMED: sse / norm, O.CCORR_NORMED: c / norm}
    for m, e in exp.items():
        assert np.allclose(O.match_template_f64(img, tpl, m), e, rtol=1e-12, atol=1e-12), m
        assert np.allclose(O.match_template(img, tpl, m), e, rtol=1e-4, atol=1e-4), m


def test_match_template_parallel_equals_serial(rng):
    img = rng.integers(0, 256, (40, 41), dtype=np.uint8)
    tpl = rng.integers(0, 256, (7, 9), dtype=np.uint8)
    for m in range(4):
        assert np.array_equal(O.match_template(img, tpl, m), O.match_template(img, tpl, m, nthreads=4))


# ---- find_extremes ---------------------------------------------------------------------

def test_find_extremes_kat():
    # template_matching.rs:818-833
    img = gray_image([[10, 7, 8, 1], [9, 15, 4, 2]])
    mx, mn, mxl, mnl = O.find_extremes(img)
    assert (mx, mn, mxl, mnl) == (15, 1, (1, 1), (3, 0))


def test_find_extremes_first_occurrence(rng):
    for dtype in (np.uint8, np.float32):
        img = rng.integers(0, 5, (37, 29)).astype(dtype)
        mx, mn, mxl, mnl = O.find_extremes(img)
        fi, fa = int(np.argmin(img)), int(np.argmax(img))  # numpy: first occurrence
        assert mnl == (fi % 29, fi // 29) and mxl == (fa % 29, fa // 29)
        assert mn == img.min() and mx == img.max()


# ---- integral image ----------------------------------------------------------------------

def test_integral_image_against_cumsum(rng):
    img = rng.integers(0, 256, (13, 17), dtype=np.uint8)
    for sq in (False, True):
        v = img.astype(np.uint64) ** (2 if sq else 1)
        exp = np.zeros((14, 18), np.uint64)
        exp[1:, 1:] = v.cumsum(0).cumsum(1)
        assert np.array_equal(O.integral_image(img, sq), (exp % (1 << 32)).astype(np.uint32))
    # integral_image.rs:25-48 doc example
    img = gray_image([[1, 2, 3], [4, 5, 6]])
    assert np.array_equal(O.integral_image(img),
                          np.array([[0, 0, 0, 0], [0, 1, 3, 6], [0, 5, 12, 21]], np.uint32))
    # integral_image.rs:66-90 doc example
    assert np.array_equal(O.integral_image(img, True),
                          np.array([[0, 0, 0, 0], [0, 1, 5, 14], [0, 17, 46, 91]], np.uint32))


# ---- adaptive threshold (contrast.rs:19-47) ----------------------------------------------

def _adaptive_brute(img, r, delta):
    """Block means by direct summation over the clipped window (no integral
    image), then the reference's comparison."""
    h, w = img.shape
    out = np.zeros_like(img)
    for y in range(h):
        for x in range(w):
            blk = img[max(0, y - r):min(h - 1, y + r) + 1, max(0, x - r):min(w - 1, x + r) + 1]
            mean = int(blk.astype(np.int64).sum()) // blk.size
            out[y, x] = 255 if int(img[y, x]) >= mean - delta else 0
    return out


def test_adaptive_threshold_reference_unit_tests():
    # contrast.rs:471-477 adaptive_threshold_constant
    assert (O.adaptive_threshold(np.full((3, 3), 100, np.uint8), 1, 0) == 255).all()
    # contrast.rs:479-492 one darker pixel
    for y in range(3):
        for x in range(3):
            img = np.full((3, 3), 200, np.uint8)
            img[y, x] = 100
            exp = np.full((3, 3), 255, np.uint8)
            exp[y, x] = 0
            assert np.array_equal(O.adaptive_threshold(img, 1, 0), exp)
    # contrast.rs:494-523 one lighter pixel
    for y in range(5):
        for x in range(5):
            img = np.full((5, 5), 100, np.uint8)
            img[y, x] = 200
            got = O.adaptive_threshold(img, 1, 0)
            for yb in range(5):
                for xb in range(5):
                    if (yb, xb) == (y, x):
                        assert got[yb, xb] == 255
                    elif abs(yb - y) <= 1 and abs(xb - x) <= 1:
                        assert got[yb, xb] == 0
                    else:
                        assert got[yb, xb] == 255
    # contrast.rs:525-540 delta
    img = np.full((3, 3), 100, np.uint8)
    img[2, 2] = 0
    assert (O.adaptive_threshold(img, 1, 100) == 255).all()
    exp = np.full((3, 3), 255, np.uint8)
    exp[2, 2] = 0
    assert np.array_equal(O.adaptive_threshold(img, 1, 50), exp)
    with pytest.raises(ValueError, match="block_radius > 0"):
        O.adaptive_threshold(img, 0, 0)


def test_adaptive_threshold_brute_force_and_truth(rng, fixtures):
    for shape, r, delta in [((9, 14), 2, 0), ((17, 6), 4, -7), ((12, 12), 30, 3), ((1, 9), 1, 0)]:
        img = rng.integers(0, 256, shape, dtype=np.uint8)
        assert np.array_equal(O.adaptive_threshold(img, r, delta), _adaptive_brute(img, r, delta))
    # tests/regression.rs:576-582
    assert np.array_equal(O.adaptive_threshold(fixtures["zebra"], 41, 0),
                          fixtures["zebra_adaptive_threshold"])


# ---- canny (edges.rs:30-155) -------------------------------------------------------------

def test_canny_direction_and_magnitude_over_the_sobel_domain():
    """Every (gx, gy) a Sobel response of a u8 image can take: the oracle's
    f32 atan2f binning equals the exact angle's bin (decided in integers:
    tan 22.5 = sqrt2 - 1, tan 67.5 = sqrt2 + 1) except at (-985, -408), and
    hypotf equals the correctly rounded sqrt of the exact gx^2 + gy^2."""
    r = 1020
    cls, mag = O.canny_direction_grid(r)
    gy, gx = np.meshgrid(np.arange(-r, r + 1, dtype=np.int64), np.arange(-r, r + 1, dtype=np.int64),
                         indexing="ij")
    x = np.where(gy < 0, -gx, gx)
    y = np.abs(gy)
    ax = np.abs(x)
    above_t1 = (ax + y) ** 2 > 2 * ax * ax
    below_t2 = (y - ax <= 0) | ((y - ax) ** 2 < 2 * ax * ax)
    exact = np.zeros(cls.shape, np.uint8)
    pos = y > 0
    exact[pos & (x > 0) & above_t1 & below_t2] = 45
    exact[pos & ((x == 0) | ((x > 0) & ~below_t2) | ((x < 0) & ~below_t2))] = 90
    exact[pos & (x < 0) & above_t1 & below_t2] = 135
    diff = np.argwhere(cls != exact)
    assert [(int(gx[i, j]), int(gy[i, j])) for i, j in diff] == [(-985, -408)]
    assert cls[-408 + r, -985 + r] == 45 and exact[-408 + r, -985 + r] == 0
    n = gx * gx + gy * gy
    assert np.array_equal(mag, np.sqrt(n.astype(np.float64)).astype(np.float32))
    assert np.array_equal(mag, np.sqrt(n.astype(np.float32)))
    # sqrtf is strictly increasing on 0 ..= 2 * 1020^2, so magnitudes compare
    # like their integer squares
    roots = np.sqrt(np.arange(0, 2 * 1020 * 1020 + 1, dtype=np.float32))
    assert (np.diff(roots) > 0).all()
    # spot values of the bins
    assert O.canny_direction(1, 0) == 0 and O.canny_direction(-1, 0) == 0
    assert O.canny_direction(0, 1) == 90 and O.canny_direction(0, -1) == 90
    assert O.canny_direction(1, 1) == 45 and O.canny_direction(-1, 1) == 135
    assert O.canny_direction(-1, -1) == 45 and O.canny_direction(1, -1) == 135


def _hysteresis_brute(thin, low, high):
    """Fixed point of the marking rule, by repeated sweeps (no stack)."""
    h, w = thin.shape
    out = np.zeros((h, w), bool)
    out[1:h - 1, 1:w - 1] = thin[1:h - 1, 1:w - 1] >= high
    nbrs = [(1, 0), (1, 1), (0, 1), (-1, -1), (-1, 0), (-1, 1)]
    changed = True
    while changed:
        changed = False
        for yy, xx in zip(*np.nonzero(out)):
            for dx, dy in nbrs:
                nx, ny = xx + dx, yy + dy
                if 0 <= nx < w and 0 <= ny < h and not out[ny, nx] and thin[ny, nx] >= low:
                    out[ny, nx] = True
                    changed = True
    return out


def test_canny_truth_and_hysteresis(fixtures, rng):
    # tests/regression.rs:555-560
    assert np.array_equal(O.canny(fixtures["zebra"], 250.0, 300.0), fixtures["zebra_canny"])
    # the oracle's depth-first search reaches the same set as a sweep to the
    # fixed point, on images whose thinned maps come from the oracle's own
    # stages
    from scipy.ndimage import gaussian_filter
    for seed in range(3):
        img = np.clip(gaussian_filter(rng.normal(128, 60, (40, 50)), 2.0), 0, 255).astype(np.uint8)
        got = O.canny(img, 20.0, 60.0)
        assert set(np.unique(got)) <= {0, 255}
        assert not got[0].any() and not got[-1].any() and not got[:, 0].any() and not got[:, -1].any()
        # raising low can only remove edges, lowering high can only add them
        assert not (O.canny(img, 40.0, 60.0) & ~got).any()
        assert not (got & ~O.canny(img, 20.0, 40.0)).any()
        thin = O.canny_thinned(img)
        assert np.array_equal(O.canny_hysteresis(thin, 20.0, 60.0) == 255,
                              _hysteresis_brute(thin, 20.0, 60.0))
        assert np.array_equal(O.canny_hysteresis(thin, 20.0, 60.0), got)
    # synthetic thinned maps ([y, x] indexing): marks travel only along the
    # six neighbours (x+1, y), (x+1, y+1), (x, y+1), (x-1, y-1), (x-1, y),
    # (x-1, y+1); there is no step up or up-right
    thin = np.zeros((7, 7), np.float32)
    thin[5, 1] = 10.0  # seed (1, 5)
    thin[5, 2] = 5.0  # (2, 5): right of the seed
    thin[4, 1] = 5.0  # (1, 4): up-left of (2, 5)
    thin[4, 3] = 5.0  # (3, 4): up-right of (2, 5) only
    got = O.canny_hysteresis(thin, 1.0, 8.0)
    assert got[5, 1] == 255 and got[5, 2] == 255 and got[4, 1] == 255 and got[4, 3] == 0
    assert np.array_equal(got == 255, _hysteresis_brute(thin, 1.0, 8.0))
    thin2 = np.zeros((7, 7), np.float32)
    thin2[5, 1] = 10.0
    thin2[4, 1] = 5.0  # straight up from the seed
    assert O.canny_hysteresis(thin2, 1.0, 8.0)[4, 1] == 0
    with pytest.raises(ValueError, match="high_threshold >= low_threshold"):
        O.canny(np.zeros((5, 5), np.uint8), 2.0, 1.0)
    img = np.zeros((9, 9), np.uint8)
    img[3:6, 3:6] = 255
    with pytest.raises(IndexError):
        O.canny(img, 0.0, 10.0)
    assert not O.canny(img, -1.0, 5000.0).any()
