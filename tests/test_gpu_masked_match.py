"""match_template_with_mask (template_matching.rs:125-163, masked methods
:295-376) on the GPU: the reference's masked known-answer tests (:659-816),
bit-identity of the direct path with the oracle's restatement on random
inputs, and the MFMA path (CrossCorrelation and CrossCorrelationNormalized as
BF16x3 correlations) against exact integer sums in float64."""
import numpy as np
import pytest
import torch

import oracle as O
from conftest import gray_image
from imageproc_amd import device as dev
from imageproc_amd._lib import Panic
from imageproc_amd.template_matching import MatchTemplateMethod as M
from imageproc_amd.template_matching import (match_template_with_mask,
                                             match_template_with_mask_parallel)

pytestmark = pytest.mark.gpu

IMG = gray_image([[1, 4, 2], [2, 1, 3], [3, 3, 4]])
TPL = gray_image([[1, 2], [3, 4]])
MASK = gray_image([[0, 1], [2, 3]])


def _f(rows):
    return np.array(rows, dtype=np.float32)


def test_masked_kats():
    # template_matching.rs:659-816, f32 constants compared bit for bit
    exp = {
        M.SumOfSquaredErrors: _f([[89., 25.], [10., 1.]]),
        M.SumOfSquaredErrorsNormalized: _f([[1.0246822, 0.19536021], [0.067865655, 0.005362412]]),
        M.CrossCorrelation: _f([[68., 124.], [146., 186.]]),
        M.CrossCorrelationNormalized: _f([[0.78290325, 0.96898663], [0.9908386, 0.9974086]]),
    }
    for m, e in exp.items():
        assert np.array_equal(match_template_with_mask(IMG, TPL, m, MASK, "direct"), e), m
        assert np.array_equal(match_template_with_mask_parallel(IMG, TPL, m, MASK), e), m
        # the default (MFMA for the cross-correlations): these sums are small
        # integers, exact on either path
        assert np.array_equal(match_template_with_mask(IMG, TPL, m, MASK), e), m


@pytest.mark.parametrize("shape,tshape", [((40, 57), (7, 5)), ((120, 97), (16, 16)),
                                          ((75, 300), (33, 9)), ((64, 64), (64, 64))])
def test_masked_random_bit_identical(shape, tshape):
    rng = np.random.default_rng(shape[0] * 31 + tshape[1])
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    tpl = rng.integers(0, 256, tshape, dtype=np.uint8)
    mask = rng.integers(0, 256, tshape, dtype=np.uint8)
    mask[rng.random(tshape) < 0.3] = 0
    for m in M:
        got = match_template_with_mask(img, tpl, m, mask, "direct")
        want = O.match_template_with_mask(img, tpl, mask, int(m))
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), m


def test_masked_zero_mask_and_device_entry():
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (50, 60), dtype=np.uint8)
    tpl = rng.integers(0, 256, (9, 11), dtype=np.uint8)
    zero = np.zeros((9, 11), np.uint8)
    for m in M:  # zero norm: the raw score (0) is returned, :330 / :366
        for path in ("auto", "direct"):
            assert np.array_equal(match_template_with_mask(img, tpl, m, zero, path),
                                  O.match_template_with_mask(img, tpl, zero, int(m)))
    mask = rng.integers(0, 256, (9, 11), dtype=np.uint8)
    gi, gt, gm = (torch.from_numpy(a).cuda() for a in (img, tpl, mask))
    for m in M:
        out = dev.match_template_with_mask(gi, gt, m, gm, path="direct").cpu().numpy()
        assert np.array_equal(out, O.match_template_with_mask(img, tpl, mask, int(m))), m
    for m in (M.CrossCorrelation, M.CrossCorrelationNormalized):
        out = dev.match_template_with_mask(gi, gt, m, gm, path="mfma").cpu().numpy()
        host = match_template_with_mask(img, tpl, m, mask, "mfma")
        assert np.array_equal(out, host), m


def test_masked_panics():
    with pytest.raises(Panic, match="the template and mask must be the same size"):
        match_template_with_mask(IMG, TPL, M.CrossCorrelation, gray_image([[1, 2, 3]]))
    with pytest.raises(Panic, match="image width must be greater than or equal"):
        match_template_with_mask(TPL, IMG, M.CrossCorrelation, IMG)


def _masked_f64(img, tpl, mask):
    """Exact sum t i m^2, sum (i m)^2 and mult_square_sum per window (int64),
    and the normalised score in float64."""
    th, tw = tpl.shape
    win = np.lib.stride_tricks.sliding_window_view(img.astype(np.int64), (th, tw))
    t = tpl.astype(np.int64)
    m2 = mask.astype(np.int64) ** 2
    cc = np.einsum("yxij,ij->yx", win, t * m2)
    imim = np.einsum("yxij,ij->yx", win * win, m2)
    tms = float(((t * mask.astype(np.int64)) ** 2).sum())
    norm = np.sqrt(tms * imim.astype(np.float64))
    ccn = np.where(norm > 0, cc / np.where(norm > 0, norm, 1.0), cc)
    return cc.astype(np.float64), ccn


@pytest.mark.parametrize("shape,tshape", [((90, 131), (7, 5)), ((200, 170), (32, 32)),
                                          ((150, 260), (33, 65)), ((140, 140), (128, 100))])
def test_masked_mfma_against_exact_sums(shape, tshape):
    """MFMA path: every output within the BF16x3 bound of the exact sums (the
    template T*M^2 keeps 16 of its 24 bits: relative 2^-16, plus the f32
    accumulation), and the best window equal to the direct path's."""
    rng = np.random.default_rng(shape[1] * 7 + tshape[0])
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    th, tw = tshape
    y0, x0 = int(rng.integers(0, shape[0] - th + 1)), int(rng.integers(0, shape[1] - tw + 1))
    tpl = img[y0:y0 + th, x0:x0 + tw].copy()
    mask = rng.integers(0, 256, tshape, dtype=np.uint8)
    mask[rng.random(tshape) < 0.3] = 0
    cc, ccn = _masked_f64(img, tpl, mask)
    got = match_template_with_mask(img, tpl, M.CrossCorrelation, mask, "mfma").astype(np.float64)
    assert np.max(np.abs(got - cc) / np.maximum(cc, 1.0)) < 1e-4
    got = match_template_with_mask(img, tpl, M.CrossCorrelationNormalized, mask,
                                   "mfma").astype(np.float64)
    assert np.max(np.abs(got - ccn)) < 1e-4
    direct = match_template_with_mask(img, tpl, M.CrossCorrelationNormalized, mask, "direct")
    assert np.unravel_index(np.argmax(got), got.shape) == \
        np.unravel_index(np.argmax(direct), direct.shape) == (y0, x0)


def test_masked_mfma_refuses_sse_and_large_templates():
    img = np.zeros((300, 300), np.uint8)
    with pytest.raises(Exception, match="MFMA path of match_template_with_mask takes"):
        match_template_with_mask(img, img[:8, :8], M.SumOfSquaredErrors, img[:8, :8], "mfma")
    with pytest.raises(Exception, match="MFMA path of match_template_with_mask takes"):
        match_template_with_mask(img, img[:129, :8], M.CrossCorrelation, img[:129, :8], "mfma")
    # the default path takes both on the direct kernel
    out = match_template_with_mask(img, img[:129, :8], M.CrossCorrelation, img[:129, :8])
    assert out.shape == (172, 293) and not out.any()
