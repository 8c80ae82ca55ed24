"""match_template_with_mask (template_matching.rs:125-163, masked methods
:295-376) on the GPU: the reference's masked known-answer tests (:659-816),
and bit-identity with the oracle's restatement on random inputs."""
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
        assert np.array_equal(match_template_with_mask(IMG, TPL, m, MASK), e), m
        assert np.array_equal(match_template_with_mask_parallel(IMG, TPL, m, MASK), e), m


@pytest.mark.parametrize("shape,tshape", [((40, 57), (7, 5)), ((120, 97), (16, 16)),
                                          ((75, 300), (33, 9)), ((64, 64), (64, 64))])
def test_masked_random_bit_identical(shape, tshape):
    rng = np.random.default_rng(shape[0] * 31 + tshape[1])
    img = rng.integers(0, 256, shape, dtype=np.uint8)
    tpl = rng.integers(0, 256, tshape, dtype=np.uint8)
    mask = rng.integers(0, 256, tshape, dtype=np.uint8)
    mask[rng.random(tshape) < 0.3] = 0
    for m in M:
        got = match_template_with_mask(img, tpl, m, mask)
        want = O.match_template_with_mask(img, tpl, mask, int(m))
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), m


def test_masked_zero_mask_and_device_entry():
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (50, 60), dtype=np.uint8)
    tpl = rng.integers(0, 256, (9, 11), dtype=np.uint8)
    zero = np.zeros((9, 11), np.uint8)
    for m in M:  # zero norm: the raw score (0) is returned, :330 / :366
        assert np.array_equal(match_template_with_mask(img, tpl, m, zero),
                              O.match_template_with_mask(img, tpl, zero, int(m)))
    mask = rng.integers(0, 256, (9, 11), dtype=np.uint8)
    gi, gt, gm = (torch.from_numpy(a).cuda() for a in (img, tpl, mask))
    for m in M:
        out = dev.match_template_with_mask(gi, gt, m, gm).cpu().numpy()
        assert np.array_equal(out, O.match_template_with_mask(img, tpl, mask, int(m))), m


def test_masked_panics():
    with pytest.raises(Panic, match="the template and mask must be the same size"):
        match_template_with_mask(IMG, TPL, M.CrossCorrelation, gray_image([[1, 2, 3]]))
    with pytest.raises(Panic, match="image width must be greater than or equal"):
        match_template_with_mask(TPL, IMG, M.CrossCorrelation, IMG)
