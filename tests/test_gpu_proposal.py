"""The trained proposal network (imageproc_amd/proposal_weights.npz): its
search agrees with the full-map search (match_template + find_extremes,
template_matching.rs:78-92, 479-511) on held-out pairs, and ranks the true
place higher than the coarse scores alone do."""
import numpy as np
import pytest
import torch

from imageproc_amd.proposal import ProposalNet, ProposalSearch
from imageproc_amd.proposal_train import CCN, pair_stream

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pairs():
    stream = pair_stream(777001, 2048, per_image=4)
    ps = [next(stream) for _ in range(96)]
    judge_search = ProposalSearch(None, CCN)
    judge = [judge_search.full_map(img, tpl) for img, tpl, _ in ps]
    return ps, judge


def test_proposal_search_agrees_with_full_map(pairs):
    ps, judge = pairs
    srch = ProposalSearch(ProposalNet.load(), CCN)
    agree = 0
    for (img, tpl, _), (gx, gy, gv) in zip(ps, judge):
        x, y, v, _ = srch.search(img, tpl)
        if (x, y) == (gx, gy):
            agree += 1
            assert np.float32(v) == np.float32(gv)  # the exact window score is the map's value
    assert agree >= 0.99 * len(ps)


def test_network_ranks_true_place_higher_than_coarse_scores(pairs):
    """With one exact window (n_top = 1), the network's first choice is the
    full map's argmax more often than the best coarse tile is."""
    ps, judge = pairs
    net = ProposalNet.load()
    s1 = ProposalSearch(net, CCN, n_top=1, margin=-1e30, min_prob=-1.0)
    with_net = sum(s1.search(img, tpl)[:2] == j[:2] for (img, tpl, _), j in zip(ps, judge))
    coarse = sum(s1.search(img, tpl, use_net=False)[:2] == j[:2] for (img, tpl, _), j in zip(ps, judge))
    assert with_net >= coarse + 0.05 * len(ps), (with_net, coarse)


def test_fallback_returns_full_map_answer(pairs):
    ps, judge = pairs
    srch = ProposalSearch(ProposalNet.load(), CCN, margin=1e30)  # always unsure
    for (img, tpl, _), j in list(zip(ps, judge))[:4]:
        x, y, v, fell_back = srch.search(img, tpl)
        assert fell_back and (x, y, v) == j


def test_search_agreement_on_held_out_8192_pairs():
    """8192x8192 pairs from a seed used neither in training nor in the margin
    calibration: the search with the calibrated fallback agrees with the full
    map on at least 99.5 % of them, and the network's ranking alone agrees at
    least as often as ranking by coarse score alone."""
    stream = pair_stream(555001, 8192, per_image=4)
    srch = ProposalSearch(ProposalNet.load(), CCN)
    raw = ProposalSearch(ProposalNet.load(), CCN, margin=-1e30, min_prob=-1.0)
    n, agree, net_only, coarse_only = 240, 0, 0, 0
    for _ in range(n):
        img, tpl, _ = next(stream)
        j = srch.full_map(img, tpl)[:2]
        agree += srch.search(img, tpl)[:2] == j
        net_only += raw.search(img, tpl)[:2] == j
        coarse_only += raw.search(img, tpl, use_net=False)[:2] == j
    assert agree >= 0.995 * n, (agree, n)
    assert net_only >= coarse_only, (net_only, coarse_only)


def test_low_margin_pair_takes_the_full_map():
    """A template that occurs twice in the image (an unperturbed crop pasted
    at a second place far away) leaves no margin between the best window and
    the runner-up: the calibrated search falls back to the full map and
    returns its answer, the first of the two places in row-major order."""
    img = next(pair_stream(555002, 2048, per_image=1))[0].clone()
    tpl = img[700:748, 900:948].clone().contiguous()
    img[1500:1548, 300:348] = tpl
    srch = ProposalSearch(ProposalNet.load(), CCN)
    x, y, v, fell_back = srch.search(img, tpl)
    assert fell_back
    assert (x, y, v) == srch.full_map(img, tpl)
    assert (x, y) == (900, 700)
