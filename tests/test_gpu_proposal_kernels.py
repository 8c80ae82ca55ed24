"""The proposal search's selection kernels against direct numpy statements of
what they select, its exact windows against the full score map
(template_matching.rs:78-92), and the one-call search against the same search
made from the step-wise entry points."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
from imageproc_amd import device as dev
from imageproc_amd._lib import check, lib
from imageproc_amd.proposal import PIX, ProposalNet, ProposalSearch
from imageproc_amd.proposal_train import CCN, pair_stream
from imageproc_amd.template_matching import MatchTemplateMethod as M

pytestmark = pytest.mark.gpu


def p(t):
    return C.c_void_p(t.data_ptr())


def order_key(v, i):
    """Candidate order: larger value first, NaN after numbers, then lower index."""
    return (1, 0.0, i) if np.isnan(v) else (0, -float(v), i)


def candidates_reference(m, k):
    ch, cw = m.shape
    tiles = []
    for ty in range(0, ch, 16):
        for tx in range(0, cw, 16):
            blk = m[ty:ty + 16, tx:tx + 16].astype(np.float64)
            ys, xs = np.mgrid[ty:ty + blk.shape[0], tx:tx + blk.shape[1]]
            idx, v = (ys * cw + xs).ravel(), blk.ravel()
            nan = np.isnan(v)
            o = np.lexsort((idx, np.where(nan, 0.0, -v), nan))[0]
            tiles.append((int(nan[o]), 0.0 if nan[o] else -float(v[o]), int(idx[o])))
    tiles.sort()
    out = [((t[2] % cw, t[2] // cw), np.nan if t[0] else -t[1]) for t in tiles[:k]]
    out += [((0, 0), -np.inf)] * (k - len(out))
    return out


@pytest.mark.parametrize("shape", [(37, 53), (600, 700), (2100, 2077)])
def test_candidates_match_sorted_tile_maxima(shape):
    rng = np.random.default_rng(shape[0])
    m = np.round(rng.random(shape), 2).astype(np.float32)  # many equal values
    m[5:40, 7:30] = np.nan
    k = 64
    cm = torch.from_numpy(m).cuda()
    cand = torch.empty((k, 2), dtype=torch.int32, device="cuda")
    cs = torch.empty(k, dtype=torch.float32, device="cuda")
    ws = torch.empty(lib().iph_dev_proposal_candidates_workspace(shape[1], shape[0]),
                     dtype=torch.uint8, device="cuda")
    check(lib().iph_dev_proposal_candidates(p(cm), shape[1], shape[0], k, p(cand), p(cs), p(ws),
                                            ws.numel(), None))
    got_c, got_s = cand.cpu().numpy(), cs.cpu().numpy()
    want = candidates_reference(m, k)
    for i, ((x, y), s) in enumerate(want):
        assert tuple(got_c[i]) == (x, y), i
        assert (np.isnan(s) and np.isnan(got_s[i])) or got_s[i] == s, i


def top_reference(lg, cand, cw, ch, n_top):
    k = lg.shape[0]
    ent = []
    for c in range(k):
        for pix in range(PIX):
            x, y = cand[c, 0] - 8 + pix % 16, cand[c, 1] - 8 + pix // 16
            v = lg[c, pix]
            if 0 <= x < cw and 0 <= y < ch and not np.isnan(v):
                ent.append((order_key(v, c * PIX + pix), x, y, v))
    ent.sort(key=lambda e: e[0])
    taken, out = set(), []
    for _, x, y, v in ent:
        if (x, y) in taken:
            continue
        taken.add((x, y))
        out.append((x, y, v))
        if len(out) == n_top:
            break
    fin = lg[~np.isnan(lg)].astype(np.float64)
    mx = fin.max()
    se = np.exp(lg.astype(np.float64) - mx).sum()
    res = [((x, y), np.exp(float(v) - mx) / se) for x, y, v in out]
    return res + [((0, 0), 0.0)] * (n_top - len(res))


# (256, 64, 3, 20): 256 patches over a 3-column map hold fewer places than
# top slots, so every list is short and the answer ends in empty slots
@pytest.mark.parametrize("k,n_top,cw,ch", [(64, 16, 200, 150), (8, 64, 40, 30), (16, 64, 5, 6),
                                           (256, 64, 3, 20)])
def test_top_places_match_sequential_choice(k, n_top, cw, ch):
    rng = np.random.default_rng(k * 1000 + n_top)
    # candidates close together (overlapping patches) and at the map's edges
    cand = np.stack([rng.integers(-3, cw + 3, k), rng.integers(-3, ch + 3, k)], 1).astype(np.int32)
    cand[: k // 2] = cand[0] + rng.integers(-6, 7, (k // 2, 2))
    lg = np.round(rng.standard_normal((k, PIX)) * 2, 1).astype(np.float32)  # ties
    if k == 8:
        lg[1, :20] = np.nan  # never chosen; the probabilities turn NaN
    z = np.zeros((k, PIX, 32), np.float32)
    z[:, :, 0] = lg
    zt, ct = torch.from_numpy(z).cuda(), torch.from_numpy(cand).cuda()
    centers = torch.empty((n_top, 2), dtype=torch.int32, device="cuda")
    probs = torch.empty(n_top, dtype=torch.float32, device="cuda")
    ws = torch.empty(lib().iph_dev_proposal_top_workspace(k, n_top), dtype=torch.uint8,
                     device="cuda")
    check(lib().iph_dev_proposal_top(p(zt), k, p(ct), cw, ch, n_top, p(centers), p(probs), p(ws),
                                     ws.numel(), None))
    gc, gp = centers.cpu().numpy(), probs.cpu().numpy()
    for r, ((x, y), pr) in enumerate(top_reference(lg, cand, cw, ch, n_top)):
        assert tuple(gc[r]) == (x, y), r
        assert abs(gp[r] - pr) <= 1e-5 * pr + 1e-30 or (np.isnan(gp[r]) and np.isnan(pr)), r


@pytest.mark.parametrize("method", list(M))
@pytest.mark.parametrize("tsize", [(48, 48), (37, 29), (7, 128), (128, 5)])
def test_refine_windows_equal_full_map(method, tsize):
    th, tw = tsize
    rng = np.random.default_rng(th * 131 + tw)
    img_np = O.gaussian_blur_f32(rng.integers(0, 256, (300, 420), dtype=np.uint8), 2.0)
    img = torch.from_numpy(img_np).cuda()
    tpl = torch.from_numpy(rng.integers(0, 256, (th, tw), dtype=np.uint8)).cuda()
    # the exact-integer map (the windows sum exactly, as the MFMA path does)
    full = dev.match_template(img, tpl, method, "mfma").cpu().numpy()
    oh, ow = full.shape
    # centres (scaled by 4) whose windows reach past every border
    cen = np.array([[0, 0], [ow // 4, oh // 4], [(ow - 1) // 4 + 2, (oh - 1) // 4 + 1],
                    [1, (oh - 1) // 4]], np.int32)
    ct = torch.from_numpy(cen).cuda()
    ws = torch.empty(lib().iph_dev_proposal_refine_workspace(len(cen)), dtype=torch.uint8,
                     device="cuda")
    res = torch.empty(4, dtype=torch.float32, device="cuda")
    check(lib().iph_dev_proposal_refine(p(img), 420, 300, p(tpl), tw, th, int(method), p(ct),
                                        len(cen), 4, 4, p(res), p(ws), ws.numel(), None))
    scores = ws[: len(cen) * 289 * 4].view(torch.float32).cpu().numpy().reshape(len(cen), 17, 17)
    places = []
    for c, (cx, cy) in enumerate(cen):
        for dy in range(17):
            for dx in range(17):
                x, y = 4 * cx - 8 + dx, 4 * cy - 8 + dy
                s = scores[c, dy, dx]
                if 0 <= x < ow and 0 <= y < oh:
                    assert s == full[y, x] or (np.isnan(s) and np.isnan(full[y, x])), (c, dx, dy)
                    if not np.isnan(s):
                        places.append((-float(s), y, x))
                else:
                    assert np.isnan(s)
    places.sort()
    r = res.cpu().numpy()
    best, by, bx = -places[0][0], places[0][1], places[0][2]
    assert (int(r[0]), int(r[1]), np.float32(r[2])) == (bx, by, np.float32(best))
    far = [-s for s, y, x in places if abs(x - bx) > 4 or abs(y - by) > 4]
    assert np.float32(r[3]) == np.float32(max(far))


def test_downsample_vector_and_scalar_rows():
    rng = np.random.default_rng(9)
    for h, w in [(64, 512), (37, 83), (16, 4096 + 16)]:
        a = rng.integers(0, 256, (h, w), dtype=np.uint8)
        src = torch.from_numpy(a).cuda()
        dst = torch.empty((h // 4, w // 4), dtype=torch.uint8, device="cuda")
        check(lib().iph_dev_downsample4_u8(p(src), w, h, p(dst), None))
        dh, dw = h // 4, w // 4
        want = (a[: 4 * dh, : 4 * dw].reshape(dh, 4, dw, 4).astype(np.int64).sum(axis=(1, 3)) + 8) >> 4
        assert np.array_equal(dst.cpu().numpy(), want.astype(np.uint8)), (h, w)


@pytest.mark.parametrize("side", [1024, 2048])
def test_one_call_search_equals_step_search(side):
    net = ProposalNet.load()
    srch = ProposalSearch(net, CCN)
    plain = ProposalSearch(None, M.CrossCorrelation)
    stream = pair_stream(4242 + side, side, per_image=3)
    for _ in range(6):
        img, tpl, _ = next(stream)
        for use_net in (True, False):
            assert srch.search_raw(img, tpl, use_net) == srch.search_steps(img, tpl, use_net)
        assert plain.search_raw(img, tpl, False) == plain.search_steps(img, tpl, False)


@pytest.mark.parametrize("tsize", [(8, 8), (20, 24), (60, 17), (67, 67), (96, 96), (128, 100)])
def test_one_call_search_equals_step_search_across_template_sizes(tsize):
    """Small templates take the one-pass coarse kernel (1 to 4 template words
    per downsampled row), large ones the int8 MFMA map; both must equal the
    step-wise search value for value."""
    th, tw = tsize
    rng = np.random.default_rng(th * 1000 + tw)
    img_np = O.gaussian_blur_f32(rng.integers(0, 256, (700, 1030), dtype=np.uint8), 3.0)
    img = torch.from_numpy(img_np).cuda()
    net = ProposalNet.load()
    for method in (M.CrossCorrelationNormalized, M.CrossCorrelation):
        srch = ProposalSearch(net, method)
        for _ in range(3):
            y, x = int(rng.integers(0, 700 - th)), int(rng.integers(0, 1030 - tw))
            tpl = img[y:y + th, x:x + tw].contiguous()
            for use_net in (True, False):
                assert srch.search_raw(img, tpl, use_net) == srch.search_steps(img, tpl, use_net)
