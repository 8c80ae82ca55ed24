"""The proposal network's kernels against the float64 oracle (oracle/cnn.py),
on inputs that are exact in bf16, and the search pipeline's kernels against
direct computations and the full score map (template_matching.rs:78-92,
479-511)."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
from oracle import cnn
from imageproc_amd import device as dev
from imageproc_amd._lib import check, lib
from imageproc_amd.proposal import CH, KK, PIX, ProposalNet, ProposalSearch
from imageproc_amd.template_matching import MatchTemplateMethod as M

pytestmark = pytest.mark.gpu


def bits(a):
    """bf16 bit patterns (as int16) of float values that are exact in bf16."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    assert np.all((u & 0xFFFF) == 0)
    return (u >> 16).astype(np.uint16).view(np.int16)


def dev16(a):
    return torch.from_numpy(bits(a).copy()).cuda()


def p(t):
    """Device pointer of t; t must stay referenced until the launch has run
    (a temporary tensor's memory is reused by the next allocation)."""
    return C.c_void_p(t.data_ptr())


def unbits(t):
    return (t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)


def rand_bf16(rng, shape, scale=1.0):
    return cnn.bf16_round((rng.standard_normal(shape) * scale).astype(np.float32))


def tol(ref_abs_sum, k):
    """f32 accumulation of k exact products: relative 2^-23 per addition."""
    return k * 2.0 ** -23 * ref_abs_sum + 1e-30


def test_conv_forward_matches_oracle():
    rng = np.random.default_rng(1)
    P = 3
    x = rand_bf16(rng, (P, 16, 16, 32))
    w = rand_bf16(rng, (32, 3, 3, 32), 0.1)
    b = rng.standard_normal(32).astype(np.float32)
    z = torch.empty(P * PIX * CH, dtype=torch.float32, device="cuda")
    a = torch.empty(P * PIX * CH, dtype=torch.int16, device="cuda")
    xd, wd, bd = dev16(x), dev16(w.reshape(32, KK)), torch.from_numpy(b).cuda()
    check(lib().iph_dev_cnn_conv3x3(p(xd), p(wd), p(bd), 1, P, p(z), p(a), None))
    ref = cnn.conv3x3(x, w, b.astype(np.float64))
    mag = cnn.conv3x3(np.abs(x), np.abs(w)) + np.abs(b)
    got = z.cpu().numpy().reshape(P, 16, 16, 32)
    assert np.all(np.abs(got - ref) <= tol(mag, 289))
    # relu output: the bf16 rounding of max(z, 0)
    want = cnn.bf16_round(np.maximum(got, 0.0).astype(np.float32))
    assert np.array_equal(unbits(a).reshape(P, 16, 16, 32), want.astype(np.float32))


def test_conv_gradients_match_oracle():
    rng = np.random.default_rng(2)
    P = 5
    x = rand_bf16(rng, (P, 16, 16, 32))
    w = rand_bf16(rng, (32, 3, 3, 32), 0.1)
    dy = rand_bf16(rng, (P, 16, 16, 32), 0.01)
    parts = 3
    ws = torch.empty(lib().iph_dev_cnn_conv3x3_dw_workspace(parts), dtype=torch.uint8, device="cuda")
    dw = torch.empty((32, KK), dtype=torch.float32, device="cuda")
    db = torch.empty(32, dtype=torch.float32, device="cuda")
    xd, dyd = dev16(x), dev16(dy)
    check(lib().iph_dev_cnn_conv3x3_dw(p(xd), p(dyd), P, parts, p(ws), ws.numel(),
                                       p(dw), p(db), None))
    rdw, rdb, rdx = cnn.conv3x3_backward(x, w, dy)
    mdw, mdb, mdx = cnn.conv3x3_backward(np.abs(x), np.abs(w), np.abs(dy))
    assert np.all(np.abs(dw.cpu().numpy() - rdw.reshape(32, KK)) <= tol(mdw.reshape(32, KK), P * PIX))
    assert np.all(np.abs(db.cpu().numpy() - rdb) <= tol(mdb, P * PIX))
    # input gradient: the forward layer with flipped weights
    wb = dev16(w.reshape(32, KK))
    wf = torch.empty_like(wb)
    check(lib().iph_dev_cnn_flip_weights(p(wb), p(wf), None))
    dx = torch.empty(P * PIX * CH, dtype=torch.float32, device="cuda")
    check(lib().iph_dev_cnn_conv3x3(p(dyd), p(wf), None, 0, P, p(dx), None, None))
    assert np.all(np.abs(dx.cpu().numpy().reshape(P, 16, 16, 32) - rdx) <= tol(mdx, 289))


def test_relu_backward_softmax_and_adam():
    rng = np.random.default_rng(3)
    n = 4096
    dx = rng.standard_normal(n).astype(np.float32)
    z = rng.standard_normal(n).astype(np.float32)
    out = torch.empty(n, dtype=torch.int16, device="cuda")
    dxd, zd = torch.from_numpy(dx).cuda(), torch.from_numpy(z).cuda()
    check(lib().iph_dev_cnn_relu_backward(p(dxd), p(zd), n, p(out), None))
    assert np.array_equal(unbits(out), cnn.bf16_round(np.where(z > 0, dx, 0.0).astype(np.float32)).astype(np.float32))
    # softmax cross-entropy over 2 pairs of 3 patches
    kpp, pairs = 3, 2
    zl = rng.standard_normal((pairs * kpp * PIX, CH)).astype(np.float32)
    tgt = np.array([5, 700], np.int32)
    loss = torch.empty(pairs, dtype=torch.float32, device="cuda")
    dz = torch.empty(pairs * kpp * PIX * CH, dtype=torch.int16, device="cuda")
    zld, tgd = torch.from_numpy(zl).cuda(), torch.from_numpy(tgt).cuda()
    check(lib().iph_dev_cnn_softmax_xent(p(zld), pairs, kpp, p(tgd), 0.5, p(loss), p(dz), None))
    g = unbits(dz).reshape(pairs * kpp * PIX, CH)
    for q in range(pairs):
        lo = zl[q * kpp * PIX:(q + 1) * kpp * PIX, 0].astype(np.float64)
        rl, rg = cnn.softmax_xent(lo, int(tgt[q]))
        assert abs(float(loss[q]) - rl) < 1e-4 * max(1.0, abs(rl))
        assert np.allclose(g[q * kpp * PIX:(q + 1) * kpp * PIX, 0], rg * 0.5, atol=2e-6, rtol=1e-2)
    assert np.all(g[:, 1:] == 0)
    # Adam: the textbook update
    prm = rng.standard_normal(100).astype(np.float32)
    gr = rng.standard_normal(100).astype(np.float32)
    m0 = np.zeros(100, np.float32)
    v0 = np.zeros(100, np.float32)
    tp, tg = torch.from_numpy(prm.copy()).cuda(), torch.from_numpy(gr).cuda()
    tm, tv = torch.zeros(100, device="cuda"), torch.zeros(100, device="cuda")
    check(lib().iph_dev_cnn_adam(p(tp), p(tg), p(tm), p(tv), 100, 1e-2, 0.9, 0.999, 1e-8, 1, None))
    m1 = 0.1 * gr
    v1 = 0.001 * gr * gr
    want = prm - 1e-2 * (m1 / 0.1) / (np.sqrt(v1 / 0.001) + 1e-8)
    assert np.allclose(tp.cpu().numpy(), want, rtol=1e-5, atol=1e-6)


def test_training_reduces_loss_on_a_synthetic_task():
    """Targets: the brightest pixel of channel 0 of each pair's patches."""
    rng = np.random.default_rng(4)
    net = ProposalNet(seed=1)
    kpp, pairs = 4, 8
    feats = np.zeros((pairs * kpp, PIX, CH), np.float32)
    feats[:, :, 0] = rng.standard_normal((pairs * kpp, PIX))
    feats[:, :, 3] = 1.0
    x = dev16(cnn.bf16_round(feats))
    tgt = np.array([int(np.argmax(feats[q * kpp:(q + 1) * kpp, :, 0])) for q in range(pairs)], np.int32)
    t = torch.from_numpy(tgt).cuda()
    first = net.train_step(x, t, kpp, lr=3e-3)
    for _ in range(60):
        last = net.train_step(x, t, kpp, lr=3e-3)
    assert last < 0.5 * first


def test_downsample_candidates_and_refine():
    rng = np.random.default_rng(5)
    img_np = O.gaussian_blur_f32(rng.integers(0, 256, (300, 420), dtype=np.uint8), 3.0)
    img = torch.from_numpy(img_np).cuda()
    small = torch.empty((75, 105), dtype=torch.uint8, device="cuda")
    check(lib().iph_dev_downsample4_u8(p(img), 420, 300, p(small), None))
    want = (img_np[:300, :420].reshape(75, 4, 105, 4).astype(np.int64).sum(axis=(1, 3)) + 8) >> 4
    assert np.array_equal(small.cpu().numpy(), want.astype(np.uint8))
    # refine around a known place returns the full map's argmax and value there
    tpl = img[100:148, 200:248].contiguous()
    srch = ProposalSearch(None, M.CrossCorrelationNormalized)
    full = dev.match_template(img, tpl, M.CrossCorrelationNormalized, "auto").cpu().numpy()
    fy, fx = np.unravel_index(np.argmax(full), full.shape)
    b = srch._buffers(300, 420, 48, 48)
    centers = torch.tensor([[int(fx) // 4, int(fy) // 4]], dtype=torch.int32, device="cuda")
    ws = torch.empty(lib().iph_dev_proposal_refine_workspace(1), dtype=torch.uint8, device="cuda")
    res = torch.empty(4, dtype=torch.float32, device="cuda")
    check(lib().iph_dev_proposal_refine(p(img), 420, 300, p(tpl), 48, 48, int(M.CrossCorrelationNormalized),
                                        p(centers), 1, 4, 4, p(res), p(ws), ws.numel(), None))
    r = res.cpu().numpy()
    assert (int(r[0]), int(r[1])) == (int(fx), int(fy))
    assert np.float32(r[2]) == full[fy, fx]
    # the candidate list holds the best tile maxima of the coarse map, best first
    bb = srch.candidates(img, tpl)
    cand = bb["cand"].cpu().numpy()
    cs = bb["cscore"].cpu().numpy()
    coarse = bb["coarse"].cpu().numpy()
    assert coarse[cand[0, 1], cand[0, 0]] == coarse.max() == cs[0]
    assert np.all(np.diff(cs[np.isfinite(cs)]) <= 0)


def test_search_finds_full_map_argmax_without_network():
    rng = np.random.default_rng(6)
    img_np = O.gaussian_blur_f32(rng.integers(0, 256, (512, 512), dtype=np.uint8), 4.0)
    img = torch.from_numpy(img_np).cuda()
    srch = ProposalSearch(None, M.CrossCorrelationNormalized)
    agree = 0
    for k in range(6):
        y, x = rng.integers(0, 512 - 48, 2)
        tpl = img[y:y + 48, x:x + 48].contiguous()
        gx, gy, gv = srch.full_map(img, tpl)
        sx, sy, sv, fb = srch.search(img, tpl, use_net=False)
        agree += (sx, sy) == (gx, gy)
        if (sx, sy) == (gx, gy):
            assert np.float32(sv) == np.float32(gv)
    assert agree >= 5
