"""Batches of uneven images (iph_batch_*): every job of a planned batch, replayed
as a hipGraph or enqueued eagerly, equals the single call it stands for, bit
for bit; blur and filter jobs also equal the oracle (filter/mod.rs:156-164,
202-210; template_matching.rs:78-92 for the match jobs, whose single calls are
checked against the oracle in test_gpu_match.py)."""
import numpy as np
import pytest
import torch

import oracle as O
from imageproc_amd import device as dev
from imageproc_amd.batch import Batch
from imageproc_amd.kernel import Kernel
from imageproc_amd.template_matching import MatchTemplateMethod as M

pytestmark = pytest.mark.gpu


def kernel5(rng):
    """5x5 N(0,1) weights normalised to sum 1, redrawn while |sum| < 0.25."""
    while True:
        k = rng.standard_normal((5, 5)).astype(np.float32)
        s = np.float32(k.sum(dtype=np.float32))
        if abs(s) >= 0.25:
            return (k / s).astype(np.float32)


def make_batch(seed, n_img):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n_img):
        h, w = (int(v) for v in rng.integers(33, 300, size=2))
        img = torch.from_numpy(rng.integers(0, 256, (h, w), dtype=np.uint8)).cuda()
        sigma = float(rng.uniform(0.5, 8.0))
        k = kernel5(rng)
        tpls = []
        for _ in range(int(rng.integers(1, 4))):
            th = int(rng.integers(8, min(40, h) + 1))
            tw = int(rng.integers(8, min(40, w) + 1))
            y, x = int(rng.integers(0, h - th + 1)), int(rng.integers(0, w - tw + 1))
            tpls.append((img[y:y + th, x:x + tw].contiguous(), M(int(rng.integers(0, 4)))))
        jobs.append((img, sigma, k, tpls))
    return jobs


def expected(jobs):
    out = []
    for img, sigma, k, tpls in jobs:
        b = dev.gaussian_blur_f32(img, sigma)
        f = dev.filter_clamped(img, Kernel(k, 5, 5), torch.uint8)
        ms = [dev.match_template(img, t, m, "auto") for t, m in tpls]
        out.append((b, f, ms))
    torch.cuda.synchronize()
    return out


def build(jobs, autotune):
    bt = Batch()
    outs = []
    for img, sigma, k, tpls in jobs:
        b = bt.add_gaussian_blur(img, sigma)
        f = bt.add_filter_clamped(img, k)
        ms = [bt.add_match_template(img, t, m) for t, m in tpls]
        outs.append((b, f, ms))
    bt.plan(autotune=autotune)
    return bt, outs


def clear(outs):
    for b, f, ms in outs:
        b.zero_()
        f.zero_()
        for m in ms:
            m.fill_(-1.0)


def assert_same(outs, exp):
    for (b, f, ms), (eb, ef, ems) in zip(outs, exp):
        assert torch.equal(b, eb)
        assert torch.equal(f, ef)
        for m, em in zip(ms, ems):
            assert torch.equal(m.view(torch.int32), em.view(torch.int32))


@pytest.mark.parametrize("seed,autotune", [(11, True), (12, False)])
def test_batch_graph_replays_equal_single_calls(seed, autotune):
    jobs = make_batch(seed, 10)
    exp = expected(jobs)
    bt, outs = build(jobs, autotune)
    for _ in range(3):
        clear(outs)
        torch.cuda.synchronize()
        bt.run()
        torch.cuda.synchronize()
        assert_same(outs, exp)
    clear(outs)
    bt.run_eager()
    torch.cuda.synchronize()
    assert_same(outs, exp)
    if autotune:
        assert "chosen:" in bt.tuning()
    bt.close()


def test_batch_jobs_equal_oracle():
    jobs = make_batch(13, 4)
    bt, outs = build(jobs, False)
    bt.run()
    torch.cuda.synchronize()
    for (img, sigma, k, _), (b, f, _) in zip(jobs, outs):
        host = img.cpu().numpy()
        assert np.array_equal(b.cpu().numpy(), O.gaussian_blur_f32(host, sigma))
        assert np.array_equal(f.cpu().numpy(), O.filter_clamped(host, k, 5, 5, np.uint8))
    bt.close()


@pytest.mark.parametrize("shapes", [[(3, 3), (5, 5), (1, 5), (5, 3), (2, 4)], [(3, 3), (1, 3), (3, 1)]])
def test_batch_filter_kernel_sizes_equal_oracle(shapes):
    """Batched filter_clamped with kernels of several sizes in one table: the
    jobs of the table's largest size (5x5 or 3x3) take the eight-outputs-per-
    thread path, the others the generic one; every output equals the oracle."""
    rng = np.random.default_rng(21)
    bt = Batch()
    cases = []
    for i, (kh, kw) in enumerate(shapes):
        h, w = int(rng.integers(20, 150)), int(rng.integers(20, 150))
        img = torch.from_numpy(rng.integers(0, 256, (h, w), dtype=np.uint8)).cuda()
        k = rng.normal(0.0, 0.3, (kh, kw)).astype(np.float32)
        cases.append((img, k, bt.add_filter_clamped(img, k)))
    bt.plan(autotune=False)
    bt.run()
    torch.cuda.synchronize()
    for img, k, out in cases:
        kh, kw = k.shape
        assert np.array_equal(out.cpu().numpy(),
                              O.filter_clamped(img.cpu().numpy(), k.reshape(-1), kw, kh, np.uint8))
    bt.close()


def test_batch_many_small_and_one_large_image():
    rng = np.random.default_rng(14)
    imgs = [torch.from_numpy(rng.integers(0, 256, (int(h), int(w)), dtype=np.uint8)).cuda()
            for h, w in [(8, 8), (1, 300), (300, 1), (1100, 1500), (5, 7), (130, 129)]]
    bt = Batch()
    ks = [kernel5(rng) for _ in imgs]
    outs = [(bt.add_gaussian_blur(im, 1.7), bt.add_filter_clamped(im, k)) for im, k in zip(imgs, ks)]
    bt.plan(autotune=False)
    bt.run()
    torch.cuda.synchronize()
    for im, k, (b, f) in zip(imgs, ks, outs):
        assert torch.equal(b, dev.gaussian_blur_f32(im, 1.7))
        assert torch.equal(f, dev.filter_clamped(im, Kernel(k, 5, 5), torch.uint8))
    bt.close()


def _check_match_vs_oracle(got, img, tpl, m):
    """The int8 MFMA bounds of test_gpu_match.py: SSE and CC are the exact
    integer sums rounded once to f32; the normalised methods are within 4 f32
    ulps of FP64 below 64x64 windows and 0.5 ulp from there on. Templates
    above 128 px take the direct path: bit-identical to the oracle."""
    if max(tpl.shape) > 128:
        assert np.array_equal(got.view(np.uint32),
                              O.match_template(img, tpl, int(m), nthreads=8).view(np.uint32)), m
        return
    exact = O.match_template_f64(img, tpl, int(m))
    if m in (M.SumOfSquaredErrors, M.CrossCorrelation):
        assert np.array_equal(got.view(np.uint32), exact.astype(np.float32).view(np.uint32)), m
    else:
        rel = np.abs(got - exact) / np.maximum(np.abs(exact), 1e-30)
        assert rel.max() <= (4.0 if tpl.size < 4096 else 0.5) * 2.0 ** -23 + 1e-30, (m, rel.max())
    if tpl.size * 65025 < 2 ** 24:
        assert np.array_equal(got.view(np.uint32),
                              O.match_template(img, tpl, int(m), nthreads=8).view(np.uint32)), m


def test_batch_match_jobs_equal_oracle():
    """Batched matches against the oracle: templates of 8 to 128 px (every K-block
    count of the fused job-table kernel, one and several template-row chunks)
    and one of 130 columns (the direct path inside the batch), all four
    methods, images of uneven sizes."""
    rng = np.random.default_rng(15)
    shapes = [((150, 170), (8, 8)), ((200, 160), (40, 33)), ((180, 210), (65, 66)),
              ((230, 190), (96, 90)), ((240, 260), (100, 97)), ((140, 150), (128, 128)),
              ((97, 301), (31, 64)), ((260, 75), (17, 70)), ((60, 150), (20, 130))]
    bt = Batch()
    cases = []
    for k, ((h, w), (th, tw)) in enumerate(shapes):
        img = rng.integers(0, 256, (h, w), dtype=np.uint8)
        tpl = rng.integers(0, 256, (th, tw), dtype=np.uint8)
        gi, gt = torch.from_numpy(img).cuda(), torch.from_numpy(tpl).cuda()
        for m in M:
            cases.append((img, tpl, m, bt.add_match_template(gi, gt, m)))
    bt.plan(autotune=True)
    for _ in range(2):
        for _, _, _, o in cases:
            o.fill_(-1.0)
        bt.run()
        torch.cuda.synchronize()
        for img, tpl, m, o in cases:
            _check_match_vs_oracle(o.cpu().numpy(), img, tpl, m)
    assert "match table" in bt.tuning()
    bt.close()
