"""Training and calibration of the proposal network, and the seeded pair
generator shared with the benchmark.

    python -m imageproc_amd.proposal_train [--pairs 200000] [--side 2048]

Pairs: a smooth-noise GrayImage (U[0,1) field, this library's f32
gaussian_blur_f32 at sigma 4, min-max rescaled to u8) and a 48x48 template
cropped from it at a random place, then perturbed: additive N(0, s) noise with
s in {0, 8, 16}, brightness x U(0.8, 1.2), and a random rectangle of 10 % of
its area set to a random grey; rounded and clamped. The label is the place of
the full CrossCorrelationNormalized map's maximum (match_template +
find_extremes: the plain search, the judge). Everything runs on the GPU
through libimageproc_hip.so; torch only holds device memory and draws the
seeded perturbations.

After training, the fallback thresholds of ProposalSearch are calibrated on
validation pairs of 8192x8192 images: the smallest exact-score margin such
that the searches that skip the full map agree with it on at least 99.5 % of
the pairs (and the rest fall back). Weights and thresholds go to
proposal_weights.npz next to this file.

    python -m imageproc_amd.proposal_train --calibrate-only [--cal-pairs 1200] [--eval-pairs 1200]

keeps the saved network and recalibrates its margin on one seeded set of
8192x8192 pairs (margin_rule), then evaluates the search on a separate
held-out set, recording both in the weights file.
"""
from __future__ import annotations

import argparse
import time

import numpy as np
import torch

from . import device as dev
from .proposal import PIX, WEIGHTS, CH, ProposalNet, ProposalSearch, coarse_target
from .template_matching import MatchTemplateMethod

CCN = MatchTemplateMethod.CrossCorrelationNormalized
TPL = 48


def _lsr(z, k):
    return (z >> k) & ((1 << (64 - k)) - 1)


def splitmix_u8(seed: int, n: int) -> torch.Tensor:
    """Top 8 bits of n SplitMix64 outputs from `seed`, on the GPU."""
    golden, c1, c2 = -7046029254386353131, -4658895280553007687, -7723592293110705685
    z = seed + torch.arange(1, n + 1, dtype=torch.int64, device="cuda") * golden
    z = (z ^ _lsr(z, 30)) * c1
    z = (z ^ _lsr(z, 27)) * c2
    z = z ^ _lsr(z, 31)
    return _lsr(z, 56).to(torch.uint8)


def smooth_noise(seed: int, side: int, sigma: float = 4.0) -> torch.Tensor:
    u = splitmix_u8(seed, side * side).view(side, side).float() / 256.0
    v = dev.gaussian_blur_f32(u.contiguous(), sigma)
    lo, hi = v.min(), v.max()
    return (255.0 * (v - lo) / (hi - lo)).to(torch.uint8).contiguous()


def perturbed_template(img: torch.Tensor, gen: torch.Generator, rng: np.random.Generator):
    """A 48x48 crop at a random place, perturbed (see the module docstring)."""
    h, w = img.shape
    y, x = int(rng.integers(0, h - TPL + 1)), int(rng.integers(0, w - TPL + 1))
    t = img[y:y + TPL, x:x + TPL].float()
    s = float(rng.choice([0.0, 8.0, 16.0]))
    if s > 0:
        t = t + s * torch.randn((TPL, TPL), generator=gen, device="cuda")
    t = t * float(rng.uniform(0.8, 1.2))
    area = 0.1 * TPL * TPL
    rw = int(rng.integers(8, TPL // 2 + 1))
    rh = min(TPL, max(1, int(round(area / rw))))
    ry, rx = int(rng.integers(0, TPL - rh + 1)), int(rng.integers(0, TPL - rw + 1))
    t[ry:ry + rh, rx:rx + rw] = float(rng.integers(0, 256))
    return torch.clamp(torch.round(t), 0, 255).to(torch.uint8).contiguous(), (x, y)


def pair_stream(seed: int, side: int, per_image: int = 16):
    """Endless seeded (image, template, crop place) pairs."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    k = 0
    while True:
        img = smooth_noise(seed * 1000003 + k, side)
        k += 1
        for _ in range(per_image):
            tpl, place = perturbed_template(img, gen, rng)
            yield img, tpl, place


def train(pairs: int, side: int, batch: int = 32, seed: int = 7, lr: float = 2e-3,
          log_every: int = 100) -> ProposalNet:
    net = ProposalNet(seed=seed)
    srch = ProposalSearch(net, CCN)
    stream = pair_stream(seed, side)
    steps = (pairs + batch - 1) // batch
    t0 = time.time()
    hits = 0
    for step in range(steps):
        xs, tg = [], []
        for _ in range(batch):
            img, tpl, _ = next(stream)
            gx, gy, _ = srch.full_map(img, tpl)
            b = srch.candidates(img, tpl)
            t = coarse_target(b["cand"].cpu().numpy(), gx, gy)
            hits += t >= 0
            xs.append(srch.features(b).clone())
            tg.append(t)
        x = torch.cat(xs)
        target = torch.tensor(tg, dtype=torch.int32, device="cuda")
        cur_lr = lr * 0.5 * (1.0 + np.cos(np.pi * step / steps))
        loss = net.train_step(x, target, srch.k, lr=cur_lr)
        if step % log_every == 0 or step == steps - 1:
            print(f"[train] step {step + 1}/{steps} pairs {(step + 1) * batch} loss {loss:.4f} "
                  f"target-in-candidates {hits / ((step + 1) * batch):.4f} "
                  f"{time.time() - t0:.0f}s", flush=True)
    return net


def evaluate(net: ProposalNet | None, n: int, side: int, seed: int, margin: float = 0.0,
             min_prob: float = 0.0):
    """Per pair: (net search place, coarse-only search place, judge place,
    best - second of the net search, its top probability)."""
    srch = ProposalSearch(net, CCN, margin=-1e30, min_prob=-1.0)
    out = []
    stream = pair_stream(seed, side, per_image=4)
    for i in range(n):
        if i and i % 200 == 0:
            print(f"[evaluate] {i}/{n} pairs", flush=True)
        img, tpl, _ = next(stream)
        gx, gy, _ = srch.full_map(img, tpl)
        nx, ny, best, second, p = srch.search_raw(img, tpl, use_net=net is not None)
        cx, cy, _, _ = srch.search(img, tpl, use_net=False)
        out.append(((nx, ny), (cx, cy), (gx, gy), best - second, p))
    return out


def calibrate(records, target: float = 0.995):
    """Smallest margin whose kept searches plus full-map fallbacks agree with
    the judge on at least `target` of the pairs; with the fallback rate."""
    margins = sorted({m for *_, m, _ in records})
    n = len(records)
    for m in [0.0] + margins:
        agree = sum(1 for (np_, _, g, mg, _) in records if (mg < m) or np_ == g)
        if agree >= target * n:
            fb = sum(1 for r in records if r[3] < m) / n
            return m, agree / n, fb
    return float("inf"), 1.0, 1.0


def margin_rule(records, floor_q: float = 0.01):
    """Fallback margin from calibration records: just above the largest
    best - second margin at which the network search disagreed with the
    judge (so every such pair would fall back), and at least the floor_q
    quantile of all margins (so the most ambiguous pairs fall back even when
    the calibration set has no disagreement)."""
    ms = np.array([r[3] for r in records], np.float64)
    bad = [r[3] for r in records if r[0] != r[2]]
    m = float(np.quantile(ms, floor_q)) if len(ms) else 0.0
    if bad:
        m = max(m, float(np.nextafter(np.float32(max(bad)), np.float32(np.inf))))
    return m


def summarise(records, margin: float):
    """Agreement with the judge of the network search with fallback, of the
    network search alone, of the coarse-only search, and the fallback rate."""
    n = len(records)
    fb = [r[3] < margin for r in records]
    with_fb = sum(1 for r, f in zip(records, fb) if f or r[0] == r[2]) / n
    return {"pairs": n, "agreement_with_fallback": with_fb,
            "net_only_agreement": sum(1 for r in records if r[0] == r[2]) / n,
            "coarse_only_agreement": sum(1 for r in records if r[1] == r[2]) / n,
            "fallback_rate": sum(fb) / n}


def calibrate_main(args) -> None:
    """Calibrates the fallback margin of the saved network on one seeded set
    of 8192x8192 pairs and evaluates the search on a separate held-out set,
    then records both in the weights file."""
    net = ProposalNet.load(args.out)
    cal = evaluate(net, args.cal_pairs, args.val_side, seed=90001)
    margin = margin_rule(cal)
    ev = evaluate(net, args.eval_pairs, args.val_side, seed=777013)
    sc, se = summarise(cal, margin), summarise(ev, margin)
    print(f"[calibrate] margin {margin:.6g}; calibration {sc}; held-out {se}", flush=True)
    with np.load(args.out, allow_pickle=False) as d:
        arrays = {k: d[k] for k in d.files}
    arrays["margin"] = np.float64(margin)
    arrays["cal_pairs"] = np.int64(sc["pairs"])
    arrays["eval_pairs"] = np.int64(se["pairs"])
    arrays["eval_agree_with_fallback"] = np.float64(se["agreement_with_fallback"])
    arrays["eval_net_agree"] = np.float64(se["net_only_agreement"])
    arrays["eval_coarse_agree"] = np.float64(se["coarse_only_agreement"])
    arrays["eval_fallback_rate"] = np.float64(se["fallback_rate"])
    np.savez(args.out, **arrays)
    print(f"[calibrate] wrote {args.out}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--calibrate-only", action="store_true",
                    help="keep the saved network; calibrate its margin and evaluate it")
    ap.add_argument("--cal-pairs", type=int, default=1200)
    ap.add_argument("--eval-pairs", type=int, default=1200)
    ap.add_argument("--pairs", type=int, default=200_000)
    ap.add_argument("--side", type=int, default=2048)
    ap.add_argument("--val-pairs", type=int, default=200)
    ap.add_argument("--val-side", type=int, default=8192)
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()
    if args.calibrate_only:
        calibrate_main(args)
        return
    net = train(args.pairs, args.side)
    print("[train] calibrating on validation pairs", flush=True)
    rec = evaluate(net, args.val_pairs, args.val_side, seed=90001)
    margin, agree, fb = calibrate(rec)
    coarse_agree = sum(1 for (_, c, g, _, _) in rec if c == g) / len(rec)
    net_agree = sum(1 for (n_, _, g, _, _) in rec if n_ == g) / len(rec)
    print(f"[train] validation: net-only agreement {net_agree:.4f}, coarse-only {coarse_agree:.4f}, "
          f"margin {margin:.6g} -> agreement {agree:.4f} with {fb:.4f} fallbacks", flush=True)
    net.save(args.out)
    with np.load(args.out, allow_pickle=False) as d:
        arrays = {k: d[k] for k in d.files}
    arrays["margin"] = np.float64(margin)
    arrays["val_net_agree"] = np.float64(net_agree)
    arrays["val_coarse_agree"] = np.float64(coarse_agree)
    arrays["train_pairs"] = np.int64(args.pairs)
    np.savez(args.out, **arrays)
    print(f"[train] wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
