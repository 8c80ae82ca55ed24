"""Learned match proposal: find match_template's best place without the full
score map, and fall back to the full map when unsure.

No counterpart in the reference, whose only search is the full map of
match_template (template_matching.rs:78-92) read by find_extremes
(template_matching.rs:479-511); that search stays the judge of this one.

Search (`ProposalSearch.search`: one call of iph_dev_proposal_search, which
enqueues every step below as kernels of libimageproc_hip.so, then one read of
its 5-float result; `search_steps` makes the same search from the step-wise
entry points, value for value):
  1. 4x-downsample image and template (rounded 4x4 means) and compute the
     coarse score map with the library's own match_template (int8 MFMA);
  2. take the best place of each 16x16 tile of the coarse map and keep the
     k = 64 best tiles;
  3. the proposal network scores every coarse place of a 16x16 patch around
     each candidate (features: the coarse scores around it, its score against
     the best candidate's, the small image's pixels, a validity mask);
  4. the 16 best-scored places, scaled back to full resolution, get exact
     scores in a +/-8 px window each (the full map's values at those places);
  5. the best of those is the answer, unless the runner-up more than 4 px away
     is within `margin` of it or the network's top probability is below
     `min_prob`: then the full map is computed and searched instead.

Network (`ProposalNet`): 4 layers of 3x3 convolutions, 32 channels, ReLU
between, bf16 activations and weights with f32 accumulation on MFMA; trained
with softmax cross-entropy over all logits of a pair's candidate patches
against the place of the full map's argmax, by the library's own forward and
backward kernels and Adam.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

from . import device as dev
from ._lib import check, lib
from .template_matching import MatchTemplateMethod

CH = 32          # channels of every layer
PS = 16          # patch side
PIX = PS * PS
KK = 9 * CH
LAYERS = 4
REAL_IN = (4, 32, 32, 32)
REAL_OUT = (32, 32, 32, 1)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "proposal_weights.npz")


def _p(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _s() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class ProposalNet:
    """The proposal network's parameters (f32 masters, bf16 copies, flipped
    copies for input gradients) and Adam state, all on the GPU."""

    def __init__(self, seed: int = 0):
        rng = np.random.default_rng(seed)
        # one contiguous block per kind (layer l at offset l), as the fused
        # search entry point takes them
        self.w_all = torch.zeros((LAYERS, CH, KK), dtype=torch.float32, device="cuda")
        self.b_all = torch.zeros((LAYERS, CH), dtype=torch.float32, device="cuda")
        self.w, self.b = list(self.w_all.unbind(0)), list(self.b_all.unbind(0))
        for l in range(LAYERS):
            w = np.zeros((CH, 3, 3, CH), np.float32)
            fan_in = 9 * REAL_IN[l]
            w[:REAL_OUT[l], :, :, :REAL_IN[l]] = (rng.standard_normal(
                (REAL_OUT[l], 3, 3, REAL_IN[l])) * np.sqrt(2.0 / fan_in)).astype(np.float32)
            self.w[l].copy_(torch.from_numpy(w.reshape(CH, KK)))
        self._adam = [(torch.zeros_like(t), torch.zeros_like(t)) for t in self.w + self.b]
        self.step = 0
        self._dw_parts = 64
        self._dw_ws = torch.empty(lib().iph_dev_cnn_conv3x3_dw_workspace(self._dw_parts),
                                  dtype=torch.uint8, device="cuda")
        self.refresh()

    def refresh(self) -> None:
        """bf16 and flipped copies of the f32 weights."""
        self.wb_all = torch.empty((LAYERS, CH * KK), dtype=torch.int16, device="cuda")
        check(lib().iph_dev_cnn_to_bf16(_p(self.w_all), LAYERS * CH * KK, _p(self.wb_all), _s()))
        self.wb, self.wf = list(self.wb_all.unbind(0)), []
        for wb in self.wb:
            wf = torch.empty_like(wb)
            check(lib().iph_dev_cnn_flip_weights(_p(wb), _p(wf), _s()))
            self.wf.append(wf)

    def forward(self, x: torch.Tensor, keep: bool = False):
        """x: bf16 bits (int16) of shape (P, 256, 32). Returns the last layer's
        f32 output (P, 256, 32), whose channel 0 is the logit; with keep, also
        each layer's input and pre-activation."""
        n = x.shape[0]
        acts, zs = [x], []
        for l in range(LAYERS):
            last = l == LAYERS - 1
            z = torch.empty((n, PIX, CH), dtype=torch.float32, device="cuda") if (keep or last) else None
            a = None if last else torch.empty((n, PIX, CH), dtype=torch.int16, device="cuda")
            check(lib().iph_dev_cnn_conv3x3(_p(acts[-1]), _p(self.wb[l]), _p(self.b[l]),
                                            0 if last else 1, n, _p(z) if z is not None else None,
                                            _p(a) if a is not None else None, _s()))
            zs.append(z)
            if a is not None:
                acts.append(a)
        return (zs[-1], acts, zs) if keep else zs[-1]

    def gradients(self, x: torch.Tensor, target: torch.Tensor, kpp: int):
        """Mean loss over pairs with a target and the gradients of every
        parameter (w then b per layer), from the library's backward kernels."""
        n_pairs = target.numel()
        zl, acts, zs = self.forward(x, keep=True)
        n_valid = max(int((target >= 0).sum()), 1)
        loss = torch.empty(n_pairs, dtype=torch.float32, device="cuda")
        dz = torch.empty((x.shape[0], PIX, CH), dtype=torch.int16, device="cuda")
        check(lib().iph_dev_cnn_softmax_xent(_p(zl), n_pairs, kpp, _p(target), 1.0 / n_valid,
                                             _p(loss), _p(dz), _s()))
        gw, gb = [None] * LAYERS, [None] * LAYERS
        for l in reversed(range(LAYERS)):
            gw[l] = torch.empty((CH, KK), dtype=torch.float32, device="cuda")
            gb[l] = torch.empty(CH, dtype=torch.float32, device="cuda")
            check(lib().iph_dev_cnn_conv3x3_dw(_p(acts[l]), _p(dz), x.shape[0], self._dw_parts,
                                               _p(self._dw_ws), self._dw_ws.numel(), _p(gw[l]),
                                               _p(gb[l]), _s()))
            if l > 0:
                dx = torch.empty((x.shape[0], PIX, CH), dtype=torch.float32, device="cuda")
                check(lib().iph_dev_cnn_conv3x3(_p(dz), _p(self.wf[l]), None, 0, x.shape[0],
                                                _p(dx), None, _s()))
                dz = torch.empty((x.shape[0], PIX, CH), dtype=torch.int16, device="cuda")
                check(lib().iph_dev_cnn_relu_backward(_p(dx), _p(zs[l - 1]), dx.numel(), _p(dz),
                                                      _s()))
        return float(loss.sum()) / n_valid, gw + gb

    def train_step(self, x: torch.Tensor, target: torch.Tensor, kpp: int, lr: float = 1e-3) -> float:
        loss, grads = self.gradients(x, target, kpp)
        self.step += 1
        for p, g, (m, v) in zip(self.w + self.b, grads, self._adam):
            check(lib().iph_dev_cnn_adam(_p(p), _p(g), _p(m), _p(v), p.numel(), lr, 0.9, 0.999,
                                         1e-8, self.step, _s()))
        self.refresh()
        return loss

    def save(self, path: str = WEIGHTS) -> None:
        np.savez(path, **{f"w{l}": self.w[l].cpu().numpy() for l in range(LAYERS)},
                 **{f"b{l}": self.b[l].cpu().numpy() for l in range(LAYERS)})

    @classmethod
    def load(cls, path: str = WEIGHTS) -> "ProposalNet":
        net = cls()
        with np.load(path, allow_pickle=False) as d:
            for l in range(LAYERS):
                net.w[l].copy_(torch.from_numpy(d[f"w{l}"]))
                net.b[l].copy_(torch.from_numpy(d[f"b{l}"]))
        net.refresh()
        return net


class ProposalSearch:
    """The search of the module docstring for one template size (buffers are
    cached per image / template shape)."""

    def __init__(self, net: ProposalNet | None, method=MatchTemplateMethod.CrossCorrelationNormalized,
                 k: int = 64, n_top: int = 16, margin: float | None = None, min_prob: float = 0.0,
                 sep: int = 4):
        """margin None takes the margin calibrated with the shipped weights
        (proposal_train --calibrate-only): below it the full map decides."""
        self.net, self.method = net, MatchTemplateMethod(method)
        if margin is None:
            with np.load(WEIGHTS, allow_pickle=False) as d:
                margin = float(d["margin"]) if "margin" in d.files else 0.0
        if self.method not in (MatchTemplateMethod.CrossCorrelation,
                               MatchTemplateMethod.CrossCorrelationNormalized):
            raise ValueError("the proposal search serves the correlation methods")
        self.k, self.n_top, self.margin, self.min_prob, self.sep = k, n_top, margin, min_prob, sep
        self._cache: dict = {}

    def _buffers(self, h, w, th, tw):
        key = (h, w, th, tw)
        b = self._cache.get(key)
        if b is None:
            sh, sw, sth, stw = h // 4, w // 4, th // 4, tw // 4
            ch_, cw_ = sh - sth + 1, sw - stw + 1
            b = {
                "small": torch.empty((sh, sw), dtype=torch.uint8, device="cuda"),
                "stpl": torch.empty((sth, stw), dtype=torch.uint8, device="cuda"),
                "coarse": torch.empty((ch_, cw_), dtype=torch.float32, device="cuda"),
                "cand": torch.empty((self.k, 2), dtype=torch.int32, device="cuda"),
                "cscore": torch.empty(self.k, dtype=torch.float32, device="cuda"),
                "cws": torch.empty(lib().iph_dev_proposal_candidates_workspace(cw_, ch_),
                                   dtype=torch.uint8, device="cuda"),
                "patches": torch.empty((self.k, PIX, CH), dtype=torch.int16, device="cuda"),
                "centers": torch.empty((self.n_top, 2), dtype=torch.int32, device="cuda"),
                "probs": torch.empty(self.n_top, dtype=torch.float32, device="cuda"),
                "tws": torch.empty(lib().iph_dev_proposal_top_workspace(self.k, self.n_top),
                                   dtype=torch.uint8, device="cuda"),
                "rws": torch.empty(lib().iph_dev_proposal_refine_workspace(self.n_top),
                                   dtype=torch.uint8, device="cuda"),
                "sws": torch.empty(lib().iph_dev_proposal_search_workspace(
                    w, h, tw, th, int(self.method), self.k, self.n_top), dtype=torch.uint8,
                    device="cuda"),
                "result": torch.empty(5, dtype=torch.float32, device="cuda"),
                "host": torch.empty(5, dtype=torch.float32, pin_memory=True),
                "shape": (sh, sw, sth, stw, ch_, cw_),
            }
            self._cache[key] = b
        return b

    def candidates(self, image: torch.Tensor, tpl: torch.Tensor):
        """Steps 1-2; returns the buffers (coarse map, candidates, small image)."""
        h, w = image.shape
        th, tw = tpl.shape
        b = self._buffers(h, w, th, tw)
        sh, sw, sth, stw, ch_, cw_ = b["shape"]
        s = _s()
        check(lib().iph_dev_downsample4_u8(_p(image), w, h, _p(b["small"]), s))
        check(lib().iph_dev_downsample4_u8(_p(tpl), tw, th, _p(b["stpl"]), s))
        dev.match_template(b["small"], b["stpl"], self.method, "mfma", out=b["coarse"])
        check(lib().iph_dev_proposal_candidates(_p(b["coarse"]), cw_, ch_, self.k, _p(b["cand"]),
                                                _p(b["cscore"]), _p(b["cws"]), b["cws"].numel(), s))
        return b

    def features(self, b) -> torch.Tensor:
        sh, sw, sth, stw, ch_, cw_ = b["shape"]
        check(lib().iph_dev_proposal_features(_p(b["coarse"]), cw_, ch_, _p(b["small"]), sw, sh,
                                              stw, _p(b["cand"]), _p(b["cscore"]), self.k,
                                              _p(b["patches"]), _s()))
        return b["patches"]

    def top(self, zl: torch.Tensor, b) -> torch.Tensor:
        """The network's n_top best distinct coarse places (and b["probs"])."""
        sh, sw, sth, stw, ch_, cw_ = b["shape"]
        check(lib().iph_dev_proposal_top(_p(zl), self.k, _p(b["cand"]), cw_, ch_, self.n_top,
                                         _p(b["centers"]), _p(b["probs"]), _p(b["tws"]),
                                         b["tws"].numel(), _s()))
        return b["centers"]

    def refine(self, image, tpl, b, centers) -> None:
        h, w = image.shape
        th, tw = tpl.shape
        check(lib().iph_dev_proposal_refine(_p(image), w, h, _p(tpl), tw, th, int(self.method),
                                            _p(centers), centers.shape[0], 4, self.sep,
                                            _p(b["result"]), _p(b["rws"]), b["rws"].numel(), _s()))

    def full_map(self, image, tpl):
        """The plain search: full score map, then find_extremes' argmax
        (x, y, score). The map and the record stay allocated per shape."""
        h, w = image.shape
        th, tw = tpl.shape
        key = ("full", h, w, th, tw)
        b = self._cache.get(key)
        if b is None:
            b = {"map": torch.empty((h - th + 1, w - tw + 1), dtype=torch.float32, device="cuda"),
                 "rec": torch.empty(24, dtype=torch.uint8, device="cuda"),
                 "host": torch.empty(24, dtype=torch.uint8, pin_memory=True)}
            self._cache[key] = b
        dev.match_template(image, tpl, self.method, "auto", out=b["map"])
        dev.find_extremes(b["map"], out=b["rec"])
        b["host"].copy_(b["rec"], non_blocking=True)
        torch.cuda.current_stream().synchronize()
        rec = b["host"].numpy().view(np.uint32)
        return int(rec[2]), int(rec[3]), float(rec[0:1].view(np.float32)[0])

    def search_steps(self, image: torch.Tensor, tpl: torch.Tensor, use_net: bool = True):
        """The search from the step-wise entry points: (x, y, best, second,
        top probability) before the fallback decision."""
        b = self.candidates(image, tpl)
        if use_net and self.net is not None:
            centers = self.top(self.net.forward(self.features(b)), b)
            p_top = float(b["probs"][0])
        else:
            centers, p_top = b["cand"][: self.n_top], 1.0
        self.refine(image, tpl, b, centers)
        r = b["result"][:4].cpu().numpy()
        return int(r[0]), int(r[1]), float(r[2]), float(r[3]), p_top

    def search_raw(self, image: torch.Tensor, tpl: torch.Tensor, use_net: bool = True):
        """One iph_dev_proposal_search call: (x, y, best, second, top
        probability) before the fallback decision."""
        h, w = image.shape
        th, tw = tpl.shape
        b = self._buffers(h, w, th, tw)
        net = self.net if use_net else None
        check(lib().iph_dev_proposal_search(
            _p(image), w, h, _p(tpl), tw, th, int(self.method),
            _p(net.wb_all) if net is not None else None,
            _p(net.b_all) if net is not None else None, self.k, self.n_top, self.sep,
            _p(b["result"]), _p(b["sws"]), b["sws"].numel(), _s()))
        host = b["host"]
        check(lib().iph_dev_read_result(_p(b["result"]), C.c_void_p(host.data_ptr()), 5, _s()))
        r = host.numpy()
        return int(r[0]), int(r[1]), float(r[2]), float(r[3]), float(r[4])

    def search(self, image: torch.Tensor, tpl: torch.Tensor, use_net: bool = True):
        """(x, y, score, used_full_map) of match_template's best place."""
        x, y, best, second, p_top = self.search_raw(image, tpl, use_net)
        if not (best == best) or best - second < self.margin or p_top < self.min_prob:
            return (*self.full_map(image, tpl), True)
        return x, y, best, False


def coarse_target(cand: np.ndarray, lx: int, ly: int) -> int:
    """Training target: the first patch pixel whose coarse place is the full
    map's argmax (lx, ly) divided by 4 (rounded), or -1."""
    tx, ty = int(round(lx / 4)), int(round(ly / 4))
    for c, (cx, cy) in enumerate(cand):
        px, py = tx - (int(cx) - 8), ty - (int(cy) - 8)
        if 0 <= px < PS and 0 <= py < PS:
            return c * PIX + py * PS + px
    return -1
