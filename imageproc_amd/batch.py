"""Batches of uneven GrayImages on one GPU: gaussian_blur_f32, filter_clamped
(f32 kernel, u8 output) and match_template jobs planned together and replayed
as one hipGraph (iph_batch_* in include/imageproc_hip.h).

The reference has no batch API; a batch stands for a loop of per-image calls
(filter/mod.rs:156-164, filter/mod.rs:202-210, template_matching.rs:78-92),
and every job's output is bit-identical to that call's. Inputs and outputs
are device tensors owned by the caller; the batch keeps references to them so
they outlive it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, lib
from .template_matching import MatchTemplateMethod


def _u8(t: torch.Tensor) -> torch.Tensor:
    if not (t.is_cuda and t.is_contiguous() and t.dtype == torch.uint8 and t.dim() == 2):
        raise ValueError("expected a contiguous 2-D uint8 device tensor")
    return t


class Batch:
    def __init__(self):
        h = C.c_void_p()
        check(lib().iph_batch_create(C.byref(h)))
        self._h = h
        self._keep: list = []

    def close(self) -> None:
        if self._h:
            lib().iph_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_gaussian_blur(self, src: torch.Tensor, sigma: float,
                          out: torch.Tensor | None = None) -> torch.Tensor:
        src = _u8(src)
        out = torch.empty_like(src) if out is None else _u8(out)
        h, w = src.shape
        check(lib().iph_batch_add_gaussian_blur(self._h, src.data_ptr(), out.data_ptr(), w, h,
                                                float(sigma)))
        self._keep += [src, out]
        return out

    def add_filter_clamped(self, src: torch.Tensor, kernel,
                           out: torch.Tensor | None = None) -> torch.Tensor:
        """kernel: 2-D array of f32 weights (kh x kw), at most 49 taps."""
        src = _u8(src)
        k = np.ascontiguousarray(np.asarray(kernel, dtype=np.float32))
        if k.ndim != 2:
            raise ValueError("kernel must be 2-D")
        kh, kw = k.shape
        out = torch.empty_like(src) if out is None else _u8(out)
        h, w = src.shape
        check(lib().iph_batch_add_filter_clamped(self._h, src.data_ptr(), out.data_ptr(), w, h,
                                                 k.ctypes.data_as(C.c_void_p), kw, kh))
        self._keep += [src, out]
        return out

    def add_match_template(self, image: torch.Tensor, template: torch.Tensor,
                           method: MatchTemplateMethod,
                           out: torch.Tensor | None = None) -> torch.Tensor:
        image, template = _u8(image), _u8(template)
        h, w = image.shape
        th, tw = template.shape
        if out is None:
            out = torch.empty((max(h - th + 1, 0), max(w - tw + 1, 0)), dtype=torch.float32,
                              device=image.device)
        check(lib().iph_batch_add_match_template(self._h, image.data_ptr(), w, h,
                                                 template.data_ptr(), tw, th, int(method),
                                                 out.data_ptr()))
        self._keep += [image, template, out]
        return out

    def plan(self, autotune: bool = True) -> None:
        s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(lib().iph_batch_plan(self._h, 1 if autotune else 0, s))

    def run(self) -> None:
        """Replays the captured graph on torch's current stream."""
        check(lib().iph_batch_run(self._h, C.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def run_eager(self) -> None:
        """Enqueues the same launches without the graph."""
        check(lib().iph_batch_run_eager(self._h,
                                        C.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def tuning(self) -> str:
        buf = C.create_string_buffer(4096)
        check(lib().iph_batch_tuning(self._h, buf, len(buf)))
        return buf.value.decode()
