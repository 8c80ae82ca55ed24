"""imageproc::template_matching on gfx950 (template_matching.rs).

``match_template`` runs on MFMA by default: int8 MFMA with exact i32 sums
for GrayImage inputs, BF16x3 split products with FP32 accumulation for
Luma<f32> inputs. ``path="direct"`` runs the reference's own per-window f32
loop on the GPU instead, bit-identical to the reference.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass

import numpy as np

from ._lib import F32, PATH_AUTO, PATH_DIRECT, PATH_MFMA, U8, Extremes as _CExtremes
from ._lib import check, lib, ptr

__all__ = ["MatchTemplateMethod", "match_template", "match_template_parallel", "Extremes",
           "find_extremes"]


class MatchTemplateMethod(enum.IntEnum):
    """template_matching.rs:8-64, in declaration order."""

    SumOfSquaredErrors = 0
    SumOfSquaredErrorsNormalized = 1
    CrossCorrelation = 2
    CrossCorrelationNormalized = 3


_PATHS = {"auto": PATH_AUTO, "mfma": PATH_MFMA, "direct": PATH_DIRECT}


def _gray_or_f32(a, what):
    a = np.ascontiguousarray(a)
    if a.ndim != 2:
        raise TypeError(f"{what} must be a 2-D (height, width) array")
    if a.dtype == np.uint8:
        return a, U8
    if a.dtype == np.float32:
        return a, F32
    raise TypeError(f"{what} pixel type {a.dtype} is not u8 or f32")


def _gray(a, what):
    a = np.ascontiguousarray(a)
    if a.ndim != 2 or a.dtype != np.uint8:
        raise TypeError(f"{what} must be a 2-D uint8 GrayImage array")
    return a


def match_template(image, template, method: MatchTemplateMethod, path: str = "auto"):
    """template_matching.rs:78-92. Returns a float32 array of shape
    (h - th + 1, w - tw + 1). Panics if the template is larger than the image."""
    img, p = _gray_or_f32(image, "image")
    tpl, pt = _gray_or_f32(template, "template")
    if p != pt:
        raise TypeError("image and template must have the same pixel type")
    h, w = img.shape
    th, tw = tpl.shape
    m = int(MatchTemplateMethod(method))
    oh, ow = max(h - th + 1, 0), max(w - tw + 1, 0)
    out = np.empty((oh, ow), np.float32)
    check(lib().iph_match_template(ptr(img), p, w, h, ptr(tpl), tw, th, m, _PATHS[path],
                                   ptr(out)))
    return out


def match_template_parallel(image, template, method: MatchTemplateMethod, path: str = "auto"):
    """template_matching.rs:96-110: the same GPU path as match_template."""
    return match_template(image, template, method, path)


def match_template_with_mask(image, template, method: MatchTemplateMethod, mask,
                             path: str = "auto"):
    """template_matching.rs:125-140 (GrayImages; the masked methods of
    :295-376). path "direct": the reference's operation order, bit-identical;
    "mfma": CrossCorrelation(Normalized) as BF16x3 MFMA correlations of the
    image and its square with T*M^2 and M^2 (templates up to 128 px); "auto":
    "mfma" where it applies, else "direct". Panics if the mask is not the
    template's size, or the template is larger than the image."""
    img = _gray(image, "image")
    tpl = _gray(template, "template")
    msk = _gray(mask, "mask")
    h, w = img.shape
    th, tw = tpl.shape
    mh, mw = msk.shape
    m = int(MatchTemplateMethod(method))
    oh, ow = max(h - th + 1, 0), max(w - tw + 1, 0)
    out = np.empty((oh, ow), np.float32)
    check(lib().iph_match_template_with_mask(ptr(img), w, h, ptr(tpl), tw, th, ptr(msk), mw, mh,
                                             m, _PATHS[path], ptr(out)))
    return out


def match_template_with_mask_parallel(image, template, method: MatchTemplateMethod, mask,
                                      path: str = "auto"):
    """template_matching.rs:142-163: the same GPU path."""
    return match_template_with_mask(image, template, method, mask, path)


@dataclass(frozen=True)
class Extremes:
    """template_matching.rs:463-475."""

    max_value: object
    min_value: object
    max_value_location: tuple
    min_value_location: tuple


def find_extremes(image) -> Extremes:
    """template_matching.rs:479-511: first occurrence in row-major order wins
    ties. Panics on an empty image."""
    img, p = _gray_or_f32(image, "image")
    h, w = img.shape
    r = _CExtremes()
    check(lib().iph_find_extremes(ptr(img), p, w, h, ptr_struct(r)))
    if p == U8:
        mx, mn = np.uint8(r.max_raw & 0xFF), np.uint8(r.min_raw & 0xFF)
    else:
        mx = np.array([r.max_raw], np.uint32).view(np.float32)[0]
        mn = np.array([r.min_raw], np.uint32).view(np.float32)[0]
    return Extremes(mx, mn, (r.max_x, r.max_y), (r.min_x, r.min_y))


def ptr_struct(s):
    import ctypes
    return ctypes.cast(ctypes.pointer(s), ctypes.c_void_p)
