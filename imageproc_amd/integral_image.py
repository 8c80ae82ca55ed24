"""imageproc's integral images (src/integral_image.rs:50-152) for GrayImages,
computed on the GPU: (h + 1) x (w + 1) u32 tables, sums modulo 2^32."""
from __future__ import annotations

import numpy as np

from ._lib import check, lib, ptr


def _gray(image) -> np.ndarray:
    img = np.ascontiguousarray(image)
    if img.ndim != 2 or img.dtype != np.uint8:
        raise TypeError("expected a 2-D uint8 GrayImage array")
    return img


def integral_image(image) -> np.ndarray:
    """integral_image::<_, u32>, integral_image.rs:50-57."""
    img = _gray(image)
    h, w = img.shape
    out = np.empty((h + 1, w + 1), np.uint32)
    check(lib().iph_integral_image(ptr(img), w, h, 0, ptr(out)))
    return out


def integral_squared_image(image) -> np.ndarray:
    """integral_squared_image::<_, u32>, integral_image.rs:88-95."""
    img = _gray(image)
    h, w = img.shape
    out = np.empty((h + 1, w + 1), np.uint32)
    check(lib().iph_integral_image(ptr(img), w, h, 1, ptr(out)))
    return out


def sum_image_pixels(integral, left: int, top: int, right: int, bottom: int) -> int:
    """integral_image.rs:250-269: the sum of the pixels in [left, right] x
    [top, bottom] from four corners of an (h + 1, w + 1) u32 table, in
    wrapping u32 arithmetic like the reference's."""
    t = np.asarray(integral)
    a = int(t[bottom + 1, right + 1])
    b = int(t[top, left])
    c = int(t[top, right + 1])
    d = int(t[bottom + 1, left])
    return (a + b - c - d) % (1 << 32)


def variance(integral, integral_squared, left: int, top: int, right: int, bottom: int) -> float:
    """integral_image.rs:303-316, in f64 like the reference."""
    n = float(right - left + 1) * float(bottom - top + 1)
    sum_sq = float(sum_image_pixels(integral_squared, left, top, right, bottom))
    s = float(sum_image_pixels(integral, left, top, right, bottom))
    return (sum_sq - s * s / n) / n
