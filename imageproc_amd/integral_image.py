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
