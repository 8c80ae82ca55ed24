"""imageproc's canny (src/edges.rs:30-155) on the GPU."""
from __future__ import annotations

import numpy as np

from ._lib import check, lib, ptr


def canny(image, low_threshold: float, high_threshold: float) -> np.ndarray:
    """edges.rs:30-52: Gaussian blur (sigma 1.4), Sobel gradients,
    non-maximum suppression and hysteresis; 255 on edges, 0 elsewhere.
    Raises Panic where the reference panics: high < low, an empty image, or
    an edge that reaches the image border (low_threshold <= 0)."""
    img = np.ascontiguousarray(image)
    if img.ndim != 2 or img.dtype != np.uint8:
        raise TypeError("expected a 2-D uint8 GrayImage array")
    out = np.empty_like(img)
    h, w = img.shape
    check(lib().iph_canny(ptr(img), ptr(out), w, h, float(low_threshold), float(high_threshold)))
    return out
