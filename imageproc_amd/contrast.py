"""imageproc's adaptive_threshold (src/contrast.rs:19-47) on the GPU."""
from __future__ import annotations

import numpy as np

from ._lib import check, lib, ptr


def adaptive_threshold(image, block_radius: int, delta: int) -> np.ndarray:
    """contrast.rs:19-47: 255 where a pixel is at least its clipped
    (2 r + 1)^2 block mean (integer division) minus delta, else 0. Panics if
    block_radius is 0."""
    img = np.ascontiguousarray(image)
    if img.ndim != 2 or img.dtype != np.uint8:
        raise TypeError("expected a 2-D uint8 GrayImage array")
    out = np.empty_like(img)
    h, w = img.shape
    check(lib().iph_adaptive_threshold(ptr(img), ptr(out), w, h, int(block_radius), int(delta)))
    return out
