"""imageproc's gradients module (src/gradients.rs) for GrayImages on the GPU:
the magnitudes in one kernel, the directional filters through filter_clamped."""
from __future__ import annotations

import numpy as np

from . import filter as _flt
from ._lib import check, lib, ptr
from .kernel import Kernel

# kernel.rs:67-85
SOBEL_HORIZONTAL_3X3 = Kernel([-1, 0, 1, -2, 0, 2, -1, 0, 1], 3, 3)
SOBEL_VERTICAL_3X3 = Kernel([-1, -2, -1, 0, 0, 0, 1, 2, 1], 3, 3)
SCHARR_HORIZONTAL_3X3 = Kernel([-3, 0, 3, -10, 0, 10, -3, 0, 3], 3, 3)
SCHARR_VERTICAL_3X3 = Kernel([-3, -10, -3, 0, 0, 0, 3, 10, 3], 3, 3)
PREWITT_HORIZONTAL_3X3 = Kernel([-1, 0, 1, -1, 0, 1, -1, 0, 1], 3, 3)
PREWITT_VERTICAL_3X3 = Kernel([-1, -1, -1, 0, 0, 0, 1, 1, 1], 3, 3)


def _i32(kernel: Kernel) -> np.ndarray:
    k = np.ascontiguousarray(kernel.data)
    if k.dtype != np.int32:
        raise TypeError("gradients take Kernel<i32>")
    return k


def gradients_grayscale(image, horizontal_kernel: Kernel, vertical_kernel: Kernel) -> np.ndarray:
    """gradients.rs:15-21 / 99-160: u16 magnitudes of the two i16-clamped
    filter responses."""
    img = np.ascontiguousarray(image)
    if img.ndim != 2 or img.dtype != np.uint8:
        raise TypeError("expected a 2-D uint8 GrayImage array")
    hk, vk = _i32(horizontal_kernel), _i32(vertical_kernel)
    out = np.empty(img.shape, np.uint16)
    h, w = img.shape
    check(lib().iph_gradients(ptr(img), w, h, ptr(hk), horizontal_kernel.width,
                              horizontal_kernel.height, ptr(vk), vertical_kernel.width,
                              vertical_kernel.height, ptr(out)))
    return out


def sobel_gradients(image) -> np.ndarray:
    """gradients.rs:198-204."""
    return gradients_grayscale(image, SOBEL_HORIZONTAL_3X3, SOBEL_VERTICAL_3X3)


def prewitt_gradients(image) -> np.ndarray:
    """gradients.rs:207-213."""
    return gradients_grayscale(image, PREWITT_HORIZONTAL_3X3, PREWITT_VERTICAL_3X3)


def horizontal_sobel(image) -> np.ndarray:
    """gradients.rs:163-166."""
    return _flt.filter_clamped(image, SOBEL_HORIZONTAL_3X3, np.int16)


def vertical_sobel(image) -> np.ndarray:
    """gradients.rs:169-172."""
    return _flt.filter_clamped(image, SOBEL_VERTICAL_3X3, np.int16)


def horizontal_scharr(image) -> np.ndarray:
    """gradients.rs:175-178."""
    return _flt.filter_clamped(image, SCHARR_HORIZONTAL_3X3, np.int16)


def vertical_scharr(image) -> np.ndarray:
    """gradients.rs:181-184."""
    return _flt.filter_clamped(image, SCHARR_VERTICAL_3X3, np.int16)


def horizontal_prewitt(image) -> np.ndarray:
    """gradients.rs:187-190."""
    return _flt.filter_clamped(image, PREWITT_HORIZONTAL_3X3, np.int16)


def vertical_prewitt(image) -> np.ndarray:
    """gradients.rs:193-196."""
    return _flt.filter_clamped(image, PREWITT_VERTICAL_3X3, np.int16)
