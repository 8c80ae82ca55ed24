"""Kernel<K>: a borrowed row-major 2-D kernel (imageproc kernel.rs:5-64)."""
from __future__ import annotations

import numpy as np

from ._lib import Panic


class Kernel:
    """Row-major kernel data with its width and height.

    Mirrors ``Kernel::new`` (kernel.rs:24-33): panics if a side is zero or if
    width * height (widened to 64 bits) differs from the data length. The
    element type is kept: float32 for Kernel<f32>, int32 for Kernel<i32>.
    """

    __slots__ = ("data", "width", "height")

    def __init__(self, data, width: int, height: int):
        if not (width > 0 and height > 0):
            raise Panic("width and height must be non-zero")
        arr = np.asarray(data)
        if arr.dtype.kind == "f":
            arr = arr.astype(np.float32)
        elif arr.dtype.kind in "iub":
            arr = arr.astype(np.int32)
        else:
            raise TypeError(f"unsupported kernel element type {arr.dtype}")
        arr = np.ascontiguousarray(arr.reshape(-1))
        if int(width) * int(height) != arr.size:
            raise Panic("assertion failed: width as u64 * height as u64 == data.len() as u64")
        self.data = arr
        self.width = int(width)
        self.height = int(height)

    def get(self, x: int, y: int):
        """Kernel::get, kernel.rs:41-46."""
        return self.data[y * self.width + x]


# kernel.rs:66-93
SOBEL_HORIZONTAL_3X3 = Kernel([-1, 0, 1, -2, 0, 2, -1, 0, 1], 3, 3)
SOBEL_VERTICAL_3X3 = Kernel([-1, -2, -1, 0, 0, 0, 1, 2, 1], 3, 3)
SCHARR_HORIZONTAL_3X3 = Kernel([-3, 0, 3, -10, 0, 10, -3, 0, 3], 3, 3)
SCHARR_VERTICAL_3X3 = Kernel([-3, -10, -3, 0, 0, 0, 3, 10, 3], 3, 3)
PREWITT_HORIZONTAL_3X3 = Kernel([-1, 0, 1, -1, 0, 1, -1, 0, 1], 3, 3)
PREWITT_VERTICAL_3X3 = Kernel([-1, -1, -1, 0, 0, 0, 1, 1, 1], 3, 3)
ROBERTS_HORIZONTAL_2X2 = Kernel([1, 0, 0, -1], 2, 2)
ROBERTS_VERTICAL_2X2 = Kernel([0, 1, -1, 0], 2, 2)
# Kernel<i16> in the reference. Its sums stay within +-1020, so i32
# accumulation with an i16 clamp gives the same pixels as i16 accumulation.
LAPLACIAN_3X3 = Kernel([0, 1, 0, 1, -4, 1, 0, 1, 0], 3, 3)
