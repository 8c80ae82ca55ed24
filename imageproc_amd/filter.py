"""imageproc::filter on gfx950: the correlation filters of filter/mod.rs and
filter/box_filter.rs, computed by libimageproc_hip.so.

Images are numpy arrays of shape (height, width): uint8 for GrayImage,
float32 for Image<Luma<f32>>. Every function returns a new array, as the
reference returns a new Image. Panics of the reference raise
``imageproc_amd.Panic`` with the same message.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import F32, I32, U8, check, lib, ptr
from .kernel import LAPLACIAN_3X3, Kernel

__all__ = [
    "gaussian_blur_f32", "separable_filter", "separable_filter_equal", "horizontal_filter",
    "vertical_filter", "filter_clamped", "filter_clamped_parallel", "laplacian_filter",
    "laplacian_filter_parallel", "box_filter", "gaussian_kernel_f32",
]


def _image(image) -> tuple[np.ndarray, int]:
    a = np.ascontiguousarray(image)
    if a.ndim != 2:
        raise TypeError("images are 2-D (height, width) arrays of single-channel pixels")
    if a.dtype == np.uint8:
        return a, U8
    if a.dtype == np.float32:
        return a, F32
    raise TypeError(f"pixel type {a.dtype} is not Luma<u8> or Luma<f32>")


def _kernel_1d(k) -> tuple[np.ndarray, int]:
    a = np.asarray(k)
    if a.dtype.kind == "f":
        return np.ascontiguousarray(a, dtype=np.float32).reshape(-1), F32
    if a.dtype.kind in "iub":
        return np.ascontiguousarray(a, dtype=np.int32).reshape(-1), I32
    raise TypeError(f"kernel element type {a.dtype} is not f32 or i32")


def gaussian_kernel_f32(sigma: float) -> np.ndarray:
    """The normalised Gaussian taps (filter/mod.rs:131-145), computed on the host."""
    n = lib().iph_gaussian_kernel_f32(float(sigma), None, 0)
    out = np.zeros(max(n, 0), np.float32)
    lib().iph_gaussian_kernel_f32(float(sigma), ptr(out), n)
    return out


def gaussian_blur_f32(image, sigma: float) -> np.ndarray:
    """filter/mod.rs:156-164. Panics if sigma <= 0."""
    img, p = _image(image)
    out = np.empty_like(img)
    h, w = img.shape
    check(lib().iph_gaussian_blur_f32(ptr(img), ptr(out), w, h, p, float(sigma)))
    return out


def _gray_u8(image) -> np.ndarray:
    img = np.ascontiguousarray(image)
    if img.ndim != 2 or img.dtype != np.uint8:
        raise TypeError("expected a 2-D uint8 GrayImage array")
    return img


def sharpen3x3(image) -> np.ndarray:
    """filter/sharpen.rs:13-16: filter_clamped with identity minus the 3x3
    Laplacian (i32 kernel), u8 out."""
    img = _gray_u8(image)
    out = np.empty_like(img)
    h, w = img.shape
    check(lib().iph_sharpen3x3(ptr(img), ptr(out), w, h))
    return out


def sharpen3x3_parallel(image) -> np.ndarray:
    """filter/sharpen.rs:20-24: the same GPU path."""
    return sharpen3x3(image)


def sharpen_gaussian(image, sigma: float, amount: float) -> np.ndarray:
    """filter/sharpen.rs:31-38: output = clamp((1 + amount) * input - amount *
    gaussian_blur_f32(input as f32, sigma)). Panics if sigma <= 0."""
    img = _gray_u8(image)
    out = np.empty_like(img)
    h, w = img.shape
    check(lib().iph_sharpen_gaussian(ptr(img), ptr(out), w, h, float(sigma), float(amount)))
    return out


def separable_filter(image, h_kernel, v_kernel) -> np.ndarray:
    """filter/mod.rs:169-183: horizontal pass, then vertical pass on the
    pixel-typed intermediate."""
    img, p = _image(image)
    hk, kt = _kernel_1d(h_kernel)
    vk, kt2 = _kernel_1d(v_kernel)
    if kt != kt2:
        raise TypeError("both 1D kernels must have the same element type")
    out = np.empty_like(img)
    h, w = img.shape
    check(lib().iph_separable_filter(ptr(img), ptr(out), w, h, p, ptr(hk), hk.size, ptr(vk),
                                     vk.size, kt))
    return out


def separable_filter_equal(image, kernel) -> np.ndarray:
    """filter/mod.rs:188-195."""
    return separable_filter(image, kernel, kernel)


def horizontal_filter(image, kernel) -> np.ndarray:
    """filter/mod.rs:232-314."""
    img, p = _image(image)
    k, kt = _kernel_1d(kernel)
    out = np.empty_like(img)
    h, w = img.shape
    check(lib().iph_horizontal_filter(ptr(img), ptr(out), w, h, p, ptr(k), k.size, kt))
    return out


def vertical_filter(image, kernel) -> np.ndarray:
    """filter/mod.rs:319-405."""
    img, p = _image(image)
    k, kt = _kernel_1d(kernel)
    out = np.empty_like(img)
    h, w = img.shape
    check(lib().iph_vertical_filter(ptr(img), ptr(out), w, h, p, ptr(k), k.size, kt))
    return out


def filter_clamped(image, kernel: Kernel, out_dtype=None) -> np.ndarray:
    """filter/mod.rs:202-210: 2-D correlation at the kernel's type, clamped
    to ``out_dtype`` (the reference's S; defaults to the pixel type)."""
    img, p = _image(image)
    if not isinstance(kernel, Kernel):
        raise TypeError("kernel must be an imageproc_amd.Kernel")
    kt = F32 if kernel.data.dtype == np.float32 else I32
    st_dtype = np.dtype(out_dtype if out_dtype is not None else img.dtype)
    if st_dtype not in _lib.DTYPE_CODE:
        raise TypeError(f"output type {st_dtype} is not supported")
    out = np.empty(img.shape, st_dtype)
    h, w = img.shape
    check(lib().iph_filter_clamped(ptr(img), p, w, h, ptr(kernel.data), kt, kernel.width,
                                   kernel.height, ptr(out), _lib.DTYPE_CODE[st_dtype]))
    return out


def filter_clamped_parallel(image, kernel: Kernel, out_dtype=None) -> np.ndarray:
    """filter/mod.rs:214-226: the same GPU kernel as filter_clamped."""
    return filter_clamped(image, kernel, out_dtype)


def laplacian_filter(image) -> np.ndarray:
    """filter/mod.rs:440-442: filter_clamped with LAPLACIAN_3X3 to i16."""
    return filter_clamped(image, LAPLACIAN_3X3, np.int16)


def laplacian_filter_parallel(image) -> np.ndarray:
    """filter/mod.rs:444-449."""
    return laplacian_filter(image)


def box_filter(image, x_radius: int, y_radius: int) -> np.ndarray:
    """filter/box_filter.rs:19-70 (GrayImage only)."""
    img = np.ascontiguousarray(image)
    if img.dtype != np.uint8 or img.ndim != 2:
        raise TypeError("box_filter takes a GrayImage (2-D uint8 array)")
    out = np.zeros_like(img)
    h, w = img.shape
    check(lib().iph_box_filter(ptr(img), ptr(out), w, h, int(x_radius), int(y_radius)))
    return out
