"""CPU oracle for imageproc's sliding-window correlation path.

TEST INFRASTRUCTURE ONLY: this package is the checker. It is imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the CPU-baseline leg of
``bench.py``; the product package ``imageproc_amd`` never imports it.

numpy wrappers over ``liboracle.so`` (built from ``oracle.c`` by
``oracle/Makefile``), a plain-C restatement of the reference's loops. Each C
function cites the imageproc 0.27.0 file:line it restates. Parity status of
every function is pinned by ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

U8, U16, I16, I32, F32 = 0, 1, 2, 3, 4
SSE, SSE_NORMED, CCORR, CCORR_NORMED = 0, 1, 2, 3

_DT = {U8: np.uint8, U16: np.uint16, I16: np.int16, I32: np.int32, F32: np.float32}
_CODE = {np.dtype(np.uint8): U8, np.dtype(np.uint16): U16, np.dtype(np.int16): I16,
         np.dtype(np.int32): I32, np.dtype(np.float32): F32}


def build() -> str:
    """Compile liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def gaussian_kernel_f32(sigma: float) -> np.ndarray:
    n = lib().or_gaussian_kernel_f32(C.c_float(sigma), None, 0)
    out = np.zeros(max(n, 0), np.float32)
    lib().or_gaussian_kernel_f32(C.c_float(sigma), _p(out), n)
    return out


def _img(a):
    return np.ascontiguousarray(a)


def horizontal_filter(img, kernel):
    img = _img(img)
    k = _img(kernel)
    out = np.empty_like(img)
    h, w = img.shape
    rc = lib().or_horizontal_filter(_p(img), _p(out), w, h, _CODE[img.dtype], _p(k),
                                    len(k), _CODE[k.dtype])
    assert rc == 0, rc
    return out


def vertical_filter(img, kernel):
    img = _img(img)
    k = _img(kernel)
    out = np.empty_like(img)
    h, w = img.shape
    rc = lib().or_vertical_filter(_p(img), _p(out), w, h, _CODE[img.dtype], _p(k),
                                  len(k), _CODE[k.dtype])
    assert rc == 0, rc
    return out


def separable_filter(img, hk, vk):
    img, hk, vk = _img(img), _img(hk), _img(vk)
    out = np.empty_like(img)
    h, w = img.shape
    rc = lib().or_separable_filter(_p(img), _p(out), w, h, _CODE[img.dtype], _p(hk), len(hk),
                                   _p(vk), len(vk), _CODE[hk.dtype])
    assert rc == 0, rc
    return out


def gaussian_blur_f32(img, sigma):
    img = _img(img)
    out = np.empty_like(img)
    h, w = img.shape
    rc = lib().or_gaussian_blur_f32(_p(img), _p(out), w, h, _CODE[img.dtype], C.c_float(sigma))
    assert rc == 0, rc
    return out


def filter_clamped(img, kernel, kw, kh, out_dtype, nthreads=0):
    img = _img(img)
    k = _img(kernel)
    h, w = img.shape
    out = np.empty((h, w), out_dtype)
    if nthreads:
        rc = lib().or_filter_clamped_parallel(_p(img), _CODE[img.dtype], w, h, _p(k),
                                              _CODE[k.dtype], kw, kh, _p(out),
                                              _CODE[np.dtype(out_dtype)], nthreads)
    else:
        rc = lib().or_filter_clamped(_p(img), _CODE[img.dtype], w, h, _p(k), _CODE[k.dtype],
                                     kw, kh, _p(out), _CODE[np.dtype(out_dtype)])
    assert rc == 0, rc
    return out


def box_filter(img, rx, ry):
    img = _img(img).astype(np.uint8, copy=False)
    out = np.zeros_like(img)
    h, w = img.shape
    lib().or_box_filter(_p(img), _p(out), w, h, rx, ry)
    return out


def match_template(img, tpl, method, nthreads=0):
    img, tpl = _img(img), _img(tpl)
    h, w = img.shape
    th, tw = tpl.shape
    out = np.empty((h - th + 1, w - tw + 1), np.float32)
    if img.dtype == np.uint8:
        if nthreads:
            rc = lib().or_match_template_u8_parallel(_p(img), w, h, _p(tpl), tw, th, method,
                                                     _p(out), nthreads)
        else:
            rc = lib().or_match_template_u8(_p(img), w, h, _p(tpl), tw, th, method, _p(out))
    else:
        if nthreads:
            rc = lib().or_match_template_f32_parallel(_p(img), w, h, _p(tpl), tw, th, method,
                                                      _p(out), nthreads)
        else:
            rc = lib().or_match_template_f32(_p(img), w, h, _p(tpl), tw, th, method, _p(out))
    assert rc == 0, rc
    return out


def match_template_f64(img, tpl, method):
    img, tpl = _img(img), _img(tpl)
    h, w = img.shape
    th, tw = tpl.shape
    out = np.empty((h - th + 1, w - tw + 1), np.float64)
    fn = lib().or_match_template_u8_f64 if img.dtype == np.uint8 else lib().or_match_template_f32_f64
    rc = fn(_p(img), w, h, _p(tpl), tw, th, method, _p(out))
    assert rc == 0, rc
    return out


def match_template_with_mask(img, tpl, mask, method):
    img, tpl, mask = _img(img), _img(tpl), _img(mask)
    h, w = img.shape
    th, tw = tpl.shape
    mh, mw = mask.shape
    out = np.empty((h - th + 1, w - tw + 1), np.float32)
    rc = lib().or_match_template_with_mask_u8(_p(img), w, h, _p(tpl), tw, th, _p(mask), mw, mh,
                                              method, _p(out))
    assert rc == 0, rc
    return out


def find_extremes(img):
    """Returns (max_value, min_value, (max_x, max_y), (min_x, min_y))."""
    img = _img(img)
    h, w = img.shape
    locs = np.zeros(4, np.uint32)
    if img.dtype == np.uint8:
        vals = np.zeros(2, np.uint8)
        rc = lib().or_find_extremes_u8(_p(img), w, h, _p(vals), _p(locs))
    else:
        vals = np.zeros(2, np.float32)
        rc = lib().or_find_extremes_f32(_p(img), w, h, _p(vals), _p(locs))
    assert rc == 0, rc
    return vals[0], vals[1], (int(locs[0]), int(locs[1])), (int(locs[2]), int(locs[3]))


def integral_image(img, squared=False):
    img = _img(img)
    h, w = img.shape
    out = np.zeros((h + 1, w + 1), np.uint32)
    lib().or_integral_image_u32(_p(img), w, h, _p(out), int(squared))
    return out


def adaptive_threshold(img, block_radius, delta):
    img = _img(img)
    h, w = img.shape
    out = np.zeros((h, w), np.uint8)
    rc = lib().or_adaptive_threshold(_p(img), w, h, C.c_uint32(block_radius),
                                     C.c_int32(delta), _p(out))
    if rc == -1:
        raise ValueError("assertion failed: block_radius > 0")
    assert rc == 0
    return out


def canny_direction(gx: int, gy: int) -> int:
    return lib().or_canny_direction(int(gx), int(gy))


def canny_direction_grid(r: int):
    """(classes, magnitudes) of every (gx, gy) in [-r, r]^2, indexed [gy + r, gx + r]."""
    n = 2 * r + 1
    cls = np.zeros((n, n), np.uint8)
    mag = np.zeros((n, n), np.float32)
    lib().or_canny_direction_grid(int(r), _p(cls), _p(mag))
    return cls, mag


def canny(img, low, high):
    img = _img(img)
    h, w = img.shape
    out = np.zeros((h, w), np.uint8)
    rc = lib().or_canny(_p(img), w, h, C.c_float(low), C.c_float(high), _p(out))
    if rc == -1:
        raise ValueError("assertion failed: high_threshold >= low_threshold")
    if rc == -4:
        raise IndexError("hysteresis reached the image border")
    assert rc == 0, rc
    return out


def canny_thinned(img):
    img = _img(img)
    h, w = img.shape
    thin = np.zeros((h, w), np.float32)
    assert lib().or_canny_thinned(_p(img), w, h, _p(thin)) == 0
    return thin


def canny_hysteresis(thin, low, high):
    thin = np.ascontiguousarray(thin, np.float32)
    h, w = thin.shape
    out = np.zeros((h, w), np.uint8)
    stack = np.zeros(2 * h * w, np.uint32)
    rc = lib().or_canny_hysteresis(_p(thin), w, h, C.c_float(low), C.c_float(high), _p(stack),
                                   _p(out))
    if rc == -4:
        raise IndexError("hysteresis reached the image border")
    assert rc == 0, rc
    return out
