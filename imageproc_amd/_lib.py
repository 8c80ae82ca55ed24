"""ctypes binding of libimageproc_hip.so (declared in include/imageproc_hip.h).

The library is loaded from this package directory. It is built by
``python -m imageproc_amd.build``; there is no other implementation to fall
back to, so a missing library raises at first use.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libimageproc_hip.so")

# Status codes, element types, methods and paths (include/imageproc_hip.h).
OK, EINVAL, ENOMEM, EDEVICE, EUNSUPPORTED = 0, 1, 2, 3, 4
U8, U16, I16, I32, F32 = 0, 1, 2, 3, 4
PATH_AUTO, PATH_MFMA, PATH_DIRECT = 0, 1, 2

DTYPE_CODE = {
    np.dtype(np.uint8): U8,
    np.dtype(np.uint16): U16,
    np.dtype(np.int16): I16,
    np.dtype(np.int32): I32,
    np.dtype(np.float32): F32,
}


class Panic(AssertionError):
    """A reference assertion failed (imageproc would panic with this message)."""


class DeviceError(RuntimeError):
    """The HIP runtime reported an error, or no gfx950 device is usable."""


class Extremes(C.Structure):
    """iph_extremes: find_extremes' result (template_matching.rs:463-475)."""

    _fields_ = [
        ("max_raw", C.c_uint32),
        ("min_raw", C.c_uint32),
        ("max_x", C.c_uint32),
        ("max_y", C.c_uint32),
        ("min_x", C.c_uint32),
        ("min_y", C.c_uint32),
    ]


_SIGS = {
    "iph_last_error": (C.c_char_p, []),
    "iph_device_info": (C.c_int32, [C.c_char_p, C.c_size_t]),
    "iph_gaussian_kernel_f32": (C.c_int32, [C.c_float, C.c_void_p, C.c_int32]),
    "iph_gaussian_blur_f32": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                          C.c_int32, C.c_float]),
    "iph_separable_filter": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                         C.c_int32, C.c_void_p, C.c_uint32, C.c_void_p,
                                         C.c_uint32, C.c_int32]),
    "iph_horizontal_filter": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                          C.c_int32, C.c_void_p, C.c_uint32, C.c_int32]),
    "iph_vertical_filter": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                        C.c_int32, C.c_void_p, C.c_uint32, C.c_int32]),
    "iph_filter_clamped": (C.c_int32, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32,
                                       C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32,
                                       C.c_void_p, C.c_int32]),
    "iph_box_filter": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.c_uint32]),
    "iph_match_template": (C.c_int32, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32,
                                       C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32,
                                       C.c_int32, C.c_void_p]),
    "iph_find_extremes": (C.c_int32, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32,
                                      C.c_void_p]),
    "iph_dev_gaussian_blur_workspace": (C.c_size_t, [C.c_uint32, C.c_uint32, C.c_uint32,
                                                     C.c_int32, C.c_float]),
    "iph_dev_gaussian_blur_f32": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                              C.c_uint32, C.c_size_t, C.c_int32, C.c_float,
                                              C.c_void_p, C.c_size_t, C.c_void_p]),
    "iph_dev_separable_filter_workspace": (C.c_size_t, [C.c_uint32, C.c_uint32, C.c_uint32,
                                                        C.c_int32, C.c_uint32, C.c_int32]),
    "iph_dev_separable_filter": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                             C.c_uint32, C.c_size_t, C.c_int32, C.c_void_p,
                                             C.c_uint32, C.c_void_p, C.c_uint32, C.c_int32,
                                             C.c_void_p, C.c_size_t, C.c_void_p]),
    "iph_dev_filter_clamped_workspace": (C.c_size_t, [C.c_int32, C.c_uint32, C.c_uint32]),
    "iph_dev_filter_clamped": (C.c_int32, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32,
                                           C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32,
                                           C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t,
                                           C.c_void_p]),
    "iph_dev_box_filter_workspace": (C.c_size_t, [C.c_uint32, C.c_uint32]),
    "iph_dev_box_filter": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                       C.c_uint32, C.c_uint32, C.c_void_p, C.c_size_t,
                                       C.c_void_p]),
    "iph_integral_image": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32,
                                       C.c_void_p]),
    "iph_adaptive_threshold": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                           C.c_uint32, C.c_int32]),
    "iph_dev_integral_image_workspace": (C.c_size_t, [C.c_uint32, C.c_uint32]),
    "iph_dev_integral_image": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32,
                                           C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "iph_dev_adaptive_threshold_workspace": (C.c_size_t, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "iph_dev_adaptive_threshold": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                               C.c_uint32, C.c_int32, C.c_void_p, C.c_size_t,
                                               C.c_void_p]),
    "iph_canny": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float,
                              C.c_float]),
    "iph_dev_canny_workspace": (C.c_size_t, [C.c_uint32, C.c_uint32]),
    "iph_dev_canny": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float,
                                  C.c_float, C.c_void_p, C.c_size_t, C.c_void_p]),
    "iph_dev_gradients": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32,
                                      C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                      C.c_void_p]),
    "iph_dev_sharpen_gaussian_workspace": (C.c_size_t, [C.c_uint32, C.c_uint32, C.c_float]),
    "iph_dev_sharpen_gaussian": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                             C.c_float, C.c_float, C.c_void_p, C.c_size_t,
                                             C.c_void_p]),
    "iph_gradients": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32,
                                  C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    "iph_sharpen3x3": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]),
    "iph_sharpen_gaussian": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                         C.c_float, C.c_float]),
    "iph_match_template_with_mask": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32,
                                                 C.c_void_p, C.c_uint32, C.c_uint32,
                                                 C.c_void_p, C.c_uint32, C.c_uint32,
                                                 C.c_int32, C.c_int32, C.c_void_p]),
    "iph_dev_match_template_with_mask_workspace": (C.c_size_t, [C.c_uint32, C.c_uint32,
                                                               C.c_uint32, C.c_uint32,
                                                               C.c_int32, C.c_int32]),
    "iph_dev_match_template_with_mask": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32,
                                                     C.c_void_p, C.c_uint32, C.c_uint32,
                                                     C.c_void_p, C.c_uint32, C.c_uint32,
                                                     C.c_int32, C.c_int32, C.c_void_p,
                                                     C.c_void_p, C.c_size_t, C.c_void_p]),
    "iph_dev_match_template_workspace": (C.c_size_t, [C.c_int32, C.c_uint32, C.c_uint32,
                                                      C.c_uint32, C.c_uint32, C.c_int32,
                                                      C.c_int32]),
    "iph_dev_match_template": (C.c_int32, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32,
                                           C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32,
                                           C.c_int32, C.c_void_p, C.c_void_p, C.c_size_t,
                                           C.c_void_p]),
    "iph_batch_create": (C.c_int32, [C.POINTER(C.c_void_p)]),
    "iph_batch_destroy": (None, [C.c_void_p]),
    "iph_batch_add_gaussian_blur": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                                C.c_uint32, C.c_float]),
    "iph_batch_add_filter_clamped": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                                 C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32]),
    "iph_batch_add_match_template": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                                 C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32,
                                                 C.c_void_p]),
    "iph_batch_plan": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p]),
    "iph_batch_run": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "iph_batch_run_eager": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "iph_batch_tuning": (C.c_int32, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "iph_dev_cnn_conv3x3": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                        C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "iph_dev_cnn_conv3x3_dw_workspace": (C.c_size_t, [C.c_int32]),
    "iph_dev_cnn_conv3x3_dw": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                           C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    "iph_dev_cnn_flip_weights": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "iph_dev_cnn_relu_backward": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                              C.c_void_p]),
    "iph_dev_cnn_softmax_xent": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                             C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]),
    "iph_dev_cnn_adam": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                     C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32,
                                     C.c_void_p]),
    "iph_dev_cnn_to_bf16": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "iph_dev_downsample4_u8": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                           C.c_void_p]),
    "iph_dev_proposal_candidates_workspace": (C.c_size_t, [C.c_uint32, C.c_uint32]),
    "iph_dev_proposal_candidates": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32,
                                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                                C.c_void_p]),
    "iph_dev_proposal_features": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                              C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                              C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "iph_dev_proposal_top_workspace": (C.c_size_t, [C.c_int32, C.c_int32]),
    "iph_dev_proposal_top": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint32,
                                         C.c_uint32, C.c_int32, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_size_t, C.c_void_p]),
    "iph_dev_proposal_refine_workspace": (C.c_size_t, [C.c_int32]),
    "iph_dev_proposal_refine": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                            C.c_uint32, C.c_uint32, C.c_int32, C.c_void_p,
                                            C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                            C.c_void_p, C.c_size_t, C.c_void_p]),
    "iph_dev_read_result": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "iph_dev_proposal_search_workspace": (C.c_size_t, [C.c_uint32, C.c_uint32, C.c_uint32,
                                                       C.c_uint32, C.c_int32, C.c_int32,
                                                       C.c_int32]),
    "iph_dev_proposal_search": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                            C.c_uint32, C.c_uint32, C.c_int32, C.c_void_p,
                                            C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                            C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "iph_dev_find_extremes_workspace": (C.c_size_t, []),
    "iph_dev_find_extremes": (C.c_int32, [C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32,
                                          C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "iph_dev_find_extremes_noalloc": (C.c_int32, [C.c_void_p, C.c_int32, C.c_uint32,
                                                  C.c_uint32, C.c_void_p, C.c_void_p]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def lib() -> C.CDLL:
    """Loads libimageproc_hip.so once and declares every entry point."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `python -m imageproc_amd.build` "
                "(imageproc_amd has no CPU implementation)")
        # torch ships its own libamdhip64.so.7. Loading it first makes this
        # library bind to the same HIP runtime (same soname), so one process
        # never holds two runtimes that would contend for the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    """Raises the Python counterpart of a non-zero iph_status."""
    if status == OK:
        return
    msg = lib().iph_last_error().decode()
    if status == EINVAL:
        raise Panic(msg)
    if status == EUNSUPPORTED:
        raise TypeError(msg)
    if status == ENOMEM:
        raise MemoryError(msg)
    raise DeviceError(msg)


def ptr(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


def device_info() -> str:
    buf = C.create_string_buffer(256)
    check(lib().iph_device_info(buf, len(buf)))
    return buf.value.decode()
