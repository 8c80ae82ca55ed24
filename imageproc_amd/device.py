"""Device-resident entry points: the iph_dev_* functions over torch tensors.

torch is used only to own device memory and to name the current HIP stream;
every computation is a kernel of libimageproc_hip.so. Workspaces are cached
per (function, size) so repeated calls do not allocate.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import torch

from ._lib import F32, I16, I32, U8, U16, Extremes, check, lib
from .template_matching import MatchTemplateMethod, _PATHS

_TORCH_CODE = {torch.uint8: U8, torch.int16: I16, torch.int32: I32, torch.float32: F32,
               torch.uint16: U16}
_ws_cache: dict = {}


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _workspace(key, nbytes: int, device) -> torch.Tensor:
    t = _ws_cache.get((key, device))
    if t is None or t.numel() < nbytes:
        t = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
        _ws_cache[(key, device)] = t
    return t


def _code(t: torch.Tensor) -> int:
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    return _TORCH_CODE[t.dtype]


def gaussian_blur_f32(src: torch.Tensor, sigma: float, out: torch.Tensor | None = None):
    """Batched gaussian_blur_f32 over src of shape (n, h, w) or (h, w)."""
    p = _code(src)
    n, h, w = (1, *src.shape) if src.dim() == 2 else src.shape
    out = torch.empty_like(src) if out is None else out
    wsb = lib().iph_dev_gaussian_blur_workspace(w, h, n, p, float(sigma))
    ws = _workspace("blur", wsb, src.device)
    check(lib().iph_dev_gaussian_blur_f32(src.data_ptr(), out.data_ptr(), w, h, n, h * w, p,
                                          float(sigma), ws.data_ptr(), ws.numel(), _stream()))
    return out


def separable_filter(src: torch.Tensor, h_kernel, v_kernel, out: torch.Tensor | None = None):
    """Batched separable_filter; kernels are host sequences (f32 or i32)."""
    import numpy as np

    p = _code(src)
    n, h, w = (1, *src.shape) if src.dim() == 2 else src.shape
    hk = np.asarray(h_kernel)
    kt = F32 if hk.dtype.kind == "f" else I32
    dt = np.float32 if kt == F32 else np.int32
    hk = np.ascontiguousarray(hk, dtype=dt)
    vk = np.ascontiguousarray(np.asarray(v_kernel), dtype=dt)
    out = torch.empty_like(src) if out is None else out
    wsb = lib().iph_dev_separable_filter_workspace(w, h, n, p, hk.size, kt)
    ws = _workspace("sep", wsb, src.device)
    check(lib().iph_dev_separable_filter(src.data_ptr(), out.data_ptr(), w, h, n, h * w, p,
                                         hk.ctypes.data, hk.size, vk.ctypes.data, vk.size, kt,
                                         ws.data_ptr(), ws.numel(), _stream()))
    return out


def filter_clamped(src: torch.Tensor, kernel, out_dtype=None, out: torch.Tensor | None = None):
    p = _code(src)
    h, w = src.shape
    kt = F32 if kernel.data.dtype.kind == "f" else I32
    if out is None:
        out = torch.empty_like(src, dtype=out_dtype or src.dtype)
    wsb = lib().iph_dev_filter_clamped_workspace(kt, kernel.width, kernel.height)
    ws = _workspace("f2d", wsb, src.device)
    check(lib().iph_dev_filter_clamped(src.data_ptr(), p, w, h, kernel.data.ctypes.data, kt,
                                       kernel.width, kernel.height, out.data_ptr(), _code(out),
                                       ws.data_ptr(), ws.numel(), _stream()))
    return out


def box_filter(src: torch.Tensor, x_radius: int, y_radius: int, out=None):
    _code(src)
    h, w = src.shape
    out = torch.empty_like(src) if out is None else out
    wsb = lib().iph_dev_box_filter_workspace(w, h)
    ws = _workspace("box", wsb, src.device)
    check(lib().iph_dev_box_filter(src.data_ptr(), out.data_ptr(), w, h, x_radius, y_radius,
                                   ws.data_ptr(), ws.numel(), _stream()))
    return out


def match_template(image: torch.Tensor, template: torch.Tensor, method, path: str = "auto",
                   out: torch.Tensor | None = None):
    p = _code(image)
    h, w = image.shape
    th, tw = template.shape
    m = int(MatchTemplateMethod(method))
    if out is None:
        out = torch.empty((max(h - th + 1, 0), max(w - tw + 1, 0)), dtype=torch.float32,
                          device=image.device)
    wsb = lib().iph_dev_match_template_workspace(p, w, h, tw, th, m, _PATHS[path])
    ws = _workspace(("match", path), wsb, image.device)
    check(lib().iph_dev_match_template(image.data_ptr(), p, w, h, template.data_ptr(), tw, th,
                                       m, _PATHS[path], out.data_ptr(), ws.data_ptr(),
                                       ws.numel(), _stream()))
    return out


def match_template_with_mask(image: torch.Tensor, template: torch.Tensor, method,
                             mask: torch.Tensor, out: torch.Tensor | None = None,
                             path: str = "auto"):
    """match_template_with_mask on device uint8 tensors (template_matching.rs:125-140);
    path as in template_matching.match_template_with_mask."""
    for t in (image, template, mask):
        if _code(t) != _TORCH_CODE[torch.uint8]:
            raise TypeError("match_template_with_mask takes GrayImages (uint8)")
    h, w = image.shape
    th, tw = template.shape
    mh, mw = mask.shape
    if out is None:
        out = torch.empty((max(h - th + 1, 0), max(w - tw + 1, 0)), dtype=torch.float32,
                          device=image.device)
    m = int(MatchTemplateMethod(method))
    wsb = lib().iph_dev_match_template_with_mask_workspace(w, h, tw, th, m, _PATHS[path])
    ws = _workspace(("masked", path), wsb, image.device)
    check(lib().iph_dev_match_template_with_mask(
        image.data_ptr(), w, h, template.data_ptr(), tw, th, mask.data_ptr(), mw, mh, m,
        _PATHS[path], out.data_ptr(), ws.data_ptr(), ws.numel(), _stream()))
    return out


def integral_image(image: torch.Tensor, squared: bool = False,
                   out: torch.Tensor | None = None) -> torch.Tensor:
    """integral_image / integral_squared_image::<_, u32> (integral_image.rs:50-95)
    of a device uint8 tensor, as an (h + 1, w + 1) int32 tensor holding the u32
    bits."""
    if _code(image) != _TORCH_CODE[torch.uint8]:
        raise TypeError("integral_image takes GrayImages (uint8)")
    h, w = image.shape
    if out is None:
        out = torch.empty((h + 1, w + 1), dtype=torch.int32, device=image.device)
    wsb = lib().iph_dev_integral_image_workspace(w, h)
    ws = _workspace("integral", wsb, image.device)
    check(lib().iph_dev_integral_image(image.data_ptr(), w, h, int(bool(squared)),
                                       out.data_ptr(), ws.data_ptr(), ws.numel(), _stream()))
    return out


def adaptive_threshold(image: torch.Tensor, block_radius: int, delta: int,
                       out: torch.Tensor | None = None) -> torch.Tensor:
    """adaptive_threshold (contrast.rs:19-47) of a device uint8 tensor."""
    if _code(image) != _TORCH_CODE[torch.uint8]:
        raise TypeError("adaptive_threshold takes GrayImages (uint8)")
    h, w = image.shape
    if out is None:
        out = torch.empty_like(image)
    wsb = lib().iph_dev_adaptive_threshold_workspace(w, h, int(block_radius))
    ws = _workspace("adaptive", wsb, image.device)
    check(lib().iph_dev_adaptive_threshold(image.data_ptr(), out.data_ptr(), w, h,
                                           int(block_radius), int(delta), ws.data_ptr(),
                                           ws.numel(), _stream()))
    return out


def canny(image: torch.Tensor, low_threshold: float, high_threshold: float,
          out: torch.Tensor | None = None) -> torch.Tensor:
    """canny (edges.rs:30-155) of a device uint8 tensor. Blocks until the
    hysteresis has converged."""
    if _code(image) != _TORCH_CODE[torch.uint8]:
        raise TypeError("canny takes GrayImages (uint8)")
    h, w = image.shape
    if out is None:
        out = torch.empty_like(image)
    wsb = lib().iph_dev_canny_workspace(w, h)
    ws = _workspace("canny", wsb, image.device)
    check(lib().iph_dev_canny(image.data_ptr(), out.data_ptr(), w, h, float(low_threshold),
                              float(high_threshold), ws.data_ptr(), ws.numel(), _stream()))
    return out


def gradients(image: torch.Tensor, horizontal_kernel, vertical_kernel,
              out: torch.Tensor | None = None) -> torch.Tensor:
    """gradients_grayscale (gradients.rs:15-21, 99-160) of a device uint8
    tensor with Kernel<i32> taps: u16 magnitudes (an int16 tensor holding the
    u16 bits when `out` is not given)."""
    if _code(image) != _TORCH_CODE[torch.uint8]:
        raise TypeError("gradients takes GrayImages (uint8)")
    hk = np.ascontiguousarray(horizontal_kernel.data, np.int32)
    vk = np.ascontiguousarray(vertical_kernel.data, np.int32)
    h, w = image.shape
    if out is None:
        out = torch.empty((h, w), dtype=torch.int16, device=image.device)
    check(lib().iph_dev_gradients(image.data_ptr(), w, h, hk.ctypes.data, horizontal_kernel.width,
                                  horizontal_kernel.height, vk.ctypes.data, vertical_kernel.width,
                                  vertical_kernel.height, out.data_ptr(), _stream()))
    return out


def sharpen_gaussian(image: torch.Tensor, sigma: float, amount: float,
                     out: torch.Tensor | None = None) -> torch.Tensor:
    """sharpen_gaussian (filter/sharpen.rs:31-38) of a device uint8 tensor."""
    if _code(image) != _TORCH_CODE[torch.uint8]:
        raise TypeError("sharpen_gaussian takes GrayImages (uint8)")
    h, w = image.shape
    if out is None:
        out = torch.empty_like(image)
    wsb = lib().iph_dev_sharpen_gaussian_workspace(w, h, float(sigma))
    ws = _workspace("sharpen", wsb, image.device)
    check(lib().iph_dev_sharpen_gaussian(image.data_ptr(), out.data_ptr(), w, h, float(sigma),
                                         float(amount), ws.data_ptr(), ws.numel(), _stream()))
    return out


def find_extremes(image: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """Writes the 24-byte iph_extremes record into `out` (a device uint8 tensor)."""
    p = _code(image)
    h, w = image.shape
    if out is None:
        out = torch.empty(C.sizeof(Extremes), dtype=torch.uint8, device=image.device)
    wsb = lib().iph_dev_find_extremes_workspace()
    ws = _workspace("ext", wsb, image.device)
    check(lib().iph_dev_find_extremes(image.data_ptr(), p, w, h, out.data_ptr(), ws.data_ptr(),
                                      ws.numel(), _stream()))
    return out
