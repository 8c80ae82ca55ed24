"""iph_dev_find_extremes_noalloc: find_extremes with its partial results in the
library's own device array (no allocation), through the C ABI. Its record must
equal iph_dev_find_extremes' (the workspace entry) byte for byte, on two calls
in a row, the second with a smaller image, and the oracle's answer."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _record(buf):
    from imageproc_amd._lib import Extremes
    raw = bytes(buf.cpu().numpy().tobytes())
    return Extremes.from_buffer_copy(raw)


def _as_tuple(r, dtype):
    if dtype == np.uint8:
        mx, mn = r.max_raw & 0xFF, r.min_raw & 0xFF
    else:
        mx = np.array([r.max_raw], np.uint32).view(np.float32)[0]
        mn = np.array([r.min_raw], np.uint32).view(np.float32)[0]
    return mx, mn, (r.max_x, r.max_y), (r.min_x, r.min_y)


def _both(img_np):
    import torch
    from imageproc_amd._lib import F32, U8, check, lib
    p = U8 if img_np.dtype == np.uint8 else F32
    h, w = img_np.shape
    dev = torch.from_numpy(img_np).cuda()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    out_ws = torch.zeros(C.sizeof(C.c_uint32) * 6, dtype=torch.uint8, device="cuda")
    out_na = torch.full_like(out_ws, 0xAB)
    ws = torch.empty(lib().iph_dev_find_extremes_workspace(), dtype=torch.uint8, device="cuda")
    check(lib().iph_dev_find_extremes(dev.data_ptr(), p, w, h, out_ws.data_ptr(), ws.data_ptr(),
                                      ws.numel(), stream))
    check(lib().iph_dev_find_extremes_noalloc(dev.data_ptr(), p, w, h, out_na.data_ptr(),
                                              stream))
    torch.cuda.synchronize()
    return _record(out_ws), _record(out_na), bytes(out_ws.cpu().numpy()), bytes(out_na.cpu().numpy())


@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_noalloc_equals_workspace_entry_on_two_calls(dtype):
    rng = np.random.default_rng(7)
    # the second call is smaller: fewer workgroups than the first wrote pairs for
    for h, w in [(2048, 2017), (37, 29), (1, 1)]:
        if dtype == np.uint8:
            img = rng.integers(0, 200, (h, w), dtype=np.uint8)
        else:
            img = rng.standard_normal((h, w)).astype(np.float32)
        print(f"noalloc find_extremes {dtype.__name__} {w}x{h}", flush=True)
        r_ws, r_na, b_ws, b_na = _both(img)
        assert b_ws == b_na, (h, w)
        mx, mn, mxl, mnl = O.find_extremes(img)
        got = _as_tuple(r_na, dtype)
        assert (got[0], got[1]) == (mx, mn) and (got[2], got[3]) == (mxl, mnl), (h, w)


def test_noalloc_ties_and_nan_rules():
    rng = np.random.default_rng(11)
    big = rng.integers(0, 5, (1500, 1600)).astype(np.float32)
    small = np.array([[np.nan, 1.0], [2.0, -1.0]], np.float32)
    for img in (big, small):
        _, _, b_ws, b_na = _both(img)
        assert b_ws == b_na


def test_noalloc_rejects_bad_arguments():
    import torch
    from imageproc_amd._lib import EINVAL, EUNSUPPORTED, U8, U16, lib
    dev = torch.zeros(16, dtype=torch.uint8, device="cuda")
    out = torch.zeros(24, dtype=torch.uint8, device="cuda")
    assert lib().iph_dev_find_extremes_noalloc(dev.data_ptr(), U8, 0, 4, out.data_ptr(),
                                               None) == EINVAL
    assert b"non-empty" in lib().iph_last_error()
    assert lib().iph_dev_find_extremes_noalloc(dev.data_ptr(), U16, 4, 4, out.data_ptr(),
                                               None) == EUNSUPPORTED
    assert lib().iph_dev_find_extremes_noalloc(dev.data_ptr(), U8, 65536, 65536,
                                               out.data_ptr(), None) == EINVAL
    assert b"2^32 - 1" in lib().iph_last_error()
