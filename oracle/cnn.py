"""Float64 restatement of the proposal network's layer arithmetic (test
infrastructure only: the checker of the gfx950 kernels, never the measured or
shipped path). No counterpart in the reference.

Tensors are NHWC patches: x[p, y, x, c] with 16x16 pixels and 32 channels;
weights w[co, ky, kx, ci]; 3x3 correlation with zero padding of 1."""
import numpy as np


def bf16_round(a):
    """Round float32 values to the nearest bfloat16 (ties to even), as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def conv3x3(x, w, b=None):
    """y[p, y, x, co] = sum_{ky, kx, ci} xpad[p, y + ky, x + kx, ci] w[co, ky, kx, ci] + b[co]."""
    P, H, W, _ = x.shape
    xp = np.zeros((P, H + 2, W + 2, x.shape[3]))
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((P, H, W, w.shape[0]))
    for ky in range(3):
        for kx in range(3):
            y += np.einsum("phwc,oc->phwo", xp[:, ky:ky + H, kx:kx + W], w[:, ky, kx, :])
    if b is not None:
        y += b
    return y


def conv3x3_backward(x, w, dy):
    """Gradients of conv3x3 with respect to w, b and x, given dL/dy."""
    P, H, W, _ = x.shape
    xp = np.zeros((P, H + 2, W + 2, x.shape[3]))
    xp[:, 1:-1, 1:-1] = x
    dw = np.zeros_like(w)
    dxp = np.zeros_like(xp)
    for ky in range(3):
        for kx in range(3):
            dw[:, ky, kx, :] = np.einsum("phwo,phwc->oc", dy, xp[:, ky:ky + H, kx:kx + W])
            dxp[:, ky:ky + H, kx:kx + W] += np.einsum("phwo,oc->phwc", dy, w[:, ky, kx, :])
    return dw, dy.sum(axis=(0, 1, 2)), dxp[:, 1:-1, 1:-1]


def softmax_xent(logits, target):
    """Loss and gradient of softmax cross-entropy over a flat logit vector."""
    m = logits.max()
    e = np.exp(logits - m)
    s = e.sum()
    g = e / s
    g[target] -= 1.0
    return np.log(s) + m - logits[target], g
