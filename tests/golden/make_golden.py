"""Writes tests/golden/regression_fixtures.npz from the reference's own
regression data (tests/data/*.png of imageproc 0.27.0).

Run once in a container that holds the reference tree:
    python tests/golden/make_golden.py /root/reference
Inputs are converted to GrayImage the way tests/regression.rs:59-87 does:
RGB(A) -> l = (r*2126 + g*7152 + b*722) / 10000 (integer), Luma8 as is.
Only decoded pixel arrays (inputs and expected outputs) and the SHA-256 of
each source PNG are stored.
"""
import hashlib
import os
import sys

import numpy as np
from PIL import Image


def from_dynamic_gray(path):
    im = Image.open(path)
    if im.mode in ("RGB", "RGBA"):
        a = np.asarray(im.convert("RGBA") if im.mode == "RGBA" else im, dtype=np.uint32)
        r, g, b = a[..., 0], a[..., 1], a[..., 2]
        return ((r * 2126 + g * 7152 + b * 722) // 10000).astype(np.uint8)
    if im.mode == "L":
        return np.asarray(im, dtype=np.uint8).copy()
    raise ValueError(f"unexpected mode {im.mode} in {path}")


def main(ref_root):
    data = os.path.join(ref_root, "tests", "data")
    items = {
        "zebra": "zebra.png",
        "robin": "robin.png",
        "elephant": "elephant.png",
        "zebra_gaussian_3": "truth/zebra_gaussian_3.png",
        "zebra_gaussian_10": "truth/zebra_gaussian_10.png",
        "robin_sharpen3x3": "truth/robin_sharpen3x3.png",
        "robin_sharpen_gaussian": "truth/robin_sharpen_gaussian.png",
        "elephant_gradients": "truth/elephant_gradients.png",
        "zebra_adaptive_threshold": "truth/zebra_adaptive_threshold.png",
        "zebra_canny": "truth/zebra_canny.png",
    }
    arrays, shas = {}, {}
    for key, rel in items.items():
        p = os.path.join(data, rel)
        arrays[key] = from_dynamic_gray(p)
        shas[key] = hashlib.sha256(open(p, "rb").read()).hexdigest()
    arrays["source_sha256"] = np.array([f"{k}:{rel}:{shas[k]}" for k, rel in items.items()])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "regression_fixtures.npz")
    np.savez_compressed(out, **arrays)
    for k, v in arrays.items():
        print(k, v.shape, v.dtype)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
