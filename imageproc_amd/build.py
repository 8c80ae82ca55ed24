"""Build step for the gfx950 library: compiles imageproc_amd/csrc/*.hip into
imageproc_amd/libimageproc_hip.so with hipcc (cross-compiles without a GPU).
Each source is compiled to its own object in parallel, then linked.

    python -m imageproc_amd.build
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libimageproc_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    # Parity kernels: never contract a*b+c into an FMA behind the source's back.
    "-ffp-contract=off",
    # IEEE-correct f32 sqrt and division on the device (the reference's
    # normalised scores are one correctly rounded sqrt and one division).
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # Packed FP32 instructions stay off: ordinary float code only.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
    return r.returncode


def build(verbose: bool = False, jobs: int | None = None) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    inc = ["-I", os.path.join(ROOT, "include")]
    tmpdir = tempfile.mkdtemp(prefix="iph_build_")
    try:
        objs = [os.path.join(tmpdir, os.path.basename(s) + ".o") for s in srcs]
        cmds = [[hipcc, *HIPCC_FLAGS, *inc, "-c", "-o", o, s] for s, o in zip(srcs, objs)]
        n = jobs or min(len(cmds), max(1, min(os.cpu_count() or 1, 16)))
        with ThreadPoolExecutor(max_workers=n) as ex:
            codes = list(ex.map(lambda c: _run(c, verbose), cmds))
        if any(codes):
            bad = [os.path.basename(s) for s, c in zip(srcs, codes) if c]
            raise RuntimeError(f"hipcc failed on {', '.join(bad)}")
        tmp = LIB + ".tmp"
        if _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose):
            raise RuntimeError("hipcc failed linking libimageproc_hip.so")
        os.replace(tmp, LIB)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
