"""Builds libaxr.so in-tree with hipcc for gfx950.

FP contraction is off and HIP's default correctly rounded fp32 division and
sqrt stay on: the raster arithmetic must be bit-identical to the IEEE
restatement of the reference kernels.  Packed-FP32 instruction selection is
disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["raster.hip", "shade.hip", "batch.hip", "post.hip", "ssr.hip", "check.hip", "axr_api.cpp", "batch_api.cpp"]
HEADERS = ["axr_common.h", "axr_tri.h", "axr_launch.h", "axr_pixel.h", "axr_bin.h", "axr_ctx.h"]
OUT = os.path.join(HERE, "libaxr.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AXR_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def build_native(verbose: bool = False) -> str:
    """Compiles every HIP source into axi_visualizer_amd/libaxr.so."""
    tmp = OUT + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-x", "hip",
           *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(line for line in (res.stdout + res.stderr).splitlines()
                     if "packed-fp32-ops" not in line)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{errs}")
    if verbose and errs.strip():
        print(errs)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_native(verbose=True))
