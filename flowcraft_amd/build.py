"""Builds libfc_lbm.so for gfx950 in-tree (flowcraft_amd/lib/).

    python -m flowcraft_amd.build

Flags: -ffp-contract=off keeps every float operation a separate IEEE operation
(HIP's default would fuse a*b+c into FMA), so the FP32 kernels round exactly as
FlowCraft's LbmReference does; float division and sqrt stay correctly rounded
(HIP's default).  Packed-FP32 instruction selection is disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["--offload-arch=" + ARCH, "-std=c++20", "-O3", "-fPIC", "-shared", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(ROOT, "include")]

TARGETS = {
    "libfc_lbm.so": ["fc_lbm.hip"],
}


def build(verbose: bool = False) -> list[str]:
    os.makedirs(LIB_DIR, exist_ok=True)
    built = []
    for lib, srcs in TARGETS.items():
        out = os.path.join(LIB_DIR, lib)
        tmp = out + ".tmp"
        cmd = [HIPCC] + COMMON + ["-o", tmp] + [os.path.join(CSRC, s) for s in srcs]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed for {lib}")
        os.replace(tmp, out)
        built.append(out)
    return built


if __name__ == "__main__":
    for p in build(verbose=True):
        print("built", p)
