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
OBJ_DIR = os.path.join(LIB_DIR, "obj")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMPILE = ["--offload-arch=" + ARCH, "-std=c++20", "-O3", "-fPIC", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
           "-I", os.path.join(ROOT, "include")]
COMMON = COMPILE + ["-shared"]

# Shared libraries: name -> list of (source, extra defines) translation units, compiled
# in parallel and linked together.  The step kernel's instantiations are split per
# lattice and storage format (step_inst.hip), the rest of the engine is fc_lbm.hip.
_STEP_UNITS = [("step_inst.hip", [f"-DFC_Q={q}", f"-DFC_STORE={st}"])
               for q in (19, 15) for st in ("StoreF32", "StoreF16", "StoreBF16", "StoreF64", "StoreF16SR",
                                            "StoreBF16SR")]
TARGETS = {
    "libfc_lbm.so": [("fc_lbm.hip", []), ("voxelize.hip", []), ("smoke.hip", [])] + _STEP_UNITS,
    "libfc_nn.so": [("nn_kernels.hip", [])],
}
# Standalone programs (not linked into the libraries): the copy kernel that calibrates
# rocprofv3's traffic counters.
EXECUTABLES = {
    "hbm_copy": ["hbm_copy.hip"],
}


def build_dropin_test(verbose: bool = False) -> str:
    """Compiles tests/cpp/dropin_test.cpp against include/flowcraft/LbmSolver.h, linking
    libfc_lbm.so and the oracle (a test binary; it needs a GPU to run)."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_test.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "dropin_test")
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include", "flowcraft"), "-o", out, src,
           "-L", LIB_DIR, "-lfc_lbm", "-L", os.path.join(ROOT, "oracle"), "-llbm_oracle",
           "-Wl,-rpath," + LIB_DIR + ":" + os.path.join(ROOT, "oracle"), "-Wl,--disable-new-dtags"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed for the drop-in test")
    return out


def _unit_object(lib, src, defines):
    tag = "".join(d.split("=")[-1].lower() for d in defines)
    return os.path.join(OBJ_DIR, f"{lib.split('.')[0]}_{os.path.splitext(src)[0]}{('_' + tag) if tag else ''}.o")


def build(verbose: bool = False, jobs: int | None = None) -> list[str]:
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = jobs or max(1, min(8, os.cpu_count() or 1))
    units = [(lib, src, defs) for lib, us in TARGETS.items() for src, defs in us]
    pending = list(units)
    running = []
    failed = None
    while pending or running:
        while pending and len(running) < jobs and failed is None:
            lib, src, defs = pending.pop(0)
            obj = _unit_object(lib, src, defs)
            cmd = [HIPCC] + COMPILE + defs + ["-c", "-o", obj, os.path.join(CSRC, src)]
            if verbose:
                print(" ".join(cmd), flush=True)
            running.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True), obj))
        if failed is not None and not running:
            break
        proc, obj = running.pop(0)
        out, _ = proc.communicate()
        if proc.returncode != 0:
            sys.stderr.write(out)
            failed = obj
    if failed is not None:
        raise RuntimeError(f"hipcc failed for {failed}")
    built = []
    for lib, us in TARGETS.items():
        out = os.path.join(LIB_DIR, lib)
        tmp = out + ".tmp"
        cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + \
              [_unit_object(lib, src, defs) for src, defs in us]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"link failed for {lib}")
        os.replace(tmp, out)
        built.append(out)
    for exe, srcs in EXECUTABLES.items():
        out = os.path.join(LIB_DIR, exe)
        cmd = [HIPCC] + [c for c in COMMON if c not in ("-shared", "-fPIC")] + ["-o", out] + \
              [os.path.join(CSRC, s) for s in srcs]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed for {exe}")
        built.append(out)
    return built


if __name__ == "__main__":
    for p in build(verbose=True):
        print("built", p)
