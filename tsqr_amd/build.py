"""Builds libtsqr_amd.so in-tree with hipcc for gfx950.

Usage: python -m tsqr_amd.build [--verbose]

Each translation unit under csrc/ is compiled for gfx950 only, in parallel,
and the objects are linked into tsqr_amd/libtsqr_amd.so.  Packed FP32 VALU
instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtsqr_amd.so")
# the slowest first: every file compiles in a process of its own
SOURCES = [os.path.join(CSRC, f) for f in ("tsqr_w32s.hip", "tsqr_w32d.hip", "batch_f32.hip", "batch_f64.hip",
                                           "tsqr_w16s.hip", "tsqr_w16d.hip", "tsqr_w8s.hip", "tsqr_w8d.hip",
                                           "tsqr_f32.hip", "tsqr_f64.hip", "tsqr_amd.hip", "batch.hip")]
OBJDIR = os.path.join(HERE, "_build")
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with status {res.returncode}: {' '.join(cmd)}")


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    cc = hipcc()
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
        objs.append(obj)
        cmds.append([cc, *COMMON, "-c", "-o", obj, src])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    build_harness(verbose)
    return LIB


def build_harness(verbose: bool = False) -> str:
    """Builds tests/dropin/ref_harness: a caller written like the reference's
    test_tsqr.cu / test_qr.cu against include/tsqr_compat.hpp, linked with
    -ltsqr_amd (host code only)."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "dropin", "ref_harness.cpp")
    out = os.path.join(root, "tests", "dropin", "ref_harness")
    tmp = out + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(root, "include"),
          src, "-L", HERE, "-ltsqr_amd", "-Wl,-rpath,$ORIGIN/../../tsqr_amd", "-o", tmp], verbose)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
