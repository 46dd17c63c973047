"""Build libhipclaw.so (the C-ABI library) in-tree for gfx950 with hipcc.

    python -m forestclaw_amd.build

Each translation unit gets its own floating-point flags: the FP64 step and
the ghost / time-sync kernels are built with -ffp-contract=off so that they
round like the reference's FP64 Fortran; the FP32 step contracts a*b + c
within one source expression (-ffp-contract=on), not across statements, so
that every instantiation of a row's arithmetic (whole-patch or banded, one or
two columns per lane) rounds the same way.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libhipclaw.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(PKG, "..", "include")]

# per-source extra flags
STRICT = ["-ffp-contract=off"]
FAST = ["-ffp-contract=on"]
SOURCES = {
    "hc_api.hip": STRICT,
    "hc_step2_fp64.hip": STRICT,
    "hc_step2_fp32.hip": FAST,
    "hc_amr_kernels.hip": STRICT,
    "hc_timesync.hip": STRICT,
    "hc_batch.hip": STRICT,
    "hc_cnn.hip": STRICT,
    "hc_ctx.hip": STRICT,
    "hc_forest.cpp": STRICT,
    "hc_plan.cpp": STRICT,
    "hc_output.cpp": STRICT,
    "hc_source.hip": STRICT,
}


def _compile(src: str, extra: list[str]) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, SOURCES[s]), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
