"""Builds libvkvg.so in-tree: gfx950 HIP kernels (hipcc) + the C host recorder (gcc).

Both compilers run with -ffp-contract=off so that no fused multiply-add changes a rounding
the geometry depends on; HIP code additionally disables packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvkvg.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VKVG_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-rdc", "-Wno-unused-result", f"-I{INC}", f"-I{CSRC}",
]
C_FLAGS = ["-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-Wall",
           "-Wno-unused-function", f"-I{INC}"]

HIP_SOURCES = ["vkh_geom.hip", "vkh_raster.hip", "vkh_engine.hip", "vkh_diff.hip"]
C_SOURCES = ["vkvg_api.c", "vkvg_png.c", "vkvg_record.c"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {os.path.basename(cmd[-1])}")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        jobs.append([HIPCC, *HIP_FLAGS, "-c", "-o", o, os.path.join(CSRC, s)])
        objs.append(o)
    for s in C_SOURCES:
        o = os.path.join(BUILD, s.replace(".c", ".o"))
        jobs.append(["gcc", *C_FLAGS, "-c", "-o", o, os.path.join(CSRC, s)])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
