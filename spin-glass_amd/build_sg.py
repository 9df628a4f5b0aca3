"""Builds libsg (HIP kernels + C++ host runtime + C-ABI) for gfx950 and the
`csg` command-line driver, in-tree: spin-glass_amd/lib/libsg.so and
spin-glass_amd/bin/csg.

Usage: python spin-glass_amd/build_sg.py
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "build")
LIB = os.path.join(HERE, "lib", "libsg.so")
CSG = os.path.join(HERE, "bin", "csg")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for every HIP object; contraction into
# FMA is disabled so that host and device FP64 threshold arithmetic agree.
FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}"]

LIB_SOURCES = ["sg_kernels_resident.hip", "sg_kernels_split.hip", "sg_kernels.hip", "sg_runtime.cpp", "sg_engine.cpp",
               "sg_probe.hip", "sg_static.hip"]
# Per-source extra flags (none at present: every object takes the default
# machine scheduler).
SOURCE_FLAGS: dict[str, list[str]] = {}


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *SOURCE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", out]
    if src.endswith(".cpp"):
        cmd.insert(1, "-xhip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    os.makedirs(os.path.dirname(CSG), exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(tmp, LIB)
    csg_src = os.path.join(CSRC, "csg_main.cpp")
    if os.path.exists(csg_src):
        cmd = ["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", csg_src, "-o", CSG + ".tmp",
               f"-L{os.path.dirname(LIB)}", "-lsg", "-Wl,-rpath,$ORIGIN/../lib"]
        subprocess.run(cmd, check=True, capture_output=not verbose)
        os.replace(CSG + ".tmp", CSG)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
