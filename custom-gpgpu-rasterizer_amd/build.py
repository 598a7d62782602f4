"""Build the native libraries in-tree (custom-gpgpu-rasterizer_amd/lib/).

  librast.so         HIP kernels for gfx950 + the C ABI of include/rast.h
  librast_scenes.so  host-only deterministic scene generators (include/rast_scene.h)
  librast_io.so      host-only OBJ reader, PNG codec (system zlib) and BMP writer (include/rast_io.h)

Every call recompiles.  Float code is built with -ffp-contract=off (no FMA
contraction) and without fast-math, so device arithmetic matches the CPU
restatement operation for operation; packed-FP32 instruction selection is
disabled for the device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-Wall"]

RAST_SOURCES = ["rast_kernels.hip", "rast_api.hip", "rast_mlp.hip"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def build_scenes() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "librast_scenes.so")
    tmp = out + ".tmp"
    _run([CXX, *HOST_FLAGS, "-I", INCLUDE, "-o", tmp, os.path.join(CSRC, "scenes.cpp")])
    os.replace(tmp, out)
    return out


def build_io() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "librast_io.so")
    tmp = out + ".tmp"
    _run([CXX, *HOST_FLAGS, "-I", INCLUDE, "-o", tmp, os.path.join(CSRC, "rast_io.cpp"), "-lz"])
    os.replace(tmp, out)
    return out


def build_rast() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "librast.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in RAST_SOURCES]
    _run([HIPCC, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-o", tmp, *srcs])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found")
    build_scenes()
    build_io()
    build_rast()


if __name__ == "__main__":
    build_all()
    print("built", LIB_DIR)
