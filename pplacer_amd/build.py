"""Builds libpplacer_hip.so in-tree for gfx950 with hipcc.

    python -m pplacer_amd.build
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpplacer_hip.so")

HIP_SOURCES = ["pp_capi.hip", "pp_noalloc.hip"]
CPP_SOURCES = ["pp_host.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]
# Packed FP32 instructions stay disabled for all device code.
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd):
    print("+ " + " ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    obj_dir = os.path.join(HERE, "_objs")
    os.makedirs(obj_dir, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        _run([hipcc, "-c", "-x", "hip"] + DEVICE + COMMON + [os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CPP_SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        _run(["g++", "-c"] + COMMON + ["-I", "/opt/rocm/include", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "--offload-arch=gfx950", "-fPIC"] + objs + ["-o", tmp, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
