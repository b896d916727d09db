"""Builds the in-tree native library ``luz_amd/libluzgpu.so`` (HIP for gfx950 + C++20 host).

Host C++ (scene reader, BVH builders, light/environment tables) is compiled with
``-ffp-contract=off`` so its FP64 arithmetic rounds exactly like the reference build.
Device code is compiled for gfx950 only; contraction is allowed there except where a
``#pragma clang fp contract(off)`` scope pins the bit-exact FP64 pieces (sampler, camera).
Packed-fp32 instructions are disabled for all device code.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libluzgpu.so")
BUILD = os.path.join(HERE, "_build")

HOST_SOURCES = ["host/scene_file.cpp", "host/accel.cpp", "host/scene_desc.cpp"]
DEVICE_SOURCES = ["luzgpu.hip"]
ARCH = "gfx950"


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the luzgpu library needs ROCm's hipcc")


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return proc.stdout


def build(verbose: bool = False) -> str:
    hipcc = _hipcc()
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    common = ["-std=c++20", "-O3", "-fPIC", "-Wall", "-Wno-unused-parameter"]
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        _run(["g++", *common, "-ffp-contract=off", "-fno-fast-math", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        _run([hipcc, *common, f"--offload-arch={ARCH}", "-x", "hip",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-munsafe-fp-atomics", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
