"""Builds libwarpmi.so (the gfx950 HIP library) in-tree with hipcc.

Run as ``python -m warp_amd.build``; __graft_entry__.build() calls
:func:`build`.  Every HIP translation unit is compiled with
-ffp-contract=off so the FP64 kernels keep the reference's rounding, and
with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libwarpmi.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["mg3d.hip", "particles.hip", "lostpart.hip", "condfield.hip", "fielddiag.hip",
               "unet.hip", "sort.hip"]
CXX_SOURCES = ["conductors.cpp", "abi.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable"]
GPU = ["--offload-arch=" + ARCH, "-munsafe-fp-atomics",
       "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(jobs=4):
    os.makedirs(os.path.join(LIBDIR, "obj"), exist_ok=True)
    objs, procs = [], []
    for src in HIP_SOURCES + CXX_SOURCES:
        obj = os.path.join(LIBDIR, "obj", src + ".o")
        cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip"] + GPU
        else:
            cmd[1:1] = ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
        objs.append(obj)
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError("compile failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
