"""Builds the native libraries in-tree (no JIT cache, nothing installed).

  vikalp_amd/lib/libvikalp_hip.so   HIP kernels for gfx950 + their C ABI
                                    (include/vikalp_hip.h)
  vikalp_amd/lib/libvikalp.so       C++20 host: contracts, HipBackend,
                                    solve_lp, vikalp::solve, generators
                                    (include/vikalp_solve.h, vikalp_lpgen.h)

Usage:  python -m vikalp_amd.build   (or vikalp_amd.build.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vikalp_amd")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "lib", "obj")
INC = os.path.join(ROOT, "include")
HIP_SRC = os.path.join(PKG, "csrc", "hip")
HOST_SRC = os.path.join(PKG, "csrc", "host")
ARCH = os.environ.get("VIKALP_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 ops are disabled for every HIP translation unit.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++20", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + INC, "-I" + HIP_SRC]
CXX_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
             "-I" + INC, "-I" + os.path.join(HOST_SRC, "include"), "-I" + HOST_SRC]

HIP_UNITS = ["ops.hip", "pdhg.hip", "batch.hip", "unrolled.hip", "presolve.hip", "layout.hip", "panel.hip"]
HOST_UNITS = ["model.cpp", "hip_backend.cpp", "pdhg_lp.cpp", "batch_lp.cpp", "unrolled_lp.cpp", "qp.cpp", "solver.cpp", "capi.cpp", "mps_reader.cpp",
              "lpgen.cpp"]


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + proc.stdout + proc.stderr)
    return proc


def _compile_hip(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", out])
    return out


def _compile_cxx(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run(["g++", *CXX_FLAGS, "-c", src, "-o", out])
    return out


def build(verbose: bool = True) -> None:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as pool:
        hip_objs = list(pool.map(_compile_hip, [os.path.join(HIP_SRC, u) for u in HIP_UNITS]))
        host_objs = list(pool.map(_compile_cxx, [os.path.join(HOST_SRC, u) for u in HOST_UNITS]))
    hip_so = os.path.join(LIB, "libvikalp_hip.so")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", hip_so, *hip_objs])
    host_so = os.path.join(LIB, "libvikalp.so")
    _run(["g++", "-shared", "-fPIC", "-o", host_so, *host_objs, "-L" + LIB, "-lvikalp_hip",
          "-Wl,-rpath,$ORIGIN", "-Wl,--no-undefined"])
    if verbose:
        print("built", hip_so, host_so)


if __name__ == "__main__":
    build()
    sys.exit(0)
