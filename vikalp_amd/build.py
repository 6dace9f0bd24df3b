"""Builds the native libraries in-tree (no JIT cache, nothing installed),
through the project's CMakeLists.txt (Ninja):

  vikalp_amd/lib/libvikalp_hip.so   HIP kernels for gfx950 + their C ABI
                                    (include/vikalp_hip.h)
  vikalp_amd/lib/libvikalp.so       C++20 host: contracts, HipBackend,
                                    solve_lp, vikalp::solve, generators
                                    (include/vikalp_solve.h, vikalp_lpgen.h)
  tests/cpp/build/hip_backend_test  the HipBackend consistency test (CTest
                                    hip_consistency), with its oracle copy

The CMake build tree is vikalp_amd/lib/cmake-build.  Every call configures
and compiles everything again (clean=False lets Ninja skip what is up to
date).

Usage:  python -m vikalp_amd.build   (or vikalp_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vikalp_amd")
LIB = os.path.join(PKG, "lib")
BUILD_DIR = os.path.join(LIB, "cmake-build")
TEST_DIR = os.path.join(ROOT, "tests", "cpp", "build")
ARCH = os.environ.get("VIKALP_OFFLOAD_ARCH", "gfx950")
HIP_COMPILER = os.environ.get("VIKALP_HIP_COMPILER", "/opt/rocm/llvm/bin/clang++")


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + proc.stdout[-6000:] + proc.stderr[-6000:])
    return proc


def build(verbose: bool = True, clean: bool = True) -> None:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(TEST_DIR, exist_ok=True)
    _run(["cmake", "-S", ROOT, "-B", BUILD_DIR, "-G", "Ninja", "-DCMAKE_BUILD_TYPE=Release",
          f"-DCMAKE_HIP_COMPILER={HIP_COMPILER}", f"-DCMAKE_HIP_ARCHITECTURES={ARCH}",
          f"-DVIKALP_LIB_DIR={LIB}", f"-DVIKALP_TEST_DIR={TEST_DIR}", "-DBUILD_TESTING=ON"])
    jobs = str(max(1, min(16, os.cpu_count() or 4)))
    _run(["cmake", "--build", BUILD_DIR, "-j", jobs] + (["--clean-first"] if clean else []))
    if verbose:
        print("built", os.path.join(LIB, "libvikalp_hip.so"), os.path.join(LIB, "libvikalp.so"),
              os.path.join(TEST_DIR, "hip_backend_test"))


if __name__ == "__main__":
    build()
    sys.exit(0)
