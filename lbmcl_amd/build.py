"""Build the native parts of lbmcl_amd in-tree.

Outputs (git-ignored, they travel with the tree to the GPU host):
  lbmcl_amd/liblbm_hip.so      HIP kernels + C ABI (include/lbm_hip.h), gfx950
  lbmcl_amd/bin/lbmcl          the lbmcl command line (C++ host, links liblbm_hip.so)
  lbmcl_amd/bin/capi_example   the C ABI used from plain C (csrc/examples/capi_example.c)
  lbmcl_amd/libmembw.so        copy and access-pattern calibration kernels (bench.py)
  lbmcl_amd/bin/membw          the same as a program of its own (rocprofv3 counter calibration)

Usage: python -m lbmcl_amd.build [--only lib|cli|membw]
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BIN = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Every HIP compile: gfx950 only, no packed-FP32 instructions, no contraction.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off", "-Wall", "-Wno-unused-function",
]

LIB = os.path.join(PKG, "liblbm_hip.so")
MEMBW = os.path.join(PKG, "libmembw.so")
CLI = os.path.join(BIN, "lbmcl")
MEMBW_BIN = os.path.join(BIN, "membw")
CAPI_EXAMPLE = os.path.join(BIN, "capi_example")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


OBJ_DIR = os.path.join(PKG, "_build")


def build_lib():
    # The LBM_ARITH_OPENCL kernels need OpenCL's FP32 division precision, a
    # per-translation-unit flag: that unit is compiled on its own.
    os.makedirs(OBJ_DIR, exist_ok=True)
    opencl_obj = os.path.join(OBJ_DIR, "lbm_opencl_arith.o")
    _run([HIPCC, *HIP_FLAGS, "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-c", "-I", INCLUDE, "-I", CSRC,
          os.path.join(CSRC, "lbm_opencl_arith.hip"), "-o", opencl_obj])
    _run([HIPCC, *HIP_FLAGS, "-shared", "-I", INCLUDE, "-I", CSRC,
          os.path.join(CSRC, "lbm_capi.hip"), os.path.join(CSRC, "lbm_adjoint.hip"), os.path.join(CSRC, "lbm_noalloc.hip"),
          "-x", "none", opencl_obj, "-o", LIB])


def build_membw():
    _run([HIPCC, *HIP_FLAGS, "-shared", os.path.join(CSRC, "membw.hip"), "-o", MEMBW])
    os.makedirs(BIN, exist_ok=True)
    _run([HIPCC, *HIP_FLAGS, "-DMEMBW_MAIN", os.path.join(CSRC, "membw.hip"), "-o", MEMBW_BIN])


def build_cli():
    os.makedirs(BIN, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    common = ["-std=c++17", "-O2", "-Wall", "-Wextra", "-I", INCLUDE, "-I", CSRC]
    _run([cxx, *common, os.path.join(CSRC, "main.cpp"), "-o", CLI,
          "-L", PKG, "-llbm_hip", "-Wl,-rpath,$ORIGIN/.."])
    cc = shutil.which("gcc") or "gcc"
    _run([cc, "-std=c99", "-O2", "-Wall", "-Wextra", "-pedantic", "-I", INCLUDE,
          os.path.join(CSRC, "examples", "capi_example.c"), "-o", CAPI_EXAMPLE,
          "-L", PKG, "-llbm_hip", "-Wl,-rpath,$ORIGIN/.."])


def build(only=None):
    if only in (None, "lib"):
        build_lib()
    if only in (None, "membw"):
        build_membw()
    if only in (None, "cli"):
        build_cli()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--only", choices=["lib", "cli", "membw"])
    args = ap.parse_args(argv)
    build(args.only)


if __name__ == "__main__":
    sys.exit(main())
