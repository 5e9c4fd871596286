"""Build recipe for oracle/_ref (test infrastructure only).

Compiles the reference's own device code, kernels.cl (with common.h beside it),
from where it lies in the reference checkout, for gfx950 with ROCm's clang, once
per configuration, using exactly the options lbmcl.hpp:131-156 hands to
clBuildProgram: -Werror -I<reference> -DDIM -DLWS -DSTRIDE_DIV -DSTRIDE_MOD
-DVISCOSITY -DVELOCITY (printed by an ostream with its default 6 significant
digits) and -DFP_SINGLE -cl-single-precision-constant or -DFP_DOUBLE, plus
-cl-fast-relaxed-math for -o.  Two warnings that -Werror would make fatal under
this clang stay warnings: FP64 adds -Wno-error=ignored-pragmas (clang does not
know the cl_amd_fp64 extension that kernels.cl:25 enables) and -o adds
-Wno-error=nan-infinity-disabled (NAN under finite-math-only).  OpenCL C 1.2
and -O3 are the runtime's defaults.  Like every GPU build here the code is
compiled without packed-FP32 instructions, which changes instruction
selection only (a packed FP32 operation rounds as its scalar form).
No reference file is copied or changed; the outputs (code objects and the
runner, oracle/ref_runner.c) go to oracle/_ref/ only.  The runner loads a code
object with clCreateProgramWithBinary, so the GPU host needs the ROCm OpenCL
runtime but not the reference's sources.

Each code object is disassembled after the build and rejected by
oracle/ref_check.py if it writes through the scalar data cache (code here
writes through vector stores only).

Usage: python -m oracle.ref_build [--reference /root/reference]
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
RUNNER = os.path.join(OUT, "ref_runner")
DEFAULT_REFERENCE = "/root/reference"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
CLANG = os.path.join(ROCM, "lib", "llvm", "bin", "clang")
OBJDUMP = os.path.join(ROCM, "lib", "llvm", "bin", "llvm-objdump")
ARCH = "gfx950"

# (precision, dim, stride, work-group x, fast math): the make test8/test32
# shapes, the 64^3 (strides 64 and 1) and 128^3 parity shapes, and the stride set of the
# performance comparison (benchmark.sh's CSoA strides and SoA) at 128^3 / 256^3.
CONFIGS = []
for _p in ("f32", "f64"):
    CONFIGS += [(_p, 8, 8, 8, False), (_p, 32, 32, 32, False), (_p, 64, 64, 64, False), (_p, 64, 1, 64, False),
                (_p, 128, 64, 64, False)]
    CONFIGS += [(_p, 256, s, 64, False) for s in (1, 64, 4096, 256 ** 3)]
    CONFIGS += [(_p, 128, s, 64, False) for s in (1, 4096, 128 ** 3)]
CONFIGS += [("f32", 32, 32, 32, True), ("f32", 128, 64, 64, True), ("f32", 256, 64, 64, True),
            ("f32", 256, 4096, 64, True)]

VISCOSITY = 0.0089
VELOCITY = 0.05

def co_name(prec, dim, stride, fast):
    return f"kernels_{prec}_d{dim}_s{stride}{'_o' if fast else ''}.co"


def co_path(prec, dim, stride, fast=False):
    return os.path.join(OUT, co_name(prec, dim, stride, fast))


def ostream_double(v: float) -> str:
    """`std::ostream << double` with the default flags: %g with 6 significant digits."""
    return "%g" % v


def options(prec, dim, stride, lwx, fast, reference):
    """kernelOptionsStr() of lbmcl.hpp:131-156 as a clang argument list."""
    opts = ["-Werror", "-I" + reference, f"-DDIM={dim}", f"-DLWS={lwx}",
            f"-DSTRIDE_DIV={stride.bit_length() - 1}", f"-DSTRIDE_MOD={stride - 1}",
            f"-DVISCOSITY={ostream_double(VISCOSITY)}", f"-DVELOCITY={ostream_double(VELOCITY)}"]
    if prec == "f32":
        opts += ["-DFP_SINGLE", "-cl-single-precision-constant"]
    else:
        # kernels.cl:25 also enables cl_amd_fp64, an extension clang does not
        # know; the pragma is ignored (cl_khr_fp64 on the line before enables
        # double) but -Werror would turn that warning into an error
        opts += ["-DFP_DOUBLE", "-Wno-error=ignored-pragmas"]
    if fast:
        # under finite-math-only clang warns that the NAN stores of kernels.cl
        # (wall and non-fluid cells, :297-313) are undefined; -Werror would
        # stop the build, so that one warning stays a warning
        opts += ["-cl-fast-relaxed-math", "-Wno-error=nan-infinity-disabled"]
    return opts


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed ({r.returncode}):\n{r.stdout}{r.stderr}")
    return r.stdout


def build_kernel(cfg, reference):
    prec, dim, stride, lwx, fast = cfg
    out = co_path(prec, dim, stride, fast)
    cmd = [CLANG, "-x", "cl", "-cl-std=CL1.2", "-Xclang", "-finclude-default-header", "-target", "amdgcn-amd-amdhsa",
           f"-mcpu={ARCH}", "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           *options(prec, dim, stride, lwx, fast, reference), os.path.join(reference, "kernels.cl"), "-o", out]
    _run(cmd)
    from .ref_check import check_no_scalar_writes  # build-time only

    check_no_scalar_writes(out, OBJDUMP, ARCH)
    return out


def build_runner():
    cc = shutil.which("gcc") or "gcc"
    _run([cc, "-std=c99", "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROCM, "include"),
          os.path.join(HERE, "ref_runner.c"), "-o", RUNNER, "-lOpenCL"])


def build(reference=DEFAULT_REFERENCE, quiet=False):
    """Build everything; returns False (and builds nothing) without the reference checkout."""
    if not os.path.exists(os.path.join(reference, "kernels.cl")):
        if not quiet:
            print(f"[ref_build] {reference}/kernels.cl not found: oracle/_ref not built", file=sys.stderr)
        return False
    os.makedirs(OUT, exist_ok=True)
    build_runner()
    for cfg in CONFIGS:
        build_kernel(cfg, reference)
    if not quiet:
        print(f"[ref_build] {len(CONFIGS)} code objects and ref_runner in {OUT}", file=sys.stderr)
    return True


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--reference", default=DEFAULT_REFERENCE)
    a = ap.parse_args(argv)
    return 0 if build(a.reference) else 1


if __name__ == "__main__":
    sys.exit(main())
