"""oracle/_ref: the reference's kernels.cl built offline, and its runner (CPU checks)."""
import os
import subprocess

import pytest

from oracle import ref_build

needs_ref = pytest.mark.skipif(not os.path.exists(ref_build.RUNNER),
                               reason="oracle/_ref is not built (the reference checkout is absent)")


def test_options_are_kernel_options_str():
    """lbmcl.hpp:131-156 for dim 128, lws 64, stride 64, nu 0.0089, U 0.05."""
    o = ref_build.options("f32", 128, 64, 64, False, "/ref")
    assert o == ["-Werror", "-I/ref", "-DDIM=128", "-DLWS=64", "-DSTRIDE_DIV=6", "-DSTRIDE_MOD=63",
                 "-DVISCOSITY=0.0089", "-DVELOCITY=0.05", "-DFP_SINGLE", "-cl-single-precision-constant"]
    d = ref_build.options("f64", 256, 256 ** 3, 64, False, "/ref")
    assert "-DSTRIDE_DIV=24" in d and "-DSTRIDE_MOD=16777215" in d and "-DFP_DOUBLE" in d
    assert "-cl-single-precision-constant" not in d
    assert "-cl-fast-relaxed-math" in ref_build.options("f32", 32, 32, 32, True, "/ref")
    # std::ostream's default formatting of a double: 6 significant digits
    assert ref_build.ostream_double(0.0089) == "0.0089" and ref_build.ostream_double(1.0 / 3.0) == "0.333333"


@needs_ref
def test_every_configuration_is_built():
    missing = [ref_build.co_name(p, d, s, f) for p, d, s, _, f in ref_build.CONFIGS
               if not os.path.exists(ref_build.co_path(p, d, s, f))]
    assert not missing


@needs_ref
@pytest.mark.parametrize("args", [
    ["x.co", "f32", "512", "64", "64,1,1", "10", "0", "0", "-"],     # 19 * 512^3 overflows the reference's int index
    ["x.co", "f16", "64", "64", "64,1,1", "10", "0", "0", "-"],      # precision
    ["x.co", "f32", "8", "1024", "8,1,1", "10", "0", "0", "-"],      # stride > dim^3
    ["x.co", "f32", "64", "48", "64,1,1", "10", "0", "0", "-"],      # stride not a power of two
    ["x.co", "f32", "64", "64", "48,1,1", "10", "0", "0", "-"],      # work-group does not divide dim
    ["x.co", "f32", "64", "64", "64,4,8", "10", "0", "0", "-"],      # 2048 work-items
    ["x.co", "f32", "64", "64", "64,1,1", "10", "0", "10", "-"],     # skip >= iterations
])
def test_runner_rejects_bad_shapes_before_any_device_call(args):
    r = subprocess.run([ref_build.RUNNER, *args], capture_output=True, text=True, timeout=60,
                       stdin=subprocess.DEVNULL)
    assert r.returncode == 2, r.stderr
