"""The C ABI and the command line without a GPU (CPU only).

* liblbm_hip.so loads and exports every function include/lbm_hip.h declares;
  the ctypes layer binds exactly that set.
* Configuration errors come back as status codes with a message, before any
  device is touched (lbm_create validates first).
* The lbmcl command line parses the reference's options and rejects bad values
  with the reference's messages (libs/lbm_options.hpp:109-161).
"""
import os
import re
import subprocess

import pytest

from lbmcl_amd import build as native_build
from lbmcl_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(lbm_\w+)\s*\(", text, re.M)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "lbm_create" in names and "lbm_compute" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(capi.LIB_PATH):
        native_build.build_lib()
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT\s+(lbm_\w+)$", out, re.M))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    assert sorted(capi.EXPORTS) == declared_functions()
    L = capi.lib()
    for n in declared_functions():
        assert hasattr(L, n)


@pytest.mark.parametrize("kw,code,msg", [
    (dict(dim=12), capi.LBM_ERR_INVALID, "power of two"),
    (dict(dim=0), capi.LBM_ERR_INVALID, "power of two"),
    (dict(dim=2048), capi.LBM_ERR_INVALID, "above 1024"),
    (dict(dim=8, stride=3), capi.LBM_ERR_INVALID, "stride"),
    (dict(dim=8, stride=1024), capi.LBM_ERR_INVALID, "stride"),
    (dict(dim=8, precision="f64", storage="fp16"), capi.LBM_ERR_INVALID, "FP32"),
    (dict(dim=32, kernel=capi.LBM_KERNEL_ROW), capi.LBM_ERR_INVALID, "row kernel"),
    (dict(dim=128, stride=2, kernel=capi.LBM_KERNEL_VEC), capi.LBM_ERR_INVALID, "vector kernel"),
    (dict(dim=128, stride=64, kernel=capi.LBM_KERNEL_TILE), capi.LBM_ERR_INVALID, "tile kernel"),
    (dict(dim=8, stride=1, kernel=capi.LBM_KERNEL_TILE), capi.LBM_ERR_INVALID, "tile kernel"),
    (dict(dim=8, viscosity=-1.0), capi.LBM_ERR_INVALID, "viscosity"),
    (dict(dim=8, velocity=float("nan")), capi.LBM_ERR_INVALID, "velocity"),
    (dict(dim=8, lws=(64, 64, 1)), capi.LBM_ERR_INVALID, "work-group"),
    (dict(dim=8, optimize=3), capi.LBM_ERR_INVALID, "LBM_ARITH_OPENCL"),
    (dict(dim=8, storage="bf16", optimize=capi.LBM_ARITH_OPENCL), capi.LBM_ERR_INVALID, "native storage"),
    (dict(dim=8, device=99), capi.LBM_ERR_DEVICE, "no HIP device"),
])
def test_create_rejects_bad_configurations(kw, code, msg):
    with pytest.raises(capi.LbmError) as e:
        capi.Context(**kw)
    assert e.value.code == code
    assert msg in str(e.value)


def test_null_arguments_are_errors():
    L = capi.lib()
    assert L.lbm_create(None, None) == capi.LBM_ERR_INVALID
    assert L.lbm_destroy(None) == capi.LBM_ERR_INVALID
    assert L.lbm_initialize(None) == capi.LBM_ERR_INVALID
    assert L.lbm_compute(None, 1, 0, 0) == capi.LBM_ERR_INVALID
    assert "null" in capi._last_error()


@pytest.fixture(scope="module")
def lbmcl_bin():
    if not os.path.exists(native_build.CLI):
        native_build.build_cli()
    return native_build.CLI


def run_cli(binary, *args, stdin=subprocess.DEVNULL):
    return subprocess.run([binary, *args], capture_output=True, text=True, stdin=stdin, timeout=60)


def test_cli_help(lbmcl_bin):
    r = run_cli(lbmcl_bin, "-h")
    assert r.returncode == 1
    for opt in ("--platform", "--dim", "--viscosity", "--velocity", "--iterations", "--every", "--work_group_size",
                "--stride", "--use_double", "--optimize", "--vtk_path", "--dump_path", "--dump_map", "--dump_f",
                "--storage", "--arith"):
        assert opt in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["-d", "-4"], "Please enter a valid lattice dimension"),
    (["-n", "-0.1"], "Please enter a valid viscosity value"),
    (["-i", "-1"], "Please enter a valid number of iterations"),
    (["-e", "-2"], "Please enter a valid number for save simulation results every N iterations"),
    (["-s", "-8"], "Please enter a valid number for stride value"),
    (["-P", "-1"], "Please enter a valid platform"),
    (["-D", "-1"], "Please enter a valid device"),
    (["--storage", "fp8"], "Please enter a valid storage"),
    (["-F", "--storage", "fp16"], "single precision"),
    (["--arith", "fast"], "Please enter a valid arithmetic"),
    (["-o", "--arith", "opencl"], "drop --arith or -o"),
    (["--arith", "opencl", "--storage", "bf16"], "needs native storage"),
])
def test_cli_rejects_bad_values(lbmcl_bin, args, msg):
    r = run_cli(lbmcl_bin, *args)
    assert r.returncode == 1
    assert msg in r.stderr


def test_cli_rounds_dim_and_checks_work_group(lbmcl_bin):
    # dim 12 rounds to 8 (lbmcl.hpp:366-369); a work-group larger than the lattice exits -1 (lbmcl.hpp:377-380)
    r = run_cli(lbmcl_bin, "-d", "12", "-w", "32,32,1")
    assert "dim is rounded to the previous power of 2: 8" in r.stdout
    assert r.returncode == 255 and "good work_group_size" in r.stderr


def test_cli_without_device_asks_then_fails(lbmcl_bin):
    """No device (or an invalid id) lists the devices and asks on stdin like the reference;
    end of input is an error instead of an endless prompt."""
    if capi.device_count() > 0:
        pytest.skip("a GPU is present")
    r = run_cli(lbmcl_bin, "-d", "8", "-D", "0")
    assert "Select a device" in r.stdout
    assert r.returncode != 0 and "no device selected" in r.stderr


def test_c_example_builds_and_reports_missing_device(lbmcl_bin):
    """csrc/examples/capi_example.c compiles as C99 against the header and links the library;
    without a device lbm_create fails with lbm_last_error's message."""
    exe = native_build.CAPI_EXAMPLE
    assert os.path.exists(exe)
    if capi.device_count() > 0:
        pytest.skip("a GPU is present")
    r = run_cli(exe, "16", "2")
    assert r.returncode == 1 and "no HIP device" in r.stderr
