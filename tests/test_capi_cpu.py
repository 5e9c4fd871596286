"""The C ABI and the command line without a GPU (CPU only).

* liblbm_hip.so loads and exports every function include/lbm_hip.h declares;
  the ctypes layer binds exactly that set.
* Configuration errors come back as status codes with a message, before any
  device is touched (lbm_create validates first).
* The lbmcl command line parses the reference's options and rejects bad values
  with the reference's messages (libs/lbm_options.hpp:109-161).
"""
import ctypes
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
    (dict(dim=8, arith=3), capi.LBM_ERR_INVALID, "LBM_ARITH_REFERENCE"),
    (dict(dim=8, arith=-1), capi.LBM_ERR_INVALID, "LBM_ARITH_REFERENCE"),
    (dict(dim=8, storage="bf16", arith=capi.LBM_ARITH_EXACT), capi.LBM_ERR_INVALID, "native storage"),
    (dict(dim=8, device=99), capi.LBM_ERR_DEVICE, "no HIP device"),
])
def test_create_rejects_bad_configurations(kw, code, msg):
    with pytest.raises(capi.LbmError) as e:
        capi.Context(**kw)
    assert e.value.code == code
    assert msg in str(e.value)


def _c_struct_sizes(tmp=[]):
    """sizeof(lbm_config) and sizeof(lbm_adjoint_config) as a C compiler lays them out."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "s.c"), os.path.join(d, "s")
        with open(src, "w") as fh:
            fh.write('#include <stdio.h>\n#include "lbm_hip.h"\nint main(void) {\n'
                     '  lbm_config c = LBM_CONFIG_INIT; lbm_adjoint_config a = LBM_ADJOINT_CONFIG_INIT;\n'
                     '  printf("%u %u\\n", c.struct_size, a.struct_size); return 0; }\n')
        subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src, "-o", exe],
                       check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    return int(out[0]), int(out[1])


def test_config_struct_size_is_checked():
    """A caller built against another layout of the configuration structs is
    rejected before anything else is read: lbm_create and lbm_adjoint compare
    struct_size with their own sizeof."""
    L = capi.lib()
    assert (ctypes.sizeof(capi.LbmConfig), ctypes.sizeof(capi.LbmAdjointConfig)) == _c_struct_sizes()
    cfg = capi.LbmConfig()
    cfg.dim, cfg.stride, cfg.viscosity, cfg.velocity = 8, 8, 0.0089, 0.05
    h = ctypes.c_void_p()
    for size in (0, ctypes.sizeof(capi.LbmConfig) - 8, ctypes.sizeof(capi.LbmConfig) + 8):
        cfg.struct_size = size
        assert L.lbm_create(ctypes.byref(cfg), ctypes.byref(h)) == capi.LBM_ERR_INVALID
        err = capi._last_error()
        assert f"struct_size is {size}" in err and f"has {ctypes.sizeof(capi.LbmConfig)} bytes" in err
        assert not h.value
    acfg = capi.LbmAdjointConfig(ctypes.sizeof(capi.LbmAdjointConfig) + 4, 0, 0, 10, 0)
    res = capi.LbmAdjointResult()
    assert L.lbm_adjoint(None, ctypes.byref(acfg), ctypes.byref(res), None) == capi.LBM_ERR_INVALID
    assert "lbm_adjoint_config.struct_size is 36" in capi._last_error()
    acfg.struct_size = ctypes.sizeof(capi.LbmAdjointConfig)
    assert L.lbm_adjoint(None, ctypes.byref(acfg), ctypes.byref(res), None) == capi.LBM_ERR_INVALID
    assert "null context" in capi._last_error()


def test_zero_config_defaults_to_reference_arithmetic():
    """lbm_config's zero `optimize` is the reference's arithmetic, and the C
    header agrees with the Python constants."""
    hdr = open(os.path.join(ROOT, "include", "lbm_hip.h")).read()
    for name in ("LBM_ARITH_REFERENCE", "LBM_ARITH_OPENCL", "LBM_ARITH_FUSED", "LBM_ARITH_EXACT"):
        m = re.search(rf"#define {name} (\d+)", hdr)
        assert m and int(m.group(1)) == getattr(capi, name), name
    assert capi.LBM_ARITH_REFERENCE == capi.LBM_ARITH_OPENCL == 0
    assert capi.LbmConfig().optimize == 0


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
