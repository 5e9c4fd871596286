"""The lbmcl command line on the GPU, the way the reference tests itself.

make test8 / make test32 (Makefile:73-86) run lbmcl and then verify.py against
the Sailfish fixtures; make testall (Makefile:66-70) exercises the -f and -m
dumps.  The same flows run here through bin/lbmcl and lbmcl_amd.verify, with
thresholds, and every file is also compared with the CPU oracle: the VTIs
value for value, the f/map dumps byte for byte (formatted by the Python
statement of the reference writers in tests/test_formats.py).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from lbmcl_amd import build as native_build
from lbmcl_amd.vti import read_vti
from tests.test_formats import _f_dump, _map_dump
from tests.util import window, window_v

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = native_build.CLI


def run(args, cwd):
    print("lbmcl", " ".join(args), flush=True)
    return subprocess.run([CLI, *args], capture_output=True, text=True, cwd=cwd, timeout=300,
                          stdin=subprocess.DEVNULL)


def verify(iterations, every, target, pred, limits):
    cmd = [sys.executable, "-m", "lbmcl_amd.verify", "-i", str(iterations), "-e", str(every), "-t", target, "-p", pred,
           "--max-mse-rho", str(limits[0]), "--max-mse-u", str(limits[1]), "--max-mae-rho", str(limits[2]),
           "--max-mae-u", str(limits[3])]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=ROOT, timeout=300)
    print(r.stdout[-800:])
    return r


@pytest.mark.parametrize("optimize", [False, True])
def test_make_test8(tmp_path, golden_dir, optimize):
    # compared value for value with the oracle: the exact arithmetic (or -o)
    args = ["-P0", "-D0", "-d", "8", "-n", "0.0089", "-u", "0.05", "-i", "10", "-e", "1", "-w", "8,8,8", "-s", "8",
            "-v", str(tmp_path)] + (["-o"] if optimize else ["--arith", "exact"])
    r = run(args, tmp_path)
    assert r.returncode == 0, r.stderr
    stats = r.stderr.strip().splitlines()[-1].split(";")
    assert stats[1:8] == ["single", "8", "10", "1", "008,008,008", "8", "1" if optimize else "0"]
    assert float(stats[8]) > 0 and float(stats[9]) > 0
    v = verify(10, 1, os.path.join(golden_dir, "target_results", "8"), str(tmp_path), (1e-13, 5e-15, 1e-6, 4e-7))
    assert v.returncode == 0, v.stdout + v.stderr
    # value for value against the oracle
    sim = O.PushSim(8, stride=8, fast=optimize)
    sim.initialize()
    for it in range(0, 11):
        if it:
            sim.step(True)
        vt = read_vti(str(tmp_path / f"lbmcl.{it:02d}.vti"))
        assert np.array_equal(vt["rho"], window(sim.rho, 8), equal_nan=True), it
        assert np.array_equal(vt["v"], window_v(sim.u, 8), equal_nan=True), it


@pytest.mark.parametrize("arith", [["--arith", "exact"], []])
def test_make_test32(tmp_path, golden_dir, arith):
    """make test32 in the exact arithmetic and in the default, the reference's own
    rounding (--arith opencl): both meet verify.py's thresholds against the
    Sailfish fixtures."""
    r = run(["-P0", "-D0", "-d", "32", "-n", "0.0089", "-u", "0.05", "-i", "500", "-e", "20", "-w", "32,32,1", "-s",
             "32", "-v", str(tmp_path)] + arith, tmp_path)
    assert r.returncode == 0, r.stderr
    v = verify(500, 20, os.path.join(golden_dir, "target_results", "32"), str(tmp_path), (2e-12, 3e-14, 4e-6, 1.2e-6))
    assert v.returncode == 0, v.stdout + v.stderr


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_make_testall_dumps(tmp_path, precision):
    """-f -m: every f_<it>.dump holds the lattice at the start of iteration it, map.dump the cell codes."""
    n, its, stride = 8, 6, 4
    args = ["-P0", "-D0", "-d", str(n), "-i", str(its), "-e", "1", "-w", "8,8,8", "-s", str(stride), "-v",
            str(tmp_path), "-p", str(tmp_path), "-f", "-m", "--arith", "exact"] + (["-F"] if precision == "f64" else [])
    r = run(args, tmp_path)
    assert r.returncode == 0, r.stderr
    assert ("double" if precision == "f64" else "single") in r.stderr
    sim = O.PushSim(n, stride=stride, precision=precision)
    sim.initialize()
    assert (tmp_path / "map.dump").read_text() == _map_dump(sim.map, n)
    assert (tmp_path / "f_0.dump").read_text() == _f_dump(sim.f_collide, n, stride)
    for it in range(1, its + 1):
        expect = _f_dump(sim.input_buffer(it), n, stride)  # the lattice iteration `it` reads (lbmcl.hpp:517-519)
        assert (tmp_path / f"f_{it}.dump").read_text() == expect, it
        sim.step(True)
        vt = read_vti(str(tmp_path / f"lbmcl.{it}.vti"))
        assert np.array_equal(vt["rho"], window(sim.rho, n), equal_nan=True), it


def test_cli_fp16_storage_runs(tmp_path):
    r = run(["-P0", "-D0", "-d", "64", "-i", "20", "-e", "10", "-s", "64", "-w", "64,1,1", "--storage", "fp16", "-v",
             str(tmp_path)], tmp_path)
    assert r.returncode == 0, r.stderr
    vt = read_vti(str(tmp_path / "lbmcl.20.vti"))
    sim = O.PullSim(64, stride=64, storage=O.STORE_FP16)
    sim.initialize()
    for _ in range(20):
        sim.step(True)
    assert np.array_equal(vt["rho"], window(sim.rho, 64), equal_nan=True)


def test_capi_example_c_caller():
    """csrc/examples/capi_example.c: the C ABI from plain C (LBM_CONFIG_INIT: the
    reference's arithmetic), checked against the oracle's contracted form -- the
    same expressions, with FP32 division correctly rounded where the GPU rounds
    to OpenCL's 2.5 ulp -- within the FP32 parity bound (1e-5 relative)."""
    exe = native_build.CAPI_EXAMPLE
    n, iters = 64, 100
    print(exe, n, iters, flush=True)
    r = subprocess.run([exe, str(n), str(iters)], capture_output=True, text=True, timeout=300,
                       stdin=subprocess.DEVNULL)
    assert r.returncode == 0, r.stderr
    words = r.stdout.split()
    rho_c = float(words[words.index("rho(centre)") + 1])
    mass = float(words[words.index("mass") + 1])
    o = O.PullSim(n, stride=64, fast=O.ARITH_CONTRACTED)
    o.initialize()
    for _ in range(iters):
        o.step(update_macro=True)
    w = n - 2
    centre = (w // 2) + w * ((w // 2) + w * (w // 2))
    assert abs(rho_c - float(window(o.rho, n)[centre])) <= 1e-5, r.stdout
    assert abs(mass - o.diag[0]) <= 1e-6 * o.diag[0], (r.stdout, o.diag)
