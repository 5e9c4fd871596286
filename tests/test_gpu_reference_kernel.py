"""The HIP path against the reference's own kernels.cl run on the same GPU.

oracle/_ref holds kernels.cl compiled offline (oracle/ref_build.py, the options
of lbmcl.hpp:131-156) and a runner that drives it through the ROCm OpenCL
runtime with the host flow of lbmcl.hpp:392-521 (oracle/ref_runner.c).  This is
"the reference run at the same precision".

* LBM_ARITH_OPENCL, the value of a zero-initialised lbm_config.optimize,
  evaluates kernels.cl's expressions the way ROCm's OpenCL compiler does
  (multiply-adds contracted as clang forms them, FP32 division to 2.5 ulp):
  rho, u and the full CSoA lattice that the next iteration reads are
  bit-identical to the reference's, for every kernel form.  The main cases run
  a context whose configuration leaves `optimize` at zero.
* LBM_ARITH_EXACT, the CPU oracle's form (bit-identical to it; the cases up to
  64^3 also run the oracle), evaluates every operation as written.  It is a
  checking aid, not the reference's rounding: against the reference it is
  within 1e-12 relative in FP64, and in FP32 the difference grows with the
  number of steps (beyond 1e-5 after a few hundred), so there it is bounded by
  the reference's own FP32 rounding error -- measured against the reference run
  in FP64 -- instead.

ROCm's OpenCL limits a work-group to 256 work-items, so the make test8 and
test32 work-groups (8,8,8) and (32,32,1) (Makefile:77, 85) are replaced by
(8,8,4) and (32,8,1); no result depends on the work-group size.

The throughput test runs the reference's kernels over a sweep of strides and
work-group sizes at 256^3 (the largest lattice its int indexing allows) and
requires the HIP path to be faster than the best of them.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from lbmcl_amd import build as native_build
from lbmcl_amd import capi
from lbmcl_amd.vti import read_vti
from oracle import ref_build
from tests.util import max_rel, window, window_v

pytestmark = pytest.mark.gpu

TOL = {"f32": 1e-5, "f64": 1e-12}


@pytest.fixture(scope="module", autouse=True)
def reference_available():
    reason = O.ref_unavailable_reason()
    if reason:
        pytest.skip(f"the reference's kernels cannot run here: {reason}")


def _compare(ours, ref, tol, label, mask=None):
    a, b = np.asarray(ours, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    if mask is not None:
        a, b = a[..., mask], b[..., mask]
    assert np.array_equal(np.isnan(a), np.isnan(b)), f"{label}: NaN patterns differ"
    rel = max_rel(a, b)
    print(f"{label}: max relative difference {rel:.3e} (limit {tol:g})", flush=True)
    assert rel <= tol, f"{label}: {rel:.3e} > {tol:g}"
    return rel


def _hip_run(precision, n, stride, iterations, arith=None, kernel=capi.LBM_KERNEL_AUTO):
    """arith None: the context's default (lbm_config.optimize left at zero)."""
    kw = {} if arith is None else {"arith": arith}
    g = capi.Context(n, stride=stride, precision=precision, kernel=kernel, **kw)
    if arith is None:
        assert g.cfg.optimize == 0
    with g:
        g.initialize()
        if iterations > 1:
            g.compute_range(1, iterations - 1, 0)
        g.compute(iterations, update_macro=True)
        rho, u = g.read_rho(), g.read_u().reshape(3, n ** 3)
        # the lattice that iteration iterations + 1 reads (lbmcl.hpp:518)
        f = g.read_f(1 if (iterations + 1) % 2 == 0 else 0)
    return rho, u, f


CASES = [
    # make test8 and make test32 shapes (Makefile:73-86), 64^3, and the 128^3 x 1000 configuration
    ("f32", 8, 8, (8, 8, 4), 10),
    ("f64", 8, 8, (8, 8, 4), 10),
    ("f32", 32, 32, (32, 8, 1), 500),
    ("f64", 32, 32, (32, 8, 1), 500),
    ("f32", 64, 64, (64, 1, 1), 300),
    ("f64", 64, 64, (64, 1, 1), 300),
    ("f32", 128, 64, (64, 1, 1), 1000),
    ("f64", 128, 64, (64, 1, 1), 1000),
]


def _bits_equal(ours, ref, label):
    a, b = np.asarray(ours), np.asarray(ref)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    bad = int(np.count_nonzero(~same))
    print(f"{label}: {bad} of {a.size} values differ from the reference", flush=True)
    assert bad == 0, f"{label}: {bad} values differ (max relative {max_rel(a, b):.3e})"


@pytest.mark.parametrize("precision,n,stride,lws,iterations", CASES)
def test_default_arithmetic_is_bit_identical_to_reference_kernels(tmp_path, precision, n, stride, lws, iterations):
    """A configuration with `optimize` left at zero: the reference's results."""
    label = f"default (OpenCL) arithmetic {n}^3 {precision} stride {stride} x {iterations}"
    print(f"reference kernels.cl vs HIP: {label}", flush=True)
    ref = O.ref_run(precision, n, stride, lws, iterations, out_prefix=str(tmp_path / "ref"))
    rho, u, f = _hip_run(precision, n, stride, iterations)
    _bits_equal(rho, ref["rho"], f"{label} rho")
    _bits_equal(u, ref["u"], f"{label} u")
    _bits_equal(f, ref["f"], f"{label} f")


@pytest.mark.parametrize("stride", [1, 4096, 128 ** 3])
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_opencl_arithmetic_other_layouts(tmp_path, precision, stride):
    """128^3 x 50 in AoS (stride 1, the tile kernel), CSoA 4096 and SoA (N^3)."""
    n, iterations = 128, 50
    label = f"OpenCL arithmetic {n}^3 {precision} stride {stride} x {iterations}"
    print(f"reference kernels.cl vs HIP: {label}", flush=True)
    ref = O.ref_run(precision, n, stride, (64, 1, 1), iterations, out_prefix=str(tmp_path / "ref"))
    rho, u, f = _hip_run(precision, n, stride, iterations, arith=capi.LBM_ARITH_OPENCL)
    _bits_equal(rho, ref["rho"], f"{label} rho")
    _bits_equal(u, ref["u"], f"{label} u")
    _bits_equal(f, ref["f"], f"{label} f")


@pytest.mark.parametrize("precision,n,stride,kernel", [
    ("f32", 64, 64, capi.LBM_KERNEL_GENERIC), ("f32", 64, 64, capi.LBM_KERNEL_ROW),
    ("f32", 256, 64, capi.LBM_KERNEL_VEC), ("f32", 64, 1, capi.LBM_KERNEL_TILE), ("f32", 256, 64, capi.LBM_KERNEL_LDS),
    ("f64", 256, 64, capi.LBM_KERNEL_LDS),
    ("f64", 64, 64, capi.LBM_KERNEL_GENERIC), ("f64", 64, 64, capi.LBM_KERNEL_ROW),
    ("f64", 64, 1, capi.LBM_KERNEL_TILE),
])
def test_opencl_arithmetic_every_kernel_form(tmp_path, precision, n, stride, kernel):
    """The generic, row, 16-byte vector and tile kernels in the OpenCL form (the
    cases above run the automatic choice: 8-byte vector or tile)."""
    iterations = 100 if n <= 64 else 20
    label = f"OpenCL arithmetic {n}^3 {precision} stride {stride} kernel {kernel} x {iterations}"
    print(f"reference kernels.cl vs HIP: {label}", flush=True)
    ref = O.ref_run(precision, n, stride, (64, 1, 1), iterations, out_prefix=str(tmp_path / "ref"))
    rho, u, f = _hip_run(precision, n, stride, iterations, arith=capi.LBM_ARITH_OPENCL, kernel=kernel)
    _bits_equal(rho, ref["rho"], f"{label} rho")
    _bits_equal(u, ref["u"], f"{label} u")
    _bits_equal(f, ref["f"], f"{label} f")


@pytest.mark.parametrize("precision,n,stride,lws,iterations", CASES)
def test_exact_arithmetic_against_reference_kernels(tmp_path, precision, n, stride, lws, iterations):
    label = f"exact arithmetic {n}^3 {precision} stride {stride} x {iterations}"
    print(f"reference kernels.cl vs HIP: {label}", flush=True)
    ref = O.ref_run(precision, n, stride, lws, iterations, out_prefix=str(tmp_path / "ref"))
    rho, u, f = _hip_run(precision, n, stride, iterations, arith=capi.LBM_ARITH_EXACT)
    if n <= 64:
        # the CPU oracle itself (bit-identical to the HIP exact form) against the reference
        o = O.PullSim(n, stride=stride, precision=precision)
        o.initialize()
        for _ in range(iterations):
            o.step(update_macro=True)
        assert np.array_equal(o.rho, rho, equal_nan=True) and np.array_equal(o.u.reshape(3, -1), u, equal_nan=True)
    if precision == "f64":
        _compare(rho, ref["rho"], TOL["f64"], f"{label} rho")
        _compare(u, ref["u"], TOL["f64"], f"{label} u")
        _compare(f, ref["f"], TOL["f64"], f"{label} f")
        return
    # FP32: within the reference's own FP32 rounding error (its FP32 run
    # against its FP64 run), and never more than 1e-4 relative
    ref64 = O.ref_run("f64", n, stride, lws, iterations, out_prefix=str(tmp_path / "ref64"))
    for name, ours, theirs, exact in (("rho", rho, ref["rho"], ref64["rho"]), ("u", u, ref["u"], ref64["u"])):
        own = max_rel(theirs, exact)
        print(f"{label} {name}: reference FP32 vs reference FP64 {own:.3e}; HIP vs reference FP64 "
              f"{max_rel(ours, exact):.3e}", flush=True)
        _compare(ours, theirs, max(TOL["f32"], min(2 * own, 1e-4)), f"{label} {name}")
        assert max_rel(ours, exact) <= 1.5 * own + 1e-7, f"{label} {name}: further from FP64 than the reference"


@pytest.mark.parametrize("n,stride,lws,iterations", [(32, 32, (32, 8, 1), 500), (128, 64, (64, 1, 1), 1000)])
def test_fast_math_matches_reference_fast_math(tmp_path, n, stride, lws, iterations):
    """-o: the reference's -cl-fast-relaxed-math build against the fused collision.
    Under finite-math-only the reference's NaN stores are undefined, so rho and u
    are compared on the cells that store them (FLUID and MOVING, kernels.cl:382)."""
    label = f"-o {n}^3 f32 stride {stride} x {iterations}"
    print(f"reference kernels.cl vs HIP: {label}", flush=True)
    ref = O.ref_run("f32", n, stride, lws, iterations, out_prefix=str(tmp_path / "ref"), fast=True)
    rho, u, _ = _hip_run("f32", n, stride, iterations, arith=capi.LBM_ARITH_FUSED)
    stored = np.isfinite(rho)
    _compare(rho, ref["rho"], 1e-4, f"{label} rho", mask=stored)
    _compare(u, ref["u"], 1e-4, f"{label} u", mask=stored)


def _hip_mlups(precision, n, stride, steps, warmup, arith=capi.LBM_ARITH_REFERENCE):
    g = capi.Context(n, stride=stride, precision=precision, arith=arith)
    with g:
        g.initialize()
        g.compute_range(1, warmup, 0)
        g.finish()
        g.clear_events()
        g.compute_range(warmup + 1, steps, 0)
        g.finish()
        ms = g.kernels_time_ms()
    return (n - 2) ** 3 * steps / (ms * 1e3)


@pytest.mark.parametrize("n", [128, 256])
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_faster_than_reference_kernels(precision, n):
    """The reference's compute kernel over strides 1, 64, 4096, N^3 and work-groups
    of 64 and 256 work-items -- plus its -o build for FP32 at 256^3 -- against the HIP
    path, kernel MLUPS from profiling events (the reference's kernelsMLUPS,
    lbmcl.hpp:606-619).  256^3 (the largest lattice the reference indexes): 200
    iterations after 20, HIP at stride 4096.  128^3: the 1000-iteration
    configuration timed over iterations 100..1000, HIP at its stride 64."""
    steps, warmup = (200, 20) if n == 256 else (901, 99)
    hip_stride = 4096 if n == 256 else 64
    rows = []
    fasts = (False, True) if precision == "f32" else (False,)
    for fast in fasts:
        for stride in (1, 64, 4096, n ** 3):
            if not os.path.exists(ref_build.co_path(precision, n, stride, fast)):
                continue
            # 64 and 256 work-items (256 as (128, 2, 1) at 128^3: x must divide dim)
            for lws in ((64, 1, 1), (min(n, 256), 256 // min(n, 256), 1)):
                r = O.ref_run(precision, n, stride, lws, warmup + steps, every=0, skip=warmup, fast=fast)
                rows.append((r["mlups"], f"reference kernels.cl{' -o' if fast else ''} stride {stride} lws {lws}"))
                print(f"{n}^3 {precision} {rows[-1][1]}: {r['mlups']:.0f} MLUPS", flush=True)
    best_ref = max(rows)
    ours = _hip_mlups(precision, n, hip_stride, steps, warmup)
    ours_fast = _hip_mlups(precision, n, hip_stride, steps, warmup, arith=capi.LBM_ARITH_FUSED)
    ours_exact = _hip_mlups(precision, n, hip_stride, steps, warmup, arith=capi.LBM_ARITH_EXACT)
    print(f"{n}^3 {precision} HIP stride {hip_stride}: default (OpenCL) arithmetic {ours:.0f} MLUPS, fused (-o) "
          f"{ours_fast:.0f}, exact {ours_exact:.0f}; best reference {best_ref[0]:.0f} ({best_ref[1]}); ratio "
          f"{ours / best_ref[0]:.2f}", flush=True)
    assert min(ours, ours_fast, ours_exact) > best_ref[0]


@pytest.mark.parametrize("arith", [["--arith", "opencl"], []])
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_cli_opencl_arith_vti_is_the_reference_output(tmp_path, precision, arith):
    """`lbmcl` (whose default is --arith opencl) writes the VTI the reference's
    kernels produce on this GPU: the make test32 run (32^3, 500 iterations,
    stride 32), last VTI."""
    n, iterations = 32, 500
    args = ["-P0", "-D0", "-d", str(n), "-i", str(iterations), "-e", str(iterations), "-s", "32", "-w", "32,8,1",
            "-v", str(tmp_path)] + arith + (["-F"] if precision == "f64" else [])
    print("lbmcl", " ".join(args), flush=True)
    r = subprocess.run([native_build.CLI, *args], capture_output=True, text=True, timeout=300,
                       stdin=subprocess.DEVNULL)
    assert r.returncode == 0, r.stderr
    vt = read_vti(str(tmp_path / f"lbmcl.{iterations}.vti"))
    ref = O.ref_run(precision, n, 32, (32, 8, 1), iterations, out_prefix=str(tmp_path / "ref"))
    _bits_equal(vt["rho"], window(ref["rho"], n), f"lbmcl --arith opencl {precision} VTI rho")
    _bits_equal(vt["v"], window_v(ref["u"], n), f"lbmcl --arith opencl {precision} VTI v")
