"""Parity of the HIP kernels (through the C ABI) with the CPU oracle.

The HIP path and the oracle evaluate the same expressions in the same order
without contraction, so the expected agreement is bit for bit: rho, u, the full
CSoA lattice in the reference's push view, and the raw stored state.  The
tolerances of a same-precision comparison (1e-5 relative in FP32, 1e-12 in
FP64) are asserted as well, so a failure message distinguishes "bits differ"
from "out of tolerance".
"""
import os

import numpy as np
import pytest

import oracle as O
from lbmcl_amd import capi
from lbmcl_amd.verify import metrics
from lbmcl_amd.vti import read_vti
from tests.util import hip_arith, max_rel, mismatch_count, nan_equal, window, window_v

pytestmark = pytest.mark.gpu

TOL = {"f32": 1e-5, "f64": 1e-12}
STORE = {"native": O.STORE_NATIVE, "fp16": O.STORE_FP16, "bf16": O.STORE_BF16}


def _pair(n, stride, precision, storage="native", kernel=capi.LBM_KERNEL_AUTO, lws=(0, 0, 0),
          viscosity=0.0089, velocity=0.05, fast=False):
    # 16-bit storage has one arithmetic form (the zero default); the oracle uses it for any `fast`
    arith = hip_arith(fast) if storage == "native" else capi.LBM_ARITH_REFERENCE
    g = capi.Context(n, viscosity=viscosity, velocity=velocity, stride=stride, precision=precision,
                     storage=storage, kernel=kernel, lws=lws, arith=arith)
    o = O.PullSim(n, stride=stride, viscosity=viscosity, velocity=velocity, precision=precision,
                  storage=STORE[storage], fast=fast)
    g.initialize()
    o.initialize()
    return g, o


def _assert_macro_equal(g, o, n, precision, label):
    rg, ug = g.read_rho(), g.read_u()
    ro, uo = o.macro()
    rel_r = max_rel(rg, ro)
    rel_u = max_rel(ug, uo.reshape(-1))
    print(f"[{label}] max rel drho={rel_r:.3e} du={rel_u:.3e}", flush=True)
    assert rel_r <= TOL[precision] and rel_u <= TOL[precision], label
    assert nan_equal(rg, ro), f"{label}: rho bits differ in {mismatch_count(rg, ro)} cells"
    assert nan_equal(ug, uo.reshape(-1)), f"{label}: u bits differ in {mismatch_count(ug, uo.reshape(-1))} entries"


CASES = [
    # n, stride, precision, kernel, steps
    (8, 8, "f32", capi.LBM_KERNEL_AUTO, 10),
    (8, 1, "f64", capi.LBM_KERNEL_AUTO, 10),
    (16, 4096, "f32", capi.LBM_KERNEL_GENERIC, 16),
    (32, 32, "f64", capi.LBM_KERNEL_GENERIC, 24),
    (64, 64, "f32", capi.LBM_KERNEL_ROW, 12),
    (64, 4, "f32", capi.LBM_KERNEL_ROW, 12),
    (64, 1, "f64", capi.LBM_KERNEL_ROW, 12),
    (64, 64 ** 3, "f64", capi.LBM_KERNEL_ROW, 12),
    (64, 64, "f32", capi.LBM_KERNEL_GENERIC, 12),
    (128, 64, "f32", capi.LBM_KERNEL_AUTO, 8),
    (128, 64, "f64", capi.LBM_KERNEL_AUTO, 8),
    (256, 256 ** 3, "f32", capi.LBM_KERNEL_ROW, 4),
    (256, 4, "f32", capi.LBM_KERNEL_VEC, 5),
    (256, 256 ** 3, "f32", capi.LBM_KERNEL_VEC, 5),
    (128, 2, "f64", capi.LBM_KERNEL_VEC, 6),
    (128, 128, "f64", capi.LBM_KERNEL_VEC, 6),
    (512, 1024, "f64", capi.LBM_KERNEL_VEC, 2),
    (512, 4096, "f64", capi.LBM_KERNEL_VEC8, 2),  # two blocks per row: block-end elements from memory
    (256, 2, "f32", capi.LBM_KERNEL_VEC8, 5),
    (128, 128, "f32", capi.LBM_KERNEL_VEC8, 6),
    (64, 64 ** 3, "f64", capi.LBM_KERNEL_VEC8, 8),
    (64, 1, "f64", capi.LBM_KERNEL_VEC8, 8),
    # the row-pipeline kernel (rows staged in LDS by direct loads, one row ahead)
    (256, 4096, "f32", capi.LBM_KERNEL_LDS, 5),
    (256, 4, "f32", capi.LBM_KERNEL_LDS, 5),
    (512, 4096, "f32", capi.LBM_KERNEL_LDS, 2),
    (256, 2, "f64", capi.LBM_KERNEL_LDS, 4),
    (256, 256 ** 3, "f64", capi.LBM_KERNEL_LDS, 3),
    (512, 512, "f64", capi.LBM_KERNEL_LDS, 2),
    # the tile kernel (small strides, LDS-staged spans)
    (32, 1, "f32", capi.LBM_KERNEL_TILE, 12),
    (32, 16, "f32", capi.LBM_KERNEL_TILE, 12),
    (128, 1, "f32", capi.LBM_KERNEL_TILE, 6),
    (128, 8, "f32", capi.LBM_KERNEL_TILE, 6),
    (16, 1, "f64", capi.LBM_KERNEL_TILE, 12),
    (64, 2, "f64", capi.LBM_KERNEL_TILE, 8),
    (128, 8, "f64", capi.LBM_KERNEL_TILE, 6),
]


@pytest.mark.parametrize("n,stride,precision,kernel,steps", CASES)
def test_native_matches_oracle(n, stride, precision, kernel, steps):
    label = f"n={n} s={stride} {precision} k={kernel}"
    print(label, flush=True)
    g, o = _pair(n, stride, precision, kernel=kernel)
    with g:
        _assert_macro_equal(g, o, n, precision, label + " init")
        assert np.array_equal(g.read_map(), o.map)
        for it in range(1, steps + 1):
            g.compute(it, update_macro=True)
            o.step(update_macro=True)
            if it in (1, 2, steps):
                _assert_macro_equal(g, o, n, precision, f"{label} it={it}")
        # both lattices as stored, and (up to 256^3) in the reference's push view
        for buf, ob in ((0, o.A), (1, o.B)):
            if n <= 256:
                fg, fo = g.read_f(buf), o.view(ob)
                assert nan_equal(fg, fo), f"{label}: f view {buf} differs in {mismatch_count(fg, fo)} slots"
            sg = g.read_state(buf)
            assert nan_equal(sg, ob), f"{label}: stored state {buf} differs"


FAST_CASES = [
    (16, 16, "f32", capi.LBM_KERNEL_GENERIC, 10),
    (64, 64, "f64", capi.LBM_KERNEL_ROW, 10),
    (128, 64, "f32", capi.LBM_KERNEL_VEC8, 8),
    (256, 256, "f32", capi.LBM_KERNEL_VEC, 5),
    (128, 128, "f64", capi.LBM_KERNEL_VEC, 6),
    (64, 2, "f32", capi.LBM_KERNEL_TILE, 8),
    (32, 4, "f64", capi.LBM_KERNEL_TILE, 8),
]


@pytest.mark.parametrize("n,stride,precision,kernel,steps", FAST_CASES)
def test_fused_collision_matches_oracle(n, stride, precision, kernel, steps):
    """-o: the fused collision, bit for bit against the oracle's fused form."""
    label = f"fast n={n} s={stride} {precision} k={kernel}"
    g, o = _pair(n, stride, precision, kernel=kernel, fast=True)
    with g:
        for it in range(1, steps + 1):
            g.compute(it, update_macro=True)
            o.step(update_macro=True)
        _assert_macro_equal(g, o, n, precision, label)
        for buf, ob in ((0, o.A), (1, o.B)):
            assert nan_equal(g.read_state(buf), ob), f"{label}: stored state {buf} differs"


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
@pytest.mark.parametrize("n,stride,kernel,steps", [(16, 16, capi.LBM_KERNEL_GENERIC, 12), (64, 64, capi.LBM_KERNEL_ROW, 12),
                                                   (256, 256, capi.LBM_KERNEL_VEC8, 4), (512, 512, capi.LBM_KERNEL_VEC, 2),
                                                   (64, 1, capi.LBM_KERNEL_TILE, 8), (128, 32, capi.LBM_KERNEL_TILE, 4),
                                                   (256, 8, capi.LBM_KERNEL_LDS, 4), (512, 4096, capi.LBM_KERNEL_LDS, 2)])
def test_shifted_storage_matches_oracle(storage, n, stride, kernel, steps):
    label = f"{storage} n={n} s={stride} k={kernel}"
    g, o = _pair(n, stride, "f32", storage=storage, kernel=kernel)
    with g:
        for it in range(1, steps + 1):
            g.compute(it, update_macro=True)
            o.step(update_macro=True)
        _assert_macro_equal(g, o, n, "f32", label)
        for buf, ob in ((0, o.A), (1, o.B)):
            if n > 256:
                continue
            fg, fo = g.read_f(buf), o.view(ob)
            assert nan_equal(fg, fo), f"{label}: f view {buf} differs in {mismatch_count(fg, fo)} slots"


@pytest.mark.parametrize("n,stride,precision,kernel", [(16, 16, "f32", capi.LBM_KERNEL_GENERIC),
                                                       (64, 64, "f64", capi.LBM_KERNEL_ROW),
                                                       (128, 128, "f32", capi.LBM_KERNEL_ROW),
                                                       (256, 256, "f32", capi.LBM_KERNEL_VEC),
                                                       (128, 128, "f64", capi.LBM_KERNEL_VEC),
                                                       (64, 1, "f32", capi.LBM_KERNEL_TILE),
                                                       (32, 2, "f64", capi.LBM_KERNEL_TILE),
                                                       (256, 4096, "f32", capi.LBM_KERNEL_LDS),
                                                       (256, 64, "f64", capi.LBM_KERNEL_LDS)])
def test_diagnostics(n, stride, precision, kernel):
    g, o = _pair(n, stride, precision, kernel=kernel)
    with g:
        for it in range(1, 7):
            g.compute(it, update_macro=True, diagnostics=(it % 3 == 0))
            o.step(update_macro=True)
            if it % 3 == 0:
                mg, eg = g.read_diagnostics()
                mo, eo = o.diag
                # same per-cell FP64 terms summed in a different order
                assert abs(mg - mo) <= 1e-12 * abs(mo), (mg, mo)
                assert abs(eg - eo) <= 1e-10 * abs(eo), (eg, eo)
                ref = O.diagnostics(o.p, o.rho, o.u, precision)
                assert abs(mg - ref[0]) <= 1e-12 * abs(ref[0])


@pytest.mark.parametrize("lws", [(8, 8, 8), (32, 1, 1), (7, 3, 1), (1, 1, 1)])
def test_block_shape_invariance(lws):
    """The result does not depend on the work-group size (the reference's -w)."""
    g, o = _pair(16, 16, "f32", kernel=capi.LBM_KERNEL_GENERIC, lws=lws)
    with g:
        for it in range(1, 6):
            g.compute(it, update_macro=True, diagnostics=(it == 5))
            o.step(update_macro=True)
        _assert_macro_equal(g, o, 16, "f32", f"lws={lws}")
        mg, _ = g.read_diagnostics()
        assert abs(mg - o.diag[0]) <= 1e-12 * abs(o.diag[0])


def _fixture_run(golden_dir, n, iterations, every, thresholds):
    g = capi.Context(n, stride=n, precision="f32")
    w = len(str(iterations))
    with g:
        g.initialize()
        rows = []
        for it in range(0, iterations + 1):
            if it > 0:
                g.compute(it, update_macro=(it % every == 0))
            if it % every == 0:
                t = read_vti(os.path.join(golden_dir, "target_results", str(n), f"ldc.0.{it:0{w}d}.vti"))
                rows.append(metrics(t["rho"], t["v"], window(g.read_rho(), n), window_v(g.read_u(), n)))
    worst = np.max(np.array(rows), axis=0)
    print(f"n={n}: worst MSE_rho={worst[0]:.3e} MSE_u={worst[1]:.3e} MAE_rho={worst[2]:.3e} MAE_u={worst[3]:.3e}")
    for v, lim, name in zip(worst, thresholds, ("mse_rho", "mse_u", "mae_rho", "mae_u")):
        assert v <= lim, f"{name} {v} > {lim}"


def test_fixture_8(golden_dir):
    """make test8 (Makefile:73-78): 8^3, 10 iterations, every 1, against Sailfish."""
    _fixture_run(golden_dir, 8, 10, 1, (1e-13, 5e-15, 1e-6, 4e-7))


def test_fixture_32(golden_dir):
    """make test32 (Makefile:81-86): 32^3, 500 iterations, every 20, against Sailfish."""
    _fixture_run(golden_dir, 32, 500, 20, (2e-12, 3e-14, 4e-6, 1.2e-6))


def test_events_and_timing():
    with capi.Context(64, stride=64, profile=True) as g:
        g.initialize()
        g.compute_range(1, 5, every=5)
        g.read_rho()
        ev = g.events()
        names = [e[0] for e in ev]
        assert names[0] == "initialize" and names.count("compute") == 5 and names[-1] == "read_rho"
        kt, tt = g.kernels_time_ms(), g.total_time_ms()
        assert 0 < kt <= tt
        assert abs(kt - sum(ms for name, ms in ev if name == "compute")) < 1e-9


def test_event_log_over_many_launches():
    """Hundreds of launches: completed event pairs are read and their events
    reused; counts, order, per-event times and the totals stay consistent."""
    n, steps = 64, 700
    with capi.Context(n, stride=64, profile=True) as g:
        g.initialize()
        g.compute_range(1, steps, every=100)
        g.read_u()
        g.compute_range(steps + 1, 3, every=0)
        ev = g.events()
        names = [e[0] for e in ev]
        assert len(ev) == 1 + steps + 1 + 3
        assert names[0] == "initialize" and names[steps + 1] == "read_u" and names.count("compute") == steps + 3
        assert all(ms > 0 for _, ms in ev)
        kt, tt = g.kernels_time_ms(), g.total_time_ms()
        assert abs(kt - sum(ms for name, ms in ev if name == "compute")) < 1e-6 * kt
        assert sum(ms for _, ms in ev) <= tt * 1.0001
        g.clear_events()
        assert g.events() == [] and g.total_time_ms() == 0.0
        g.compute(steps + 4)
        assert len(g.events()) == 1 and g.kernels_time_ms() > 0


def test_state_roundtrip():
    """lbm_write_state restores what lbm_read_state returned."""
    with capi.Context(32, stride=32) as g:
        g.initialize()
        for it in range(1, 5):
            g.compute(it)
        s0, s1 = g.read_state(0), g.read_state(1)
        g.compute(5, update_macro=True)
        r5 = g.read_rho()
        g.write_state(0, s0)
        g.write_state(1, s1)
        g.compute(5, update_macro=True)
        assert nan_equal(g.read_rho(), r5)


def test_full_size_512_f32():
    """512^3 FP32, the bandwidth-scale configuration: 2 steps bit-identical to the oracle."""
    n = 512
    g = capi.Context(n, stride=n ** 3, precision="f32", arith=capi.LBM_ARITH_EXACT)
    o = O.PullSim(n, stride=n ** 3, precision="f32")
    with g:
        g.initialize()
        o.initialize()
        for it in (1, 2):
            g.compute(it, update_macro=True)
            o.step(update_macro=True)
        rg, ro = g.read_rho(), o.rho
        assert nan_equal(rg, ro), mismatch_count(rg, ro)
        ug = g.read_u()
        assert nan_equal(ug, o.u), mismatch_count(ug, o.u)


def test_full_size_1024_f32_vector_matches_generic():
    """1024^3 FP32 -- 163 GB of lattices, element offsets beyond 2^32 (19 N^3 =
    2.0e10), a size the reference's int index cannot address and the CPU oracle
    is too large for: the 8-byte-lane vector kernel (two blocks per row), the
    row-pipeline kernel (whole 1024-cell rows in LDS) and the
    one-thread-per-cell generic kernel, all bit-identical to the oracle at
    smaller sizes, agree bit for bit on rho and u after two steps, and on the
    FP64 mass and energy to their summation order."""
    n = 1024
    out = {}
    for kernel in (capi.LBM_KERNEL_VEC8, capi.LBM_KERNEL_GENERIC, capi.LBM_KERNEL_LDS):
        print(f"1024^3 FP32 stride 4096 kernel {kernel}", flush=True)
        with capi.Context(n, stride=4096, precision="f32", kernel=kernel) as g:
            g.initialize()
            g.compute(1)
            g.compute(2, update_macro=True, diagnostics=True)
            out[kernel] = (g.read_rho(), g.read_u(), g.read_diagnostics())
    (rv, uv, dv), (rg, ug, dg) = out[capi.LBM_KERNEL_VEC8], out[capi.LBM_KERNEL_GENERIC]
    rl, ul, dl = out.pop(capi.LBM_KERNEL_LDS)
    assert nan_equal(rl, rg), mismatch_count(rl, rg)
    assert nan_equal(ul, ug), mismatch_count(ul, ug)
    assert abs(dl[0] - dg[0]) <= 1e-12 * abs(dg[0]) and abs(dl[1] - dg[1]) <= 1e-10 * abs(dg[1])
    del rl, ul
    assert nan_equal(rv, rg), mismatch_count(rv, rg)
    assert int(np.isfinite(rv).sum()) > (n - 4) ** 3
    del rv, rg
    assert nan_equal(uv, ug), mismatch_count(uv, ug)
    assert abs(dv[0] - dg[0]) <= 1e-12 * abs(dg[0]) and abs(dv[1] - dg[1]) <= 1e-10 * abs(dg[1])


@pytest.mark.parametrize("n,stride,kernel", [(32, 32, capi.LBM_KERNEL_AUTO), (64, 64, capi.LBM_KERNEL_AUTO),
                                             (64, 1, capi.LBM_KERNEL_TILE), (64, 64, capi.LBM_KERNEL_GENERIC)])
def test_opencl_arithmetic_f64_matches_contracted_oracle(n, stride, kernel):
    """LBM_ARITH_OPENCL in FP64 (correctly rounded division, so the CPU can
    follow it): bit-identical to the oracle's contracted form, 100 steps."""
    g = capi.Context(n, stride=stride, precision="f64", kernel=kernel)  # the default: LBM_ARITH_OPENCL
    o = O.PullSim(n, stride=stride, precision="f64", fast=O.ARITH_CONTRACTED)
    with g:
        g.initialize()
        o.initialize()
        for it in range(1, 101):
            g.compute(it, update_macro=(it == 100))
            o.step(update_macro=(it == 100))
        _assert_macro_equal(g, o, n, "f64", f"OpenCL arithmetic FP64 {n}^3 stride {stride} kernel {kernel}")
        buf = 1 if 101 % 2 == 0 else 0
        assert nan_equal(g.read_state(buf), o.input_buffer(101))


def test_allocation_failure_is_an_error_and_leaves_the_device_usable():
    """1024^3 FP64 needs two 163 GB lattices, more than the GPU's HBM: lbm_create
    fails with LBM_ERR_MEMORY after releasing what it had allocated, and the next
    context computes correctly."""
    print("1024^3 FP64 (does not fit)", flush=True)
    with pytest.raises(capi.LbmError) as e:
        capi.Context(1024, stride=4096, precision="f64")
    assert e.value.code == capi.LBM_ERR_MEMORY and "hipMalloc" in str(e.value)
    g, o = _pair(16, 16, "f64")
    with g:
        g.compute(1, update_macro=True)
        o.step(update_macro=True)
        _assert_macro_equal(g, o, 16, "f64", "after a failed allocation")


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_cavity_128_1000_steps(precision):
    """The 128^3 GPU configuration: stride 64, 1000 iterations, rho and u after
    the last one bit-identical to the oracle at the same precision (and so
    within 1e-5 / 1e-12 relative)."""
    n, steps = 128, 1000
    g, o = _pair(n, 64, precision)
    with g:
        g.compute_range(1, steps - 1, 0)
        g.compute(steps, update_macro=True)
        for _ in range(steps - 1):
            o.step(update_macro=False)
        o.step(update_macro=True)
        _assert_macro_equal(g, o, n, precision, f"128^3 x {steps} {precision}")
