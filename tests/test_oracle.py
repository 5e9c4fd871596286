"""Pinning the CPU oracle (CPU only).

The oracle is the checker of the HIP kernels, so it is itself checked against
everything that fixes the reference's behaviour independently of it:

* the reference's own golden data: the Sailfish outputs in target_results/8
  and /32 (make test8 / make test32, Makefile:73-86), with verify.py's metrics
  and explicit thresholds;
* closed forms derived by hand for step 2 of the 8^3 cavity;
* the constant the reference states in a comment (INV_TAU, kernels.cl:62);
* invariants: layout (stride) invariance, y-mirror symmetry, the U = 0 rest
  state, push form == pull form bit for bit;
* planted mistakes: oracle variants with one plausible error each must fail
  the fixture check;
* the 16-bit codecs against numpy (binary16) and torch (bfloat16);
* the contracted form (kernels.cl as clang compiles OpenCL C): push == pull,
  and within rounding of the exact form.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from lbmcl_amd.verify import metrics
from lbmcl_amd.vti import read_vti
from tests.util import max_rel, nan_equal, window, window_v


def fixture_errors(golden_dir, n, iterations, every, sim_factory=None):
    sim = (sim_factory or (lambda: O.PushSim(n, stride=n)))()
    sim.initialize()
    w = len(str(iterations))
    rows = []
    for it in range(0, iterations + 1):
        if it > 0:
            sim.step(update_macro=(it % every == 0))
        if it % every == 0:
            t = read_vti(os.path.join(golden_dir, "target_results", str(n), f"ldc.0.{it:0{w}d}.vti"))
            rho, u = sim.macro()
            rows.append(metrics(t["rho"], t["v"], window(rho, n), window_v(u.reshape(-1), n)))
    return np.max(np.array(rows), axis=0)


# worst (MSE rho, MSE u, MAE rho, MAE u) over the dumped iterations; the oracle
# measures 1.3e-14 / 1.0e-15 / 5.4e-7 / 1.9e-7 (8^3) and 6.8e-13 / 1.1e-14 /
# 2.1e-6 / 5.8e-7 (32^3): FP32 noise between two different codes.
LIMITS = {8: (1e-13, 5e-15, 1e-6, 4e-7), 32: (2e-12, 3e-14, 4e-6, 1.2e-6)}


def test_fixture_8(golden_dir):
    worst = fixture_errors(golden_dir, 8, 10, 1)
    assert all(v <= lim for v, lim in zip(worst, LIMITS[8])), worst


def test_fixture_32(golden_dir):
    worst = fixture_errors(golden_dir, 32, 500, 20)
    assert all(v <= lim for v, lim in zip(worst, LIMITS[32])), worst


def test_fixture_8_pull_form(golden_dir):
    worst = fixture_errors(golden_dir, 8, 10, 1, lambda: O.PullSim(8, stride=8))
    assert all(v <= lim for v, lim in zip(worst, LIMITS[8])), worst


def test_closed_forms_step2():
    """Hand-derived values at VTI 2 of the 8^3 cavity (after one collide-stream step)."""
    n = 8
    s = O.PushSim(n, stride=n)
    s.initialize()
    s.step()
    s.step()
    r = s.rho.reshape(n, n, n)
    u = s.u.reshape(3, n, n, n)
    U = 0.05
    for y in range(3, n - 3):  # rows away from the y edges
        assert abs(r[n - 2, y, 2] - 29 / 30) < 1e-6      # lid cell next to the x=1 edge: 1 - 1/30
        assert abs(r[n - 2, y, n - 3] - 31 / 30) < 1e-6  # lid cell next to the x=N-2 edge: 1 + 1/30
        assert abs(r[n - 3, y, 2] - 119 / 120) < 1e-6    # fluid under the lid edge
        assert abs(r[n - 3, y, n - 3] - 121 / 120) < 1e-6
        for x in range(3, n - 3):
            assert abs(u[0, n - 3, y, x] - U / 6) < 1e-8  # fluid under the lid interior
    # the same values are in the Sailfish fixture
    t = read_vti(os.path.join(os.path.dirname(__file__), "golden", "target_results", "8", "ldc.0.02.vti"))
    tr = t["rho"].reshape(6, 6, 6)
    assert abs(tr[5, 2, 1] - 29 / 30) < 1e-6 and abs(tr[4, 2, 1] - 119 / 120) < 1e-6
    assert abs(tr[5, 3, 4] - 31 / 30) < 1e-6


def test_inv_tau():
    inv64 = O.constants(O.params(8, 8, 0.0089, 0.05, "f64"), "f64")[0]
    assert abs(inv64 - 1.89861401177140698415) < 1e-15  # kernels.cl:62
    inv32 = O.constants(O.params(8, 8, 0.0089, 0.05, "f32"), "f32")[0]
    nu = np.float32(0.0089)
    assert inv32 == np.float32(1) / (np.float32(3) * nu + np.float32(0.5))


def test_kernel_literal_rounding():
    """-D values carry 6 significant digits of the T-typed value (lbmcl.hpp:140-141)."""
    assert O.kernel_literal(0.0089, "f64") == 0.0089
    assert O.kernel_literal(0.0123456789, "f64") == 0.0123457
    assert O.kernel_literal(0.0123456789, "f32") == float(np.float32(0.0123457))
    assert O.kernel_literal(1.0 / 3.0, "f32") == float(np.float32(0.333333))


def test_cell_types():
    n = 8
    t = np.array([[[O.cell_type(n, x, y, z) for x in range(n)] for y in range(n)] for z in range(n)])
    assert (t[0] == O.WALL).all() and (t[:, :, n - 1] == O.WALL).all()
    # back corners (z = 1) are CORNER, front corners are bounce-back (common.h:43-66)
    for x in (1, n - 2):
        for y in (1, n - 2):
            assert t[1, y, x] == O.CORNER
            assert t[n - 2, y, x] & O.FRONT and not t[n - 2, y, x] & O.MOVING
    assert t[n - 2, 3, 3] == O.FRONT | O.MOVING
    assert t[3, 3, 3] == O.FLUID
    assert t[3, 1, 1] == O.LEFT | O.BOTTOM
    # NaN pattern of the VTI window matches the fixtures: 136 of 216 at 8^3
    s = O.PushSim(n, stride=n)
    s.initialize()
    assert int(np.isnan(window(s.rho, n)).sum()) == 136


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_stride_invariance(precision):
    n = 16
    ref = None
    for stride in (1, 2, 16, 64, 4096):
        s = O.PushSim(n, stride=stride, precision=precision)
        s.initialize()
        for _ in range(12):
            s.step()
        if ref is None:
            ref = (s.rho.copy(), s.u.copy())
        else:
            assert nan_equal(s.rho, ref[0]) and nan_equal(s.u, ref[1]), stride


@pytest.mark.parametrize("precision,tol", [("f32", 2e-6), ("f64", 1e-14)])
def test_y_mirror_symmetry(precision, tol):
    n = 16
    s = O.PushSim(n, stride=n, precision=precision)
    s.initialize()
    for _ in range(100):
        s.step()
    r = s.rho.reshape(n, n, n)
    u = s.u.reshape(3, n, n, n)
    assert np.nanmax(np.abs(r - r[:, ::-1, :])) < tol
    assert np.nanmax(np.abs(u[0] - u[0][:, ::-1, :])) < tol
    assert np.nanmax(np.abs(u[1] + u[1][:, ::-1, :])) < tol
    assert np.nanmax(np.abs(u[2] - u[2][:, ::-1, :])) < tol
    assert np.nanmax(np.abs(u[0])) > 1e-3  # the flow is not trivially symmetric


@pytest.mark.parametrize("precision,tol", [("f32", 2e-6), ("f64", 1e-13)])
def test_rest_state(precision, tol):
    """With a resting lid the cavity stays at rho = 1, u = 0 to rounding."""
    s = O.PushSim(16, stride=16, precision=precision, velocity=0.0)
    s.initialize()
    for _ in range(50):
        s.step()
    w = np.isfinite(s.rho)
    assert np.max(np.abs(s.rho[w] - 1)) < tol
    assert np.nanmax(np.abs(s.u)) < tol


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("n,stride", [(8, 8), (16, 1), (16, 4096), (32, 64)])
def test_push_equals_pull(precision, n, stride):
    a = O.PushSim(n, stride=stride, precision=precision)
    b = O.PullSim(n, stride=stride, precision=precision)
    a.initialize()
    b.initialize()
    for it in range(1, 21):
        assert nan_equal(a.input_buffer(it), b.view(b.input_buffer(it))), it
        a.step(True)
        b.step(True)
        assert nan_equal(a.rho, b.rho) and nan_equal(a.u, b.u), it


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_contracted_form(precision):
    """kernels.cl's expressions with clang's OpenCL contractions (the form of the
    reference's kernels on AMD GPUs): push == pull bit for bit as in the exact
    form; the contractions do change bits, by rounding only."""
    n, stride = 16, 16
    a = O.PushSim(n, stride=stride, precision=precision, fast=O.ARITH_CONTRACTED)
    b = O.PullSim(n, stride=stride, precision=precision, fast=O.ARITH_CONTRACTED)
    c = O.PullSim(n, stride=stride, precision=precision)
    for s in (a, b, c):
        s.initialize()
    for it in range(1, 51):
        assert nan_equal(a.input_buffer(it), b.view(b.input_buffer(it))), it
        a.step(True)
        b.step(True)
        c.step(True)
        assert nan_equal(a.rho, b.rho) and nan_equal(a.u, b.u), it
    assert not nan_equal(b.u, c.u)
    tol = 1e-5 if precision == "f32" else 1e-13
    assert max_rel(b.rho, c.rho) <= tol and max_rel(b.u, c.u) <= 10 * tol


def test_force_zero_is_bit_identical_and_linear():
    n = 16
    a = O.PullSim(n, stride=n)
    b = O.PullSim(n, stride=n)
    b.force = np.zeros(3 * n ** 3, dtype=np.float32)
    a.initialize()
    b.initialize()
    for _ in range(10):
        a.step()
        b.step()
    assert nan_equal(a.rho, b.rho) and nan_equal(a.u, b.u)
    # the force term injects momentum 3 w_i e_i g summed over i = g, and no mass
    c = O.PullSim(n, stride=n, precision="f64")
    d = O.PullSim(n, stride=n, precision="f64")
    d.force = np.zeros(3 * n ** 3)
    d.force[n ** 3 + 8 + 8 * n + 8 * n * n] = 1e-6   # g_y at cell (8, 8, 8)
    c.initialize()
    d.initialize()
    c.step()
    d.step()
    pc, pd = c.view(c.input_buffer(2)), d.view(d.input_buffer(2))
    delta = (pd - pc).reshape(-1)
    fin = np.isfinite(delta)
    dm = delta[fin].sum()
    # momentum: sum over the 19 neighbours of e_i,y * delta
    mom = 0.0
    for i in range(19):
        sl = [O.slot(c_id, i, n) for c_id in range(n ** 3)]
        mom += O.EY[i] * np.nansum(delta[sl])
    # exact up to the FP64 rounding of f_i + F_i (|f_i| < 1)
    assert abs(dm) < 1e-15 and abs(mom - 1e-6) < 1e-15


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_diagnostics_against_numpy(precision):
    n = 16
    s = O.PullSim(n, stride=n, precision=precision)
    s.initialize()
    for _ in range(7):
        s.step(True)
    cells = np.isfinite(s.rho)
    r = s.rho[cells].astype(np.float64)
    u = s.u.reshape(3, -1)[:, cells].astype(np.float64)
    mass, ke = r.sum(), (0.5 * r * (u ** 2).sum(axis=0)).sum()
    m, e = O.diagnostics(s.p, s.rho, s.u, precision)
    assert abs(m - mass) < 1e-12 * mass and abs(e - ke) < 1e-12 * ke
    assert abs(s.diag[0] - mass) < 1e-12 * mass and abs(s.diag[1] - ke) < 1e-10 * ke


def test_fp16_codec_matches_numpy():
    rng = np.random.default_rng(7)
    vals = np.concatenate([
        rng.standard_normal(200000).astype(np.float32) * np.float32(10.0) ** rng.integers(-9, 5, 200000),
        np.array([0.0, -0.0, 65504, 65519.99, 65520, 1e9, -1e9, np.inf, -np.inf, 6.1e-5, 5.96e-8, 2.98e-8,
                  2.99e-8, 1e-10, 0.5 + 2 ** -12, 1 + 2 ** -11, 1 + 3 * 2 ** -11], dtype=np.float32),
    ]).astype(np.float32)
    ours = O.f32_to_f16(vals)
    with np.errstate(over="ignore"):
        ref = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref)
    back = O.f16_to_f32(ref)
    assert np.array_equal(back, ref.view(np.float16).astype(np.float32))
    nan = O.f32_to_f16(np.array([np.nan], dtype=np.float32))
    assert np.isnan(O.f16_to_f32(nan))[0]


def test_bf16_codec_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(8)
    vals = (rng.standard_normal(200000) * 10.0 ** rng.integers(-30, 30, 200000)).astype(np.float32)
    vals = np.concatenate([vals, np.array([0.0, -0.0, np.inf, -np.inf, 1 + 2 ** -8, 1 + 3 * 2 ** -8], np.float32)])
    ours = O.f32_to_bf16(vals)
    ref = torch.from_numpy(vals).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(ours, ref)
    assert np.array_equal(O.bf16_to_f32(ref), torch.from_numpy(ref.view(np.int16)).view(torch.bfloat16).float().numpy())


@pytest.mark.parametrize("storage", [O.STORE_FP16, O.STORE_BF16])
def test_shifted_storage_tracks_fp32(storage):
    """The 16-bit shifted storage stays close to the FP32 path (it is not exact)."""
    n = 16
    a = O.PullSim(n, stride=n)
    b = O.PullSim(n, stride=n, storage=storage)
    a.initialize()
    b.initialize()
    for _ in range(40):
        a.step()
        b.step()
    tol = 2e-4 if storage == O.STORE_FP16 else 5e-3
    assert np.nanmax(np.abs(a.rho - b.rho)) < tol
    assert np.nanmax(np.abs(a.u - b.u)) < tol * 0.05 / 0.02 + 1e-5


def _collide_pair(t, f, force=None):
    import ctypes
    p = O.params(32, 32, 0.0089, 0.05, "f32")
    f = np.ascontiguousarray(f, dtype=np.float64)
    g = None if force is None else np.ascontiguousarray(force, dtype=np.float32)
    full, delta, macro = np.empty(19), np.empty(19, np.float32), np.empty(8)
    O.lib().lbmo_collide_pair(ctypes.byref(p), t, O._ptr(f), None if g is None else O._ptr(g), O._ptr(full),
                              O._ptr(delta), O._ptr(macro))
    return full, delta, macro


def test_deviation_form_equals_full_value_form():
    """The 16-bit storage forms' step works on d_i = f_i - w_i in FP32
    (collide_delta).  For every cell type -- FLUID, lid, bounce-back, corner --
    on random near-equilibrium populations, with and without a body force, w_i
    + d_i' equals the reference's expressions evaluated in FP64 on the full
    values, and rho, u agree, to FP32 rounding (1e-7): a wrong sign, weight or
    dropped term (the rho - 1 part of the equilibrium, the force) is off by
    1e-4 or more."""
    rng = np.random.default_rng(7)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    w32 = np.array([np.float32(1) / np.float32(3)] + [np.float32(1) / np.float32(18)] * 6
                   + [np.float32(1) / np.float32(36)] * 12, dtype=np.float64)
    types = [O.FLUID, O.FRONT | O.MOVING, O.LEFT, O.LEFT | O.TOP, O.CORNER]
    for trial in range(100):
        rho = 1 + 0.01 * rng.standard_normal()
        u = 0.04 * rng.standard_normal(3)
        eu = O.EX * u[0] + O.EY * u[1] + O.EZ * u[2]
        feq = rho * w * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * (u @ u))
        f = feq + 1e-3 * w * rng.standard_normal(19)
        force = None if trial % 2 else 1e-5 * rng.standard_normal(3)
        for t in types:
            full, delta, macro = _collide_pair(t, f, force)
            assert np.max(np.abs((w32 + delta.astype(np.float64)) - full)) < 1e-7, (trial, t)
            assert abs(macro[0] - macro[4]) < 1e-7 and np.max(np.abs(macro[1:4] - macro[5:8])) < 1e-7, (trial, t)


def test_deviation_form_keeps_the_mass():
    """FP16 storage, 32^3 cavity, 2000 steps: the total mass stays within 1e-5
    of the FP64 run's.  (The full-value FP32 step on 16-bit-decoded
    populations gained 3.3e-7 of the mass per step, 6.6e-4 here: every cell of
    a smooth flow rounded values that share their low bits with w_i.)"""
    n, steps = 32, 2000
    ref = O.PullSim(n, stride=n, precision="f64", threads=8)
    h = O.PullSim(n, stride=n, storage=O.STORE_FP16, threads=8)
    ref.initialize()
    h.initialize()
    for _ in range(steps):
        ref.step(update_macro=False)
        h.step(update_macro=False)
    m64, m16 = ref.diag[0], h.diag[0]
    assert abs(m16 - m64) / m64 < 1e-5, (m16, m64)


def rule_violations(n=8, seed=3):
    """One push step from a seeded random lattice: count the slots where a corner
    cell does not pass its populations straight on, or a bounce-back cell does
    not send each population's opposite (kernels.cl:400-408, 420-425; corners
    have neither boundary condition nor collision, kernels.cl:254-260).

    The Sailfish fixtures cannot pin the corner rule: next to the back corners
    the flow stays at rest to FP32 resolution, where f_i equals f_opp(i)."""
    sim = O.PushSim(n, stride=n, precision="f64")
    sim.initialize()
    rng = np.random.default_rng(seed)
    sim.f_collide[:] = rng.uniform(0.01, 0.1, sim.f_collide.size)
    sim.step()
    bad = 0
    for z in range(n):
        for y in range(n):
            for x in range(n):
                t = O.cell_type(n, x, y, z)
                rule = "pass" if t == O.CORNER else ("swap" if (t & 0x3f0) and not t & O.MOVING else None)
                if rule is None:
                    continue
                c = x + y * n + z * n * n
                for i in range(19):
                    d = (x + O.EX[i]) + (y + O.EY[i]) * n + (z + O.EZ[i]) * n * n
                    src = i if rule == "pass" else O.OPP[i]
                    if sim.f_stream[O.slot(d, i, n)] != sim.f_collide[O.slot(c, src, n)]:
                        bad += 1
    return bad


def test_corner_and_bounceback_rules():
    assert rule_violations() == 0


MUTANTS = {1: "lid fix-up partner", 2: "corners bounce back", 3: "wrong opposite", 4: "tau for 1/tau",
           5: "population in the wrong x-moment", 6: "lid edges start at rest"}


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_mistakes_are_caught(golden_dir, tmp_path, mutant):
    """Each plausible mistake, planted in a copy of the oracle, fails the 8^3
    fixture check or the one-step rule check."""
    so = tmp_path / f"liblbm_oracle_m{mutant}.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
                    f"-DLBMO_MUTANT={mutant}", os.path.join(O.HERE, "lbm_oracle.cpp"), "-o", str(so)], check=True)
    import ctypes
    saved = O._lib
    try:
        O._lib = ctypes.CDLL(str(so))
        O._declare(O._lib)
        worst = fixture_errors(golden_dir, 8, 10, 1)
        violations = rule_violations()
    finally:
        O._lib = saved
    assert any(v > lim for v, lim in zip(worst, LIMITS[8])) or violations > 0, (MUTANTS[mutant], worst)
