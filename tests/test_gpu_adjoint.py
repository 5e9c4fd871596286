"""The body force and the adjoint on the GPU.

* the body force: a zero field is bit-identical to no force; a random field
  matches the CPU oracle's forcing term bit for bit, with native and with
  16-bit storage;
* the adjoint: J from the adjoint run equals J from a plain forward run;
  checkpointed and full-tape runs agree exactly; gradients match FP64 central
  finite differences (h = 1e-6 relative) on seeded random directions of
  (U, nu, g(x)) -- the 64^3 x 500 configuration with a Gaussian random body
  force (correlation 8, amplitude 1e-6) and J = total kinetic energy, and the
  plane objective; the FP32 adjoint against the same FP64 differences, held
  to the same 1e-4;
* the large configuration: 256^3 x 2000 in FP32 with checkpoints every 45
  steps and J = mean u_x on the mid-plane z = 128 -- the adjoint costs at most
  3x the forward pass, fits in the GPU's memory, gives the same gradients with
  checkpoints every 50 steps, and agrees with FP64 central differences.
"""
import numpy as np
import pytest

import oracle as O
from lbmcl_amd import capi
from lbmcl_amd.adjoint import (CavityObjective, cavity_gradient, cavity_objective, gaussian_random_field)
from tests.util import nan_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,kernel", [("f32", capi.LBM_KERNEL_VEC8), ("f64", capi.LBM_KERNEL_VEC),
                                              ("f32", capi.LBM_KERNEL_GENERIC), ("f64", capi.LBM_KERNEL_ROW),
                                              ("f32", capi.LBM_KERNEL_TILE), ("f64", capi.LBM_KERNEL_TILE)])
def test_force_matches_oracle(precision, kernel):
    n = 128 if kernel != capi.LBM_KERNEL_GENERIC else 32
    stride = 2 if kernel == capi.LBM_KERNEL_TILE else n
    dt = np.float32 if precision == "f32" else np.float64
    g = gaussian_random_field(n, 4.0, 1e-4, seed=5).astype(dt)
    ex = capi.LBM_ARITH_EXACT  # the oracle's form
    with capi.Context(n, stride=stride, precision=precision, kernel=kernel, arith=ex) as a, \
            capi.Context(n, stride=stride, precision=precision, kernel=kernel, arith=ex) as b:
        o = O.PullSim(n, stride=stride, precision=precision)
        o.force = g.reshape(-1)
        a.set_force(np.zeros(3 * n ** 3, dtype=dt))
        b.set_force(g)
        a.initialize()
        b.initialize()
        o.initialize()
        ref = capi.Context(n, stride=stride, precision=precision, kernel=kernel, arith=ex)
        ref.initialize()
        for it in range(1, 7):
            a.compute(it, update_macro=True)
            b.compute(it, update_macro=True)
            ref.compute(it, update_macro=True)
            o.step(True)
        assert nan_equal(a.read_rho(), ref.read_rho()) and nan_equal(a.read_u(), ref.read_u())
        assert nan_equal(b.read_rho(), o.rho) and nan_equal(b.read_u(), o.u)
        assert not nan_equal(b.read_u(), ref.read_u())
        ref.close()


@pytest.mark.parametrize("storage,kernel,n,stride", [("fp16", capi.LBM_KERNEL_VEC8, 256, 4096),
                                                      ("bf16", capi.LBM_KERNEL_VEC8, 256, 256),
                                                      ("fp16", capi.LBM_KERNEL_TILE, 64, 8),
                                                      ("bf16", capi.LBM_KERNEL_GENERIC, 32, 32)])
def test_force_16bit_storage_matches_oracle(storage, kernel, n, stride):
    """The body force with 16-bit storage: the kernels add the force term to the
    scaled deviations (times the storage scale), the oracle to the unscaled
    ones; the stored lattices and rho/u agree bit for bit."""
    g = gaussian_random_field(n, 4.0, 1e-4, seed=7).astype(np.float32)
    st = O.STORE_FP16 if storage == "fp16" else O.STORE_BF16
    with capi.Context(n, stride=stride, storage=storage, kernel=kernel) as b:
        o = O.PullSim(n, stride=stride, storage=st)
        o.force = g.reshape(-1)
        b.set_force(g)
        b.initialize()
        o.initialize()
        for it in range(1, 6):
            b.compute(it, update_macro=True)
            o.step(True)
        assert nan_equal(b.read_rho(), o.rho) and nan_equal(b.read_u(), o.u)
        for buf, ob in ((0, o.A), (1, o.B)):
            assert np.array_equal(b.read_state(buf).view(np.uint16), ob.view(np.uint16)), f"lattice {buf}"


def _direction(n, rng):
    return rng.standard_normal(), rng.standard_normal(), gaussian_random_field(n, 8.0, 1.0, int(rng.integers(1 << 30)))


def _fd(n, steps, U, nu, g, dU, dnu, dg, h, **kw):
    jp = cavity_objective(n, steps, U + h * U * dU, nu + h * nu * dnu, g + h * 1e-6 * dg, **kw)
    jm = cavity_objective(n, steps, U - h * U * dU, nu - h * nu * dnu, g - h * 1e-6 * dg, **kw)
    return (jp - jm) / (2 * h)


def _adjoint_directional(r, U, nu, dU, dnu, dg):
    return r["d_velocity"] * U * dU + r["d_viscosity"] * nu * dnu + 1e-6 * float(np.sum(r["d_force"] * dg))


@pytest.mark.parametrize("precision,n,stride", [("f32", 128, 128), ("f32", 128, 2), ("f64", 64, 64),
                                               ("f64", 128, 16), ("f32", 64, 64), ("f32", 128, 1)])
@pytest.mark.parametrize("objective", ["kinetic_energy", "mean_ux_plane"])
def test_vector_adjoint_matches_generic(precision, n, stride, objective):
    """The vector adjoint kernel (one block per row, DPP/LDS scatter; two cells
    per lane in FP32 at N >= 128 and stride >= 2, else one) against the generic
    one-thread-per-cell kernel: per-cell gradients bit-identical, the FP64 sums
    equal up to their summation order; checkpointed == full tape."""
    steps = 30
    g = gaussian_random_field(n, 8.0, 1e-6, seed=7)
    kw = dict(force=g, objective=objective, precision=precision, stride=stride)
    rv = cavity_gradient(n, steps, checkpoint_every=7, **kw)
    rf = cavity_gradient(n, steps, checkpoint_every=steps, **kw)
    rg = cavity_gradient(n, steps, checkpoint_every=7, kernel=capi.LBM_KERNEL_GENERIC, **kw)
    assert np.array_equal(rv["d_force"], rg["d_force"])
    assert np.count_nonzero(rv["d_force"]) > 0
    for k in ("objective", "d_velocity", "d_viscosity"):
        assert abs(rv[k] - rg[k]) <= 1e-9 * abs(rg[k]), (k, rv[k], rg[k])
        assert rv[k] == rf[k]
    assert np.array_equal(rv["d_force"], rf["d_force"])


def test_objective_and_checkpointing():
    n, steps = 32, 60
    g = gaussian_random_field(n, 8.0, 1e-6, seed=11)
    full = cavity_gradient(n, steps, force=g, checkpoint_every=steps)
    ck = cavity_gradient(n, steps, force=g, checkpoint_every=7)
    ck1 = cavity_gradient(n, steps, force=g, checkpoint_every=1)
    j = cavity_objective(n, steps, force=g)
    assert abs(full["objective"] - j) <= 1e-12 * abs(j)
    for r in (ck, ck1):
        assert r["objective"] == full["objective"]
        assert r["d_velocity"] == full["d_velocity"] and r["d_viscosity"] == full["d_viscosity"]
        assert np.array_equal(r["d_force"], full["d_force"])
    assert ck["checkpoints"] == 9 and ck["segment"] == 7


@pytest.mark.parametrize("objective", ["kinetic_energy", "mean_ux_plane"])
def test_gradient_matches_fp64_finite_differences(objective):
    """64^3, T = 500, g from a seeded Gaussian random field; 16 seeded directions."""
    n, steps, U, nu = 64, 500, 0.05, 0.0089
    g = gaussian_random_field(n, 8.0, 1e-6, seed=1)
    kw = dict(objective=objective)
    r64 = cavity_gradient(n, steps, U, nu, g, precision="f64", **kw)
    r32 = cavity_gradient(n, steps, U, nu, g, precision="f32", **kw)
    rng = np.random.default_rng(123)
    worst64 = worst32 = 0.0
    ndir = 16
    for d in range(ndir):
        dU, dnu, dg = _direction(n, rng)
        fd = _fd(n, steps, U, nu, g, dU, dnu, dg, 1e-6, **kw)
        a64 = _adjoint_directional(r64, U, nu, dU, dnu, dg)
        a32 = _adjoint_directional(r32, U, nu, dU, dnu, dg)
        e64, e32 = abs(a64 - fd) / abs(fd), abs(a32 - fd) / abs(fd)
        worst64, worst32 = max(worst64, e64), max(worst32, e32)
        print(f"{objective} dir {d}: fd {fd:+.10e} adj64 {a64:+.10e} ({e64:.2e}) adj32 {a32:+.10e} ({e32:.2e})")
    print(f"{objective}: worst relative error FP64 {worst64:.3e}, FP32 {worst32:.3e}; "
          f"adjoint/forward time {r64['adjoint_ms'] / r64['forward_ms']:.2f} (FP64)")
    assert worst64 < 1e-4
    assert worst32 < 1e-4
    # the body-force part alone (directions in g only; larger step, J is nearly linear in g)
    for d in range(4):
        dg = gaussian_random_field(n, 8.0, 1.0, 1000 + d)
        fd = _fd(n, steps, U, nu, g, 0.0, 0.0, dg, 1e-3, **kw)
        a64 = _adjoint_directional(r64, U, nu, 0.0, 0.0, dg)
        print(f"{objective} g-only dir {d}: fd {fd:+.10e} adj64 {a64:+.10e} ({abs(a64 - fd) / abs(fd):.2e})")
        assert abs(a64 - fd) <= 1e-4 * abs(fd)


def test_large_checkpointed_adjoint_cost_and_gradient():
    """256^3 x 2000, FP32 forward and adjoint (FP64 gradient sums), K = 45,
    J = mean u_x on z = 128, seeded Gaussian random body force."""
    torch = pytest.importorskip("torch")
    n, steps, U, nu = 256, 2000, 0.05, 0.0089
    g = gaussian_random_field(n, 8.0, 1e-6, seed=0)
    kw = dict(objective="mean_ux_plane", plane_z=128)
    print("256^3 x 2000 FP32 adjoint, checkpoints every 45 steps", flush=True)
    r45 = cavity_gradient(n, steps, U, nu, g.astype(np.float32), precision="f32", checkpoint_every=45, **kw)
    ratio = r45["adjoint_ms"] / r45["forward_ms"]
    hbm = torch.cuda.get_device_properties(0).total_memory
    print(f"forward {r45['forward_ms']:.0f} ms, adjoint {r45['adjoint_ms']:.0f} ms ({ratio:.2f}x), "
          f"{r45['checkpoints']} checkpoints, {r45['device_bytes'] / 2 ** 30:.1f} GiB of {hbm / 2 ** 30:.0f} GiB",
          flush=True)
    assert r45["checkpoints"] == 45 and r45["segment"] == 45
    assert ratio <= 3.0
    assert r45["device_bytes"] < hbm
    print("the same with checkpoints every 50 steps", flush=True)
    r50 = cavity_gradient(n, steps, U, nu, g.astype(np.float32), precision="f32", checkpoint_every=50, **kw)
    for k in ("objective", "d_velocity", "d_viscosity"):
        assert r50[k] == r45[k], k
    assert np.array_equal(r50["d_force"], r45["d_force"])
    rng = np.random.default_rng(2024)
    for d in range(2):
        dU, dnu, dg = _direction(n, rng)
        print(f"FP64 central differences, direction {d}", flush=True)
        fd = _fd(n, steps, U, nu, g, dU, dnu, dg, 1e-6, precision="f64", **kw)
        a32 = _adjoint_directional(r45, U, nu, dU, dnu, dg)
        err = abs(a32 - fd) / abs(fd)
        print(f"direction {d}: fd {fd:+.10e} adjoint (FP32) {a32:+.10e} ({err:.2e})", flush=True)
        assert err < 1e-4


def test_autograd_function():
    torch = pytest.importorskip("torch")
    n, steps = 16, 30
    g = torch.from_numpy(gaussian_random_field(n, 4.0, 1e-6, seed=3))
    U = torch.tensor(0.05, dtype=torch.float64, requires_grad=True)
    nu = torch.tensor(0.0089, dtype=torch.float64, requires_grad=True)
    gg = g.clone().requires_grad_(True)
    J = CavityObjective.apply(U, nu, gg, dict(dim=n, steps=steps, precision="f64"))
    J.backward()
    r = cavity_gradient(n, steps, 0.05, 0.0089, g.numpy(), precision="f64")
    assert float(J) == r["objective"]
    assert float(U.grad) == r["d_velocity"] and float(nu.grad) == r["d_viscosity"]
    assert np.array_equal(gg.grad.numpy(), r["d_force"])
