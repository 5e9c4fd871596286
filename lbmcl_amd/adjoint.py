"""Gradients of lid-driven-cavity objectives, and their PyTorch wrapper.

The cavity is run for T iterations from its initial state with lid velocity U,
viscosity nu and a per-cell body force g(x); the objective J is evaluated on
the state after T iterations:

  "kinetic_energy"  J = sum over FLUID/MOVING cells of rho |u|^2 / 2
  "mean_ux_plane"   J = mean u_x over the FLUID/MOVING cells of the plane z = plane_z

`cavity_gradient` returns J with dJ/dU, dJ/dnu and dJ/dg(x) from the HIP
adjoint (lbm_adjoint: reverse collide-stream step, checkpoints every K ~ sqrt(T)
steps, FP64 gradient sums).  `cavity_objective` evaluates J alone with a plain
forward run (the finite-difference reference), and `CavityObjective` wraps the
pair as a torch.autograd.Function of (U, nu, g).

`gaussian_random_field` draws the seeded body-force fields: white noise from
numpy's PCG64 generator, filtered in Fourier space by exp(-k^2 l^2 / 2) with
correlation length l, scaled to the requested standard deviation per component.
"""
from __future__ import annotations

import numpy as np

from . import capi

OBJECTIVES = {"kinetic_energy": capi.LBM_OBJECTIVE_KINETIC_ENERGY,
              "mean_ux_plane": capi.LBM_OBJECTIVE_MEAN_UX_PLANE}


def gaussian_random_field(n, correlation=8.0, amplitude=1e-6, seed=0, components=3):
    """(components, n^3) float64 field, periodic, x fastest; std `amplitude` per component."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k = 2 * np.pi * np.fft.fftfreq(n)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    filt = np.exp(-0.5 * (kx ** 2 + ky ** 2 + kz ** 2) * correlation ** 2)
    out = np.empty((components, n ** 3))
    for c in range(components):
        w = rng.standard_normal((n, n, n))
        f = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
        f *= amplitude / f.std()
        out[c] = f.reshape(-1)
    return out


def _context(dim, velocity, viscosity, precision, stride, kernel=0):
    # default CSoA stride: 4096 (the fastest measured), at most dim^3
    stride = stride or min(4096, dim ** 3)
    return capi.Context(dim, viscosity=viscosity, velocity=velocity, stride=stride, precision=precision,
                        profile=False, raw_parameters=True, kernel=kernel)


def cavity_objective(dim, steps, velocity=0.05, viscosity=0.0089, force=None, objective="kinetic_energy",
                     plane_z=None, precision="f64", stride=None, kernel=0):
    """J by a plain forward run, evaluated in FP64 from the macroscopic fields."""
    plane_z = dim // 2 if plane_z is None else plane_z
    with _context(dim, velocity, viscosity, precision, stride, kernel) as g:
        if force is not None:
            g.set_force(np.asarray(force).reshape(-1))
        g.initialize()
        g.compute_range(1, steps, 0)
        g.compute(steps + 1, update_macro=True)
        rho = g.read_rho().astype(np.float64)
        u = g.read_u().astype(np.float64).reshape(3, -1)
    cells = np.isfinite(rho)
    if objective == "kinetic_energy":
        return float(np.sum(0.5 * rho[cells] * np.sum(u[:, cells] ** 2, axis=0)))
    z = np.arange(dim ** 3) // (dim * dim)
    sel = cells & (z == plane_z)
    return float(np.mean(u[0, sel]))


def cavity_gradient(dim, steps, velocity=0.05, viscosity=0.0089, force=None, objective="kinetic_energy",
                    plane_z=None, precision="f64", stride=None, checkpoint_every=0, grad_force=True, kernel=0):
    """dict: objective J, d_velocity, d_viscosity, d_force (3, dim^3), timings and memory.

    kernel: capi.LBM_KERNEL_* of the context; LBM_KERNEL_GENERIC or LBM_KERNEL_ROW
    select the generic (one thread per cell) adjoint kernel, anything else the
    vector adjoint kernel where the lattice allows it."""
    plane_z = dim // 2 if plane_z is None else plane_z
    with _context(dim, velocity, viscosity, precision, stride, kernel) as g:
        if force is not None:
            g.set_force(np.asarray(force).reshape(-1))
        return g.adjoint(steps, OBJECTIVES[objective], plane_z, checkpoint_every, grad_force)


try:
    import torch

    class CavityObjective(torch.autograd.Function):
        """J(U, nu, g) of a cavity run; backward returns the adjoint gradients.

        Inputs are CPU tensors: U and nu scalars, g of shape (3, dim^3) (or None
        for no body force).  `options` holds the remaining arguments of
        cavity_gradient (dim, steps, objective, ...)."""

        @staticmethod
        def forward(ctx, velocity, viscosity, force, options):
            r = cavity_gradient(velocity=float(velocity), viscosity=float(viscosity),
                                force=None if force is None else force.detach().cpu().numpy(), **options)
            ctx.grads = (r["d_velocity"], r["d_viscosity"], r["d_force"])
            ctx.has_force = force is not None
            return torch.tensor(r["objective"], dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            dU, dnu, dg = ctx.grads
            gU = grad_out * torch.tensor(dU, dtype=torch.float64)
            gnu = grad_out * torch.tensor(dnu, dtype=torch.float64)
            gg = grad_out * torch.from_numpy(dg) if (ctx.has_force and dg is not None) else None
            return gU, gnu, gg, None

except ImportError:  # torch is optional for the rest of the package
    CavityObjective = None


def main(argv=None):
    """Large-cavity gradient run: prints J, the gradients' norms, timings and memory as JSON."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--dim", type=int, default=256)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--precision", default="f32")
    ap.add_argument("--objective", default="mean_ux_plane", choices=sorted(OBJECTIVES))
    ap.add_argument("--plane-z", type=int, default=None)
    ap.add_argument("--checkpoint-every", type=int, default=0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--fd-directions", type=int, default=0,
                    help="also compare with FP64 central differences (h = 1e-6 relative) on this many directions")
    a = ap.parse_args(argv)
    g = gaussian_random_field(a.dim, 8.0, 1e-6, a.seed)
    r = cavity_gradient(a.dim, a.steps, force=g, objective=a.objective, plane_z=a.plane_z, precision=a.precision,
                        checkpoint_every=a.checkpoint_every)
    dg = r.pop("d_force")
    r["d_force_norm"] = float(np.linalg.norm(dg))
    r["adjoint_over_forward"] = r["adjoint_ms"] / r["forward_ms"]
    r["J"] = r.pop("objective")
    r.update(dim=a.dim, steps=a.steps, precision=a.precision, objective_kind=a.objective)
    if a.fd_directions:
        rng = np.random.default_rng(a.seed + 100)
        U, nu, h = 0.05, 0.0089, 1e-6
        errs = []
        for _ in range(a.fd_directions):
            dU, dnu = rng.standard_normal(), rng.standard_normal()
            d = gaussian_random_field(a.dim, 8.0, 1.0, int(rng.integers(1 << 30)))
            jp = cavity_objective(a.dim, a.steps, U + h * U * dU, nu + h * nu * dnu, g + h * 1e-6 * d,
                                  a.objective, a.plane_z, "f64")
            jm = cavity_objective(a.dim, a.steps, U - h * U * dU, nu - h * nu * dnu, g - h * 1e-6 * d,
                                  a.objective, a.plane_z, "f64")
            fd = (jp - jm) / (2 * h)
            adj = r["d_velocity"] * U * dU + r["d_viscosity"] * nu * dnu + 1e-6 * float(np.sum(dg * d))
            errs.append(abs(adj - fd) / abs(fd))
        r["fd_relative_errors_vs_fp64"] = errs
    print(json.dumps(r))
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
