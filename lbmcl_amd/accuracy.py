"""Accuracy of the storage/precision forms against an FP64 run of the same cavity.

For a cavity of size N run for T iterations, every form (FP32, FP32 with the
fused collision, FP16- and BF16-shifted storage) is compared with the FP64
native run on the fluid and lid cells:

  rho      max |rho - rho64| / max |rho64|       (relative, rho ~ 1)
  u        max |u - u64| / max |u64|              (relative to the largest speed)
  u_rms    rms |u - u64| / max |u64|
  *_at     (x, y, z) of the largest difference
  *_core   the same maxima over cells at least 4 cells from every wall
  mass     relative difference of the total mass sum(rho)
  vortex   distance (lattice units) between the primary-vortex centres, in the
           mid-plane y = N/2: the centroid of the positive y-vorticity
           w_y = d ux/dz - d uz/dx (the circulation the lid drives), and the
           point of least in-plane speed below the lid (parabolic refinement),
           which is only meaningful once the vortex has formed.

Usage: python -m lbmcl_amd.accuracy --dim 256 --iterations 2000 [--stride S]
prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import capi

# (precision, storage, arithmetic); 16-bit storage always computes in the deviation form
FORMS = (("f32", "native", "reference"), ("f32", "native", "fused"), ("f32", "fp16", "reference"),
         ("f32", "bf16", "reference"))


def run_form(dim, iterations, precision, storage, arith="reference", stride=None, viscosity=0.0089, velocity=0.05):
    """(rho, u) after `iterations` collide-stream steps (the macros of iteration iterations + 1)."""
    stride = stride or dim ** 3
    with capi.Context(dim, viscosity=viscosity, velocity=velocity, stride=stride, precision=precision,
                      storage=storage, arith=arith, profile=False) as g:
        g.initialize()
        g.compute_range(1, iterations, 0)
        g.compute(iterations + 1, update_macro=True)
        rho = g.read_rho().astype(np.float64)
        u = g.read_u().astype(np.float64).reshape(3, dim, dim, dim)
    return rho.reshape(dim, dim, dim), u


def vorticity_centroid(u, n):
    """Centroid (x, z) of the positive y-vorticity in the plane y = n // 2."""
    y = n // 2
    ux, uz = u[0][:, y, :], u[2][:, y, :]          # indexed [z, x]
    wy = np.zeros_like(ux)
    wy[1:-1, 1:-1] = 0.5 * (ux[2:, 1:-1] - ux[:-2, 1:-1]) - 0.5 * (uz[1:-1, 2:] - uz[1:-1, :-2])
    w = np.nan_to_num(wy, nan=0.0)
    w[:3, :] = 0.0
    w[:, :3] = 0.0
    w[-3:, :] = 0.0
    w[:, -3:] = 0.0
    wp = np.clip(w, 0.0, None)
    s = wp.sum()
    if s == 0:
        return (float("nan"), float("nan"))
    zz, xx = np.mgrid[0:n, 0:n]
    return (float((wp * xx).sum() / s), float((wp * zz).sum() / s))


def min_speed_centre(u, n):
    """In-plane least-speed point (x, z) of the mid-plane below the lid, refined
    by a parabola through the neighbours in x and in z."""
    y = n // 2
    sp = np.sqrt(np.nan_to_num(u[0][:, y, :]) ** 2 + np.nan_to_num(u[2][:, y, :]) ** 2)
    lo, hi = max(2, n // 8), n - max(3, n // 8)
    sub = sp[n // 2:hi, lo:hi]  # upper half: the primary vortex sits below the lid
    k = int(np.argmin(sub))
    zi, xi = np.unravel_index(k, sub.shape)
    z, x = zi + n // 2, xi + lo

    def refine(a, b, c):
        d = a - 2 * b + c
        return 0.0 if d == 0 else 0.5 * (a - c) / d

    dx = refine(sp[z, x - 1], sp[z, x], sp[z, x + 1])
    dz = refine(sp[z - 1, x], sp[z, x], sp[z + 1, x])
    return (float(x + dx), float(z + dz))


def compare(ref, test, n):
    rho0, u0 = ref
    rho1, u1 = test
    m = np.isfinite(rho0)
    umax = float(np.max(np.sqrt(np.sum(np.nan_to_num(u0) ** 2, axis=0))))
    du = np.sqrt(np.sum((np.nan_to_num(u1) - np.nan_to_num(u0)) ** 2, axis=0))[m]
    c0, c1 = vorticity_centroid(u0, n), vorticity_centroid(u1, n)
    s0, s1 = min_speed_centre(u0, n), min_speed_centre(u1, n)
    # locations (arrays are indexed [z, y, x]) and the core away from the walls
    dr_full = np.where(m, np.abs(rho1 - rho0), 0.0)
    du_full = np.where(m, np.sqrt(np.sum((np.nan_to_num(u1) - np.nan_to_num(u0)) ** 2, axis=0)), 0.0)
    core = np.zeros_like(m)
    core[4:n - 4, 4:n - 4, 4:n - 4] = True

    def at(a):
        z, y, x = np.unravel_index(int(np.argmax(a)), a.shape)
        return [int(x), int(y), int(z)]

    rmax = np.max(np.abs(rho0[m]))
    return {
        "rho_rel_max": float(np.max(np.abs(rho1[m] - rho0[m])) / rmax),
        "rho_rel_max_at": at(dr_full),
        "rho_rel_max_core": float(np.max(dr_full[core]) / rmax),
        "mass_rel": float((np.sum(rho1[m]) - np.sum(rho0[m])) / np.sum(rho0[m])),
        "u_rel_max": float(np.max(du) / umax),
        "u_rel_max_at": at(du_full),
        "u_rel_max_core": float(np.max(du_full[core]) / umax),
        "u_rel_rms": float(np.sqrt(np.mean(du ** 2)) / umax),
        "vortex_centroid": c1,
        "vortex_centroid_shift": float(np.hypot(c1[0] - c0[0], c1[1] - c0[1])),
        "min_speed_point": s1,
        "min_speed_point_shift": float(np.hypot(s1[0] - s0[0], s1[1] - s0[1])),
    }


def accuracy(dim, iterations, stride=None):
    ref = run_form(dim, iterations, "f64", "native", "reference", stride)
    out = {"dim": dim, "iterations": iterations, "reference": "f64 native",
           "u_max_reference": float(np.nanmax(np.sqrt(np.sum(np.nan_to_num(ref[1]) ** 2, axis=0)))),
           "vortex_centroid_reference": vorticity_centroid(ref[1], dim),
           "min_speed_point_reference": min_speed_centre(ref[1], dim), "forms": {}}
    for precision, storage, arith in FORMS:
        name = f"{precision}_{storage}" + ("_fused" if arith == "fused" and storage == "native" else "")
        out["forms"][name] = compare(ref, run_form(dim, iterations, precision, storage, arith, stride), dim)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--dim", type=int, default=256)
    ap.add_argument("--iterations", type=int, default=2000)
    ap.add_argument("--stride", type=int, default=0)
    a = ap.parse_args(argv)
    print(json.dumps(accuracy(a.dim, a.iterations, a.stride or None)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
