"""Verification of LBMCL outputs against target VTI files (verify.py counterpart).

Same command line and printout as the reference's verify.py (verify.py:12-60):
for every dumped iteration i = 0, every, 2*every, ... <= iterations it reads
``<target>/ldc.0.<i>.vti`` and ``<prediction>/lbmcl.<i>.vti`` (i zero-padded to
the digit count of `iterations`), replaces NaN by 0 and prints the mean squared
error and the maximum absolute error of rho and v.  The reference has no
pass/fail threshold; the optional ``--max-*`` options add one (exit status 1 if
any iteration exceeds it).  Reads files with the numpy-only reader in
lbmcl_amd.vti instead of pyvista; ``--pyvista`` additionally reads every file
with pyvista, when it is importable, and requires both readers to return the
same arrays.
"""
from __future__ import annotations

import argparse
import math
import sys

import numpy as np

from .vti import read_vti


def width_of(iterations: int) -> int:
    return int(math.log10(iterations)) + 1 if iterations > 0 else 1  # verify.py:27


def metrics(target_rho, target_v, pred_rho, pred_v):
    """(MSE rho, MSE v, MAE rho, MAE v) with NaN -> 0 (verify.py:47-56).

    MSE of v is sklearn's mean_squared_error with uniform averaging over the 3
    components, which equals the mean over all entries."""
    tr, pr = np.nan_to_num(target_rho.astype(np.float64)), np.nan_to_num(pred_rho.astype(np.float64))
    tv, pv = np.nan_to_num(target_v.astype(np.float64)), np.nan_to_num(pred_v.astype(np.float64))
    return (float(np.mean((tr - pr) ** 2)), float(np.mean((tv - pv) ** 2)),
            float(np.max(np.abs(tr - pr))), float(np.max(np.abs(tv - pv))))


def pyvista_reader():
    """pyvista's VTI reader (verify.py:44-45) as a function path -> {name: array}, or None."""
    try:
        import pyvista
    except ImportError:
        return None

    def read(path):
        mesh = pyvista.read(path)
        return {name: np.asarray(mesh.point_data[name]) for name in ("rho", "v")}

    return read


def cross_check(path, ours, theirs):
    """Both readers must agree exactly (NaN where NaN)."""
    for name in ("rho", "v"):
        a = np.asarray(ours[name], dtype=np.float64).reshape(-1)
        b = np.asarray(theirs[name], dtype=np.float64).reshape(-1)
        if a.shape != b.shape or not np.array_equal(a, b, equal_nan=True):
            raise ValueError(f"{path}: {name} differs between lbmcl_amd.vti and pyvista")


def verify(iterations, every, target_path, prediction_path, out=sys.stdout, other_reader=None):
    """Return a list of (iteration, mse_rho, mse_v, mae_rho, mae_v)."""
    w = width_of(iterations)
    print("{0:^{w}}  {1:^13}  {2:^13}  {3:^13}  {4:^13}".format("#it", "MSE_RHO", "MSE_U", "MAE_RHO", "MAE_U", w=w),
          file=out)
    rows = []
    for i in range(0, iterations + 1, every):
        tpath = "{p}/ldc.0.{i:0{w}}.vti".format(p=target_path, i=i, w=w)
        ppath = "{p}/lbmcl.{i:0{w}}.vti".format(p=prediction_path, i=i, w=w)
        t, p = read_vti(tpath), read_vti(ppath)
        if other_reader is not None:
            cross_check(tpath, t, other_reader(tpath))
            cross_check(ppath, p, other_reader(ppath))
        m = metrics(t["rho"], t["v"], p["rho"], p["v"])
        rows.append((i,) + m)
        print("{i:{w}}:  {0:e}   {1:e}   {2:e}   {3:e}".format(*m, i=i, w=w), file=out)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description="MSE and max absolute error of rho and v between target and "
                                             "prediction VTI datasets.")
    ap.add_argument("-i", "--iterations", type=int, required=True)
    ap.add_argument("-e", "--every", type=int, required=True)
    ap.add_argument("-t", "--target_path", type=str, required=True)
    ap.add_argument("-p", "--prediction_path", type=str, required=True)
    ap.add_argument("--max-mse-rho", type=float, default=None)
    ap.add_argument("--max-mse-u", type=float, default=None)
    ap.add_argument("--max-mae-rho", type=float, default=None)
    ap.add_argument("--max-mae-u", type=float, default=None)
    ap.add_argument("--pyvista", action="store_true", help="also read every file with pyvista and compare")
    a = ap.parse_args(argv)
    other = None
    if a.pyvista:
        other = pyvista_reader()
        if other is None:
            print("pyvista is not installed: cross-check skipped", file=sys.stderr)
    rows = verify(a.iterations, a.every, a.target_path, a.prediction_path, other_reader=other)
    limits = (a.max_mse_rho, a.max_mse_u, a.max_mae_rho, a.max_mae_u)
    bad = [r for r in rows if any(lim is not None and v > lim for v, lim in zip(r[1:], limits))]
    if bad:
        print(f"FAIL: {len(bad)} iteration(s) above the thresholds", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
