"""Shared helpers of the test suite."""
from __future__ import annotations

import numpy as np

Q = 19


def hip_arith(fast):
    """The lbm_config.optimize value of the oracle's arithmetic form `fast`
    (False: exact, True: fused, 2: contracted as ROCm's OpenCL compiler does)."""
    from lbmcl_amd import capi
    return {0: capi.LBM_ARITH_EXACT, 1: capi.LBM_ARITH_FUSED, 2: capi.LBM_ARITH_OPENCL}[int(fast)]


def window(a, n):
    """The VTI point window x, y, z in [1, N-2] of a flat N^3 field (lbmcl.hpp:284-286)."""
    return np.asarray(a).reshape(n, n, n)[1:n - 1, 1:n - 1, 1:n - 1].reshape(-1)


def window_v(u, n):
    u = np.asarray(u).reshape(3, n ** 3)
    return np.stack([window(u[k], n) for k in range(3)], axis=1)


def nan_equal(a, b) -> bool:
    return bool(np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True))


def mismatch_count(a, b) -> int:
    a, b = np.asarray(a), np.asarray(b)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    return int(np.count_nonzero(~same))


def max_rel(a, b, scale=None) -> float:
    """max |a - b| / scale over entries that are finite in both (scale = max |b|)."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    m = np.isfinite(a) & np.isfinite(b)
    if not m.any():
        return 0.0
    s = scale if scale is not None else max(float(np.max(np.abs(b[m]))), 1e-300)
    return float(np.max(np.abs(a[m] - b[m]))) / s
