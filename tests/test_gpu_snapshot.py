"""Asynchronous snapshots of the output path (lbm_snapshot, include/lbm_hip.h).

Each snapshot kind equals the blocking read it replaces, rearranged to the
snapshot layout: the VTI window of rho and u (lbm_read_rho / lbm_read_u), the
cell-major push view of a lattice (lbm_read_f), the map (lbm_read_map).  A
snapshot captures the state at the point of the stream where it was enqueued,
even when more iterations are enqueued before it is waited for.  With every
slot held, lbm_snapshot blocks until another thread releases one.
"""
import threading
import time

import numpy as np
import pytest

from lbmcl_amd import capi
from tests.util import nan_equal

pytestmark = pytest.mark.gpu


def _window(rho, u, n):
    w = n - 2
    r = rho.reshape(n, n, n)[1:n - 1, 1:n - 1, 1:n - 1].reshape(-1)
    v = np.stack([u.reshape(3, n, n, n)[k][1:n - 1, 1:n - 1, 1:n - 1].reshape(-1) for k in range(3)], axis=1)
    assert r.size == w ** 3
    return np.concatenate([r, v.reshape(-1)])


def _cell_major(f, n, stride):
    c = np.arange(n ** 3)
    q = np.arange(19)
    slots = ((c[:, None] // stride) * 19 + q[None, :]) * stride + (c[:, None] & (stride - 1))
    return f[slots.reshape(-1)]


@pytest.mark.parametrize("n,stride,precision,storage", [(32, 8, "f32", "native"), (64, 64, "f64", "native"),
                                                        (128, 128, "f32", "fp16")])
def test_snapshots_equal_blocking_reads(n, stride, precision, storage):
    with capi.Context(n, stride=stride, precision=precision, storage=storage) as g:
        g.initialize()
        tm = g.snapshot(capi.LBM_SNAPSHOT_MAP)
        g.compute_range(1, 9, 3)  # macro fields stored at iterations 3, 6, 9
        t_macro = g.snapshot(capi.LBM_SNAPSHOT_MACRO)
        t_f = g.snapshot(capi.LBM_SNAPSHOT_F, 1)
        rho, u, f1, m = g.read_rho(), g.read_u(), g.read_f(1), g.read_map()
        g.compute_range(10, 5, 1)  # overwrites rho, u and both lattices
        assert np.array_equal(g.snapshot_wait(tm, capi.LBM_SNAPSHOT_MAP), m)
        assert nan_equal(g.snapshot_wait(t_macro, capi.LBM_SNAPSHOT_MACRO), _window(rho, u, n))
        assert nan_equal(g.snapshot_wait(t_f, capi.LBM_SNAPSHOT_F), _cell_major(f1, n, stride))
        for t in (tm, t_macro, t_f):
            g.snapshot_release(t)
        # slots are reused
        for _ in range(2 * capi.LBM_SNAPSHOT_SLOTS):
            t = g.snapshot(capi.LBM_SNAPSHOT_MACRO)
            out = g.snapshot_wait(t, capi.LBM_SNAPSHOT_MACRO)
            g.snapshot_release(t)
        assert nan_equal(out, _window(g.read_rho(), g.read_u(), n))


def test_snapshot_errors_and_backpressure():
    n = 32
    with capi.Context(n, stride=n) as g:
        g.initialize()
        with pytest.raises(capi.LbmError):
            g.snapshot(7)
        with pytest.raises(capi.LbmError):
            g.snapshot(capi.LBM_SNAPSHOT_F, 2)
        with pytest.raises(capi.LbmError):
            g.snapshot_wait(12345, capi.LBM_SNAPSHOT_MACRO)
        held = [g.snapshot(capi.LBM_SNAPSHOT_MACRO) for _ in range(capi.LBM_SNAPSHOT_SLOTS)]
        with pytest.raises(capi.LbmError):  # must be waited for first
            g.snapshot_release(held[0])
        for t in held:
            g.snapshot_wait(t, capi.LBM_SNAPSHOT_MACRO)

        def release_later():
            time.sleep(0.5)
            g.snapshot_release(held[1])

        th = threading.Thread(target=release_later)
        th.start()
        t0 = time.perf_counter()
        t = g.snapshot(capi.LBM_SNAPSHOT_MACRO)  # blocks until the thread releases a slot
        waited = time.perf_counter() - t0
        th.join()
        assert waited > 0.3
        g.snapshot_wait(t, capi.LBM_SNAPSHOT_MACRO)
        for x in (held[0], held[2], t):
            g.snapshot_release(x)
        with pytest.raises(capi.LbmError):
            g.snapshot_release(t)
