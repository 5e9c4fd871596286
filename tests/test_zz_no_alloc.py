"""lbm_compute_noalloc against lbm_compute on the GPU, through the C ABI.

The entry point that allocates nothing runs one iteration with diagnostics on
buffers the caller owns (here allocated with the HIP runtime the library
itself uses, through ctypes), its per-block partial sums in the library's
fixed device array.  Its outputs -- the written lattice, rho, u, and the FP64
mass and energy -- must equal, bit for bit, those of lbm_compute(ctx, it, 1, 1)
on the same state.  Two calls in a row, the second on a smaller lattice (and
another storage form), so the second reuses the array the first filled; then
the size limit and argument errors.
"""
import ctypes

import numpy as np
import pytest

from lbmcl_amd import capi

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    capi.lib()  # the library first: its libamdhip64 is the one loaded below
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


class _Dev:
    """A device buffer holding a copy of a numpy array's bytes."""

    def __init__(self, h, a):
        self.h, self.a = h, np.ascontiguousarray(a)
        self.p = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), max(self.a.nbytes, 1)) == 0
        assert h.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, H2D) == 0

    def read(self):
        out = np.empty_like(self.a)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        assert self.h.hipFree(self.p) == 0


def _check_against_compute(h, n, stride, precision, storage, before):
    label = f"{n}^3 {precision} {storage} stride {stride}"
    print(f"lbm_compute_noalloc vs lbm_compute: {label}", flush=True)
    with capi.Context(n, stride=stride, precision=precision, storage=storage, profile=False) as g:
        g.initialize()
        g.compute_range(1, before, 1)  # rho and u stored every iteration
        it = before + 1
        src_buf = 0 if it % 2 else 1  # odd iterations read lattice 0
        src, dst0 = g.read_state(src_buf), g.read_state(1 - src_buf)
        rho0, u0 = g.read_rho(), g.read_u()
        g.compute(it, update_macro=True, diagnostics=True)
        want_dst, want_rho, want_u = g.read_state(1 - src_buf), g.read_rho(), g.read_u()
        want_diag = g.read_diagnostics()
        cfg = g.cfg
    bufs = [_Dev(h, x) for x in (src, dst0, rho0, u0, np.full(2, np.nan))]
    d_src, d_dst, d_rho, d_u, d_diag = bufs
    try:
        # the null stream: the call is ordered after the copies above
        rc = capi.lib().lbm_compute_noalloc(ctypes.byref(cfg), d_src.p, d_dst.p, d_rho.p, d_u.p, 1, d_diag.p, None)
        assert rc == capi.LBM_OK, capi._last_error()
        assert h.hipDeviceSynchronize() == 0
        assert np.array_equal(d_dst.read().view(np.uint8), want_dst.view(np.uint8)), f"{label}: written lattice differs"
        assert np.array_equal(d_rho.read(), want_rho, equal_nan=True), f"{label}: rho differs"
        assert np.array_equal(d_u.read(), want_u, equal_nan=True), f"{label}: u differs"
        got = d_diag.read()
        assert got[0] == want_diag[0] and got[1] == want_diag[1], (label, got, want_diag)
        print(f"{label}: identical (mass {got[0]:.9f}, energy {got[1]:.6e})", flush=True)
    finally:
        for b in bufs:
            b.free()


def test_noalloc_matches_compute_two_calls():
    h = _hip()
    _check_against_compute(h, 128, 4096, "f32", "native", 7)
    _check_against_compute(h, 64, 64, "f32", "fp16", 4)  # smaller, after the first
    _check_against_compute(h, 32, 1, "f64", "native", 3)  # the tile kernel, smaller again


def test_noalloc_rejects_what_does_not_fit():
    h = _hip()
    L = capi.lib()
    buf = _Dev(h, np.zeros(16, dtype=np.float64))
    p = buf.p
    cfg = capi.LbmConfig()
    cfg.struct_size = ctypes.sizeof(capi.LbmConfig)
    cfg.dim, cfg.stride, cfg.viscosity, cfg.velocity = 1024, 4096, 0.0089, 0.05
    # a 1024^3 launch has more blocks than the fixed scratch holds: rejected before any launch
    assert L.lbm_compute_noalloc(ctypes.byref(cfg), p, ctypes.c_void_p(buf.p.value + 64), p, p, 1, p,
                                 None) == capi.LBM_ERR_INVALID
    assert "exceeds the fixed scratch" in capi._last_error()
    cfg.dim, cfg.stride = 64, 64
    assert L.lbm_compute_noalloc(ctypes.byref(cfg), p, p, p, p, 0, p, None) == capi.LBM_ERR_INVALID
    assert "different lattices" in capi._last_error()
    assert L.lbm_compute_noalloc(ctypes.byref(cfg), p, None, p, p, 0, p, None) == capi.LBM_ERR_INVALID
    cfg.struct_size = 0
    assert L.lbm_compute_noalloc(ctypes.byref(cfg), p, ctypes.c_void_p(buf.p.value + 64), p, p, 0, p,
                                 None) == capi.LBM_ERR_INVALID
    assert "struct_size" in capi._last_error()
    buf.free()
