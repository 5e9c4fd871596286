"""lbm_compute_noalloc against lbm_compute on the GPU, through the C ABI.

The entry point that allocates nothing runs one iteration with diagnostics on
buffers the caller owns (here torch tensors on cuda:0), its per-block partial
sums in the library's fixed device array.  Its outputs -- the written lattice,
rho, u, and the FP64 mass and energy -- must equal, bit for bit, those of
lbm_compute(ctx, it, 1, 1) on the same state.  Two calls in a row, the second
on a smaller lattice (and another storage form), so the second reuses the
array the first filled; then the size limit and argument errors.
"""
import ctypes

import numpy as np
import pytest

from lbmcl_amd import capi

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    """Device copy of a numpy array's bytes."""
    return torch.from_numpy(np.frombuffer(np.ascontiguousarray(a).tobytes(), dtype=np.uint8).copy()).to("cuda:0")


def _host(t, dtype):
    return np.frombuffer(t.cpu().numpy().tobytes(), dtype=dtype)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _check_against_compute(torch, n, stride, precision, storage, before):
    label = f"{n}^3 {precision} {storage} stride {stride}"
    print(f"lbm_compute_noalloc vs lbm_compute: {label}", flush=True)
    dt = np.float64 if precision == "f64" else np.float32
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
    t_src, t_dst, t_rho, t_u = _dev(torch, src), _dev(torch, dst0), _dev(torch, rho0), _dev(torch, u0)
    t_diag = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda:0")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = capi.lib().lbm_compute_noalloc(ctypes.byref(cfg), _ptr(t_src), _ptr(t_dst), _ptr(t_rho), _ptr(t_u), 1,
                                        _ptr(t_diag), stream)
    assert rc == capi.LBM_OK, capi._last_error()
    torch.cuda.synchronize()
    assert np.array_equal(_host(t_dst, np.uint8), np.frombuffer(want_dst.tobytes(), dtype=np.uint8)), \
        f"{label}: written lattice differs"
    assert np.array_equal(_host(t_rho, dt), want_rho, equal_nan=True), f"{label}: rho differs"
    assert np.array_equal(_host(t_u, dt), want_u, equal_nan=True), f"{label}: u differs"
    got = t_diag.cpu().numpy()
    assert got[0] == want_diag[0] and got[1] == want_diag[1], (label, got, want_diag)
    print(f"{label}: identical (mass {got[0]:.9f}, energy {got[1]:.6e})", flush=True)


def test_noalloc_matches_compute_two_calls():
    torch = pytest.importorskip("torch")
    _check_against_compute(torch, 128, 4096, "f32", "native", 7)
    _check_against_compute(torch, 64, 64, "f32", "fp16", 4)  # smaller, after the first
    _check_against_compute(torch, 32, 1, "f64", "native", 3)  # the tile kernel, smaller again


def test_noalloc_rejects_what_does_not_fit():
    torch = pytest.importorskip("torch")
    L = capi.lib()
    buf = torch.zeros(16, dtype=torch.float64, device="cuda:0")
    p = _ptr(buf)
    cfg = capi.LbmConfig()
    cfg.struct_size = ctypes.sizeof(capi.LbmConfig)
    cfg.dim, cfg.stride, cfg.viscosity, cfg.velocity = 1024, 4096, 0.0089, 0.05
    # a 1024^3 launch has more blocks than the fixed scratch holds: rejected before any launch
    assert L.lbm_compute_noalloc(ctypes.byref(cfg), p, ctypes.c_void_p(buf.data_ptr() + 64), p, p, 1, p,
                                 None) == capi.LBM_ERR_INVALID
    assert "exceeds the fixed scratch" in capi._last_error()
    cfg.dim, cfg.stride = 64, 64
    assert L.lbm_compute_noalloc(ctypes.byref(cfg), p, p, p, p, 0, p, None) == capi.LBM_ERR_INVALID
    assert "different lattices" in capi._last_error()
    assert L.lbm_compute_noalloc(ctypes.byref(cfg), p, None, p, p, 0, p, None) == capi.LBM_ERR_INVALID
    cfg.struct_size = 0
    assert L.lbm_compute_noalloc(ctypes.byref(cfg), p, ctypes.c_void_p(buf.data_ptr() + 64), p, p, 0, p,
                                 None) == capi.LBM_ERR_INVALID
    assert "struct_size" in capi._last_error()
