"""ctypes binding of liblbm_hip.so (include/lbm_hip.h).

This is the Python caller of the C ABI used by the validation layer, the
tests, the benchmark and the autograd wrapper.  It loads the in-tree library
and raises if it is missing: there is no CPU fallback behind it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "liblbm_hip.so")

LBM_OK = 0
LBM_ERR_INVALID = -1
LBM_ERR_DEVICE = -2
LBM_ERR_MEMORY = -3
LBM_ERR_RUNTIME = -4
LBM_ERR_STATE = -5

LBM_F32, LBM_F64 = 0, 1
LBM_STORE_NATIVE, LBM_STORE_FP16_SHIFT, LBM_STORE_BF16_SHIFT = 0, 1, 2
LBM_KERNEL_AUTO, LBM_KERNEL_GENERIC, LBM_KERNEL_ROW, LBM_KERNEL_VEC, LBM_KERNEL_VEC8, LBM_KERNEL_TILE = 0, 1, 2, 3, 4, 5
LBM_KERNEL_LDS = 6
# lbm_config.optimize: arithmetic form of the collision.  The zero value is the
# reference's own arithmetic (kernels.cl as ROCm's OpenCL compiler evaluates it;
# 16-bit storage always computes in the deviation form); EXACT is the CPU oracle's.
LBM_ARITH_REFERENCE = LBM_ARITH_OPENCL = 0
LBM_ARITH_FUSED, LBM_ARITH_EXACT = 1, 2
ARITH_NAMES = {"reference": LBM_ARITH_REFERENCE, "opencl": LBM_ARITH_OPENCL, "fused": LBM_ARITH_FUSED,
               "exact": LBM_ARITH_EXACT}

EVENT_NAMES = {0: "initialize", 1: "compute", 2: "read_map", 3: "read_f", 4: "read_rho", 5: "read_u",
               6: "diagnostics"}

STORAGE_NAMES = {"native": LBM_STORE_NATIVE, "fp16": LBM_STORE_FP16_SHIFT, "bf16": LBM_STORE_BF16_SHIFT}

EXPORTS = (
    "lbm_get_device_count", "lbm_get_device_name", "lbm_create", "lbm_destroy", "lbm_initialize",
    "lbm_compute", "lbm_compute_range", "lbm_finish", "lbm_read_map", "lbm_read_rho", "lbm_read_u",
    "lbm_read_f", "lbm_read_state", "lbm_write_state", "lbm_read_diagnostics", "lbm_get_constants",
    "lbm_event_count", "lbm_event_get", "lbm_total_time_ms", "lbm_kernels_time_ms", "lbm_clear_events",
    "lbm_set_force", "lbm_adjoint", "lbm_snapshot", "lbm_snapshot_wait", "lbm_snapshot_release", "lbm_last_error",
    "lbm_compute_noalloc",
)

LBM_SNAPSHOT_MACRO, LBM_SNAPSHOT_F, LBM_SNAPSHOT_MAP = 0, 1, 2
LBM_SNAPSHOT_SLOTS = 3

LBM_OBJECTIVE_KINETIC_ENERGY, LBM_OBJECTIVE_MEAN_UX_PLANE = 0, 1


class LbmConfig(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("dim", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("stride", ctypes.c_int64),
        ("viscosity", ctypes.c_double),
        ("velocity", ctypes.c_double),
        ("lws", ctypes.c_int32 * 3),
        ("optimize", ctypes.c_int32),
        ("kernel", ctypes.c_int32),
        ("profile", ctypes.c_int32),
        ("xcd_remap", ctypes.c_int32),
        ("raw_parameters", ctypes.c_int32),
    ]


class LbmAdjointConfig(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("objective", ctypes.c_int32),
        ("plane_z", ctypes.c_int32),
        ("steps", ctypes.c_int64),
        ("checkpoint_every", ctypes.c_int64),
    ]


class LbmAdjointResult(ctypes.Structure):
    _fields_ = [
        ("objective", ctypes.c_double),
        ("d_velocity", ctypes.c_double),
        ("d_viscosity", ctypes.c_double),
        ("forward_ms", ctypes.c_double),
        ("adjoint_ms", ctypes.c_double),
        ("device_bytes", ctypes.c_double),
        ("checkpoints", ctypes.c_int64),
        ("segment", ctypes.c_int64),
    ]


class LbmError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} failed ({code}): {_last_error()}")
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    """Load liblbm_hip.so; raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m lbmcl_amd.build` "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    P = ctypes.POINTER
    sig = {
        "lbm_get_device_count": [P(i32)],
        "lbm_get_device_name": [i32, ctypes.c_char_p, ctypes.c_size_t],
        "lbm_create": [P(LbmConfig), P(vp)],
        "lbm_destroy": [vp],
        "lbm_initialize": [vp],
        "lbm_compute": [vp, i64, i32, i32],
        "lbm_compute_range": [vp, i64, i64, i64],
        "lbm_finish": [vp],
        "lbm_read_map": [vp, vp],
        "lbm_read_rho": [vp, vp],
        "lbm_read_u": [vp, vp],
        "lbm_read_f": [vp, i32, vp],
        "lbm_read_state": [vp, i32, vp],
        "lbm_write_state": [vp, i32, vp],
        "lbm_read_diagnostics": [vp, P(dbl), P(dbl)],
        "lbm_get_constants": [vp, P(dbl)],
        "lbm_event_count": [vp, P(i64)],
        "lbm_event_get": [vp, i64, P(i32), P(dbl)],
        "lbm_total_time_ms": [vp, P(dbl)],
        "lbm_kernels_time_ms": [vp, P(dbl)],
        "lbm_clear_events": [vp],
        "lbm_set_force": [vp, vp],
        "lbm_adjoint": [vp, P(LbmAdjointConfig), P(LbmAdjointResult), vp],
        "lbm_snapshot": [vp, i32, i32, P(i64)],
        "lbm_snapshot_wait": [vp, i64, P(vp), P(i64)],
        "lbm_snapshot_release": [vp, i64],
        "lbm_compute_noalloc": [P(LbmConfig), vp, vp, vp, vp, i32, vp, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.lbm_last_error.argtypes = []
    L.lbm_last_error.restype = ctypes.c_char_p


def _last_error() -> str:
    return lib().lbm_last_error().decode(errors="replace")


def check(code, what):
    if code != LBM_OK:
        raise LbmError(code, what)


def device_count() -> int:
    n = ctypes.c_int32(0)
    check(lib().lbm_get_device_count(ctypes.byref(n)), "lbm_get_device_count")
    return n.value


def device_name(device=0) -> str:
    buf = ctypes.create_string_buffer(256)
    check(lib().lbm_get_device_name(device, buf, 256), "lbm_get_device_name")
    return buf.value.decode()


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One device lattice: a thin owner of an lbm_ctx."""

    def __init__(self, dim, viscosity=0.0089, velocity=0.05, stride=None, precision="f32", storage="native",
                 device=0, lws=(0, 0, 0), kernel=LBM_KERNEL_AUTO, profile=True, xcd_remap=True,
                 arith=LBM_ARITH_REFERENCE, raw_parameters=False):
        """`arith`: an LBM_ARITH_* value or one of "reference", "opencl", "fused", "exact"."""
        self.dim = int(dim)
        self.precision = precision
        self.storage = storage
        self.dtype = np.float64 if precision == "f64" else np.float32
        if storage == "native":
            self.sdtype = self.dtype
        else:
            self.sdtype = np.uint16
        cfg = LbmConfig()
        cfg.struct_size = ctypes.sizeof(LbmConfig)
        cfg.dim = self.dim
        cfg.precision = LBM_F64 if precision == "f64" else LBM_F32
        cfg.storage = STORAGE_NAMES[storage]
        cfg.device = device
        cfg.stride = int(stride if stride is not None else self.dim)
        cfg.viscosity = viscosity
        cfg.velocity = velocity
        cfg.lws[0], cfg.lws[1], cfg.lws[2] = lws
        cfg.optimize = ARITH_NAMES[arith] if isinstance(arith, str) else int(arith)
        cfg.kernel = kernel
        cfg.profile = int(bool(profile))
        cfg.xcd_remap = 1 if xcd_remap else -1  # 0 (a zeroed config) also means on
        cfg.raw_parameters = int(bool(raw_parameters))
        self.cfg = cfg
        self.stride = cfg.stride
        h = ctypes.c_void_p()
        check(lib().lbm_create(ctypes.byref(cfg), ctypes.byref(h)), "lbm_create")
        self._h = h

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            check(lib().lbm_destroy(self._h), "lbm_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- enqueue ------------------------------------------------------------
    def initialize(self):
        check(lib().lbm_initialize(self._h), "lbm_initialize")

    def compute(self, iteration, update_macro=False, diagnostics=False):
        check(lib().lbm_compute(self._h, int(iteration), int(bool(update_macro)), int(bool(diagnostics))),
              "lbm_compute")

    def compute_range(self, first, count, every=0):
        check(lib().lbm_compute_range(self._h, int(first), int(count), int(every)), "lbm_compute_range")

    def finish(self):
        check(lib().lbm_finish(self._h), "lbm_finish")

    # -- reads ----------------------------------------------------------------
    @property
    def n3(self):
        return self.dim ** 3

    def read_map(self):
        out = np.empty(self.n3, dtype=np.int32)
        check(lib().lbm_read_map(self._h, _ptr(out)), "lbm_read_map")
        return out

    def read_rho(self):
        out = np.empty(self.n3, dtype=self.dtype)
        check(lib().lbm_read_rho(self._h, _ptr(out)), "lbm_read_rho")
        return out

    def read_u(self):
        out = np.empty(3 * self.n3, dtype=self.dtype)
        check(lib().lbm_read_u(self._h, _ptr(out)), "lbm_read_u")
        return out

    def read_f(self, buffer):
        out = np.empty(19 * self.n3, dtype=self.dtype)
        check(lib().lbm_read_f(self._h, int(buffer), _ptr(out)), "lbm_read_f")
        return out

    def snapshot(self, what, buffer=0):
        """Enqueue an asynchronous snapshot (LBM_SNAPSHOT_*); returns its ticket."""
        t = ctypes.c_int64(0)
        check(lib().lbm_snapshot(self._h, int(what), int(buffer), ctypes.byref(t)), "lbm_snapshot")
        return t.value

    def snapshot_wait(self, ticket, what):
        """Wait for a snapshot and return a copy of its data (the slot stays held)."""
        p, nbytes = ctypes.c_void_p(), ctypes.c_int64(0)
        check(lib().lbm_snapshot_wait(self._h, int(ticket), ctypes.byref(p), ctypes.byref(nbytes)),
              "lbm_snapshot_wait")
        dt = np.int32 if what == LBM_SNAPSHOT_MAP else self.dtype
        buf = (ctypes.c_char * nbytes.value).from_address(p.value)
        return np.frombuffer(buf, dtype=dt).copy()

    def snapshot_release(self, ticket):
        check(lib().lbm_snapshot_release(self._h, int(ticket)), "lbm_snapshot_release")

    def read_state(self, buffer):
        out = np.empty(19 * self.n3, dtype=self.sdtype)
        check(lib().lbm_read_state(self._h, int(buffer), _ptr(out)), "lbm_read_state")
        return out

    def write_state(self, buffer, data):
        data = np.ascontiguousarray(data, dtype=self.sdtype)
        assert data.size == 19 * self.n3
        check(lib().lbm_write_state(self._h, int(buffer), _ptr(data)), "lbm_write_state")

    def read_diagnostics(self):
        m, e = ctypes.c_double(), ctypes.c_double()
        check(lib().lbm_read_diagnostics(self._h, ctypes.byref(m), ctypes.byref(e)), "lbm_read_diagnostics")
        return m.value, e.value

    def constants(self):
        out = (ctypes.c_double * 4)()
        check(lib().lbm_get_constants(self._h, out), "lbm_get_constants")
        return {"inv_tau": out[0], "velocity": out[1], "viscosity": out[2], "scale": out[3]}

    # -- body force and adjoint -----------------------------------------------
    def set_force(self, force):
        """Body force g as an array of 3*dim^3 values [gx | gy | gz], or None."""
        if force is None:
            check(lib().lbm_set_force(self._h, None), "lbm_set_force")
            return
        g = np.ascontiguousarray(force, dtype=self.dtype).reshape(-1)
        assert g.size == 3 * self.n3
        check(lib().lbm_set_force(self._h, _ptr(g)), "lbm_set_force")

    def adjoint(self, steps, objective=LBM_OBJECTIVE_KINETIC_ENERGY, plane_z=0, checkpoint_every=0,
                grad_force=True):
        """J and dJ/dU, dJ/dnu (and dJ/dg, shape (3, dim^3), if grad_force) after `steps` iterations."""
        cfg = LbmAdjointConfig(ctypes.sizeof(LbmAdjointConfig), objective, plane_z, steps, checkpoint_every)
        res = LbmAdjointResult()
        g = np.zeros(3 * self.n3, dtype=np.float64) if grad_force else None
        check(lib().lbm_adjoint(self._h, ctypes.byref(cfg), ctypes.byref(res), None if g is None else _ptr(g)),
              "lbm_adjoint")
        out = {f: getattr(res, f) for f, _ in LbmAdjointResult._fields_}
        out["d_force"] = None if g is None else g.reshape(3, self.n3)
        return out

    # -- timing ---------------------------------------------------------------
    def events(self):
        n = ctypes.c_int64()
        check(lib().lbm_event_count(self._h, ctypes.byref(n)), "lbm_event_count")
        out = []
        for i in range(n.value):
            k, ms = ctypes.c_int32(), ctypes.c_double()
            check(lib().lbm_event_get(self._h, i, ctypes.byref(k), ctypes.byref(ms)), "lbm_event_get")
            out.append((EVENT_NAMES.get(k.value, str(k.value)), ms.value))
        return out

    def total_time_ms(self):
        ms = ctypes.c_double()
        check(lib().lbm_total_time_ms(self._h, ctypes.byref(ms)), "lbm_total_time_ms")
        return ms.value

    def kernels_time_ms(self):
        ms = ctypes.c_double()
        check(lib().lbm_kernels_time_ms(self._h, ctypes.byref(ms)), "lbm_kernels_time_ms")
        return ms.value

    def clear_events(self):
        check(lib().lbm_clear_events(self._h), "lbm_clear_events")
