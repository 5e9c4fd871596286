"""CPU oracle for the D3Q19 lid-driven-cavity collide-stream path.

TEST INFRASTRUCTURE ONLY.  This package is the checker.  Only ``tests/``, the
smoke check in ``__graft_entry__.py`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; the product (``lbmcl_amd``) never does.

The arithmetic lives in ``lbm_oracle.cpp`` (a scalar C++ restatement of the
reference's kernels.cl:236-426 and lbmcl.hpp:131-156, 435-520); this module
loads it through ctypes and drives it the way the reference host class drives
its kernels:

* ``PushSim`` -- the reference form: ``initialize`` into both lattices, then
  ``compute`` per iteration with ping-pong buffers (lbmcl.hpp:435-467,
  490-521), push streaming and a cell map.
* ``PullSim`` -- the gather form the HIP kernels use (post-collision values at
  their source cell, aligned writes), native or shifted FP16/BF16 storage.

Pinning (what checks this oracle, see tests/test_oracle.py):
  the Sailfish fixtures of the reference (target_results/8 and /32, via the
  verify.py metrics), hand-derived closed forms at step 2 of the 8^3 cavity,
  the folded constant INV_TAU (kernels.cl:62), stride invariance, y-mirror
  symmetry, the U=0 rest state, push == pull bit for bit, and the FP16/BF16
  codecs against numpy/torch conversions.  The body-force term and the
  diagnostics have no reference counterpart; they are pinned by their
  mathematics (zero force is bit-identical to no force; the force term is
  linear and conserves mass; diagnostics against numpy sums).

``fast = ARITH_CONTRACTED`` evaluates kernels.cl's expressions with the
contractions clang forms for OpenCL C (std::fma where it emits fmuladd), the
reference's arithmetic on AMD GPUs; in FP64 that is bit-identical to the HIP
LBM_ARITH_OPENCL form.

``ref_run`` runs the reference's own kernels.cl -- compiled offline from the
reference's source by ``ref_build.py`` into ``oracle/_ref`` -- on the local GPU
through the ROCm OpenCL runtime (``ref_runner.c``), for comparisons with "the
reference run at the same precision" and for its throughput on the same GPU.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_oracle.so")

Q = 19
EX = np.array([0, 1, 0, -1, 0, 0, 0, 1, -1, -1, 1, 1, 0, -1, 0, 1, 0, -1, 0])
EY = np.array([0, 0, 1, 0, -1, 0, 0, 1, 1, -1, -1, 0, 1, 0, -1, 0, 1, 0, -1])
EZ = np.array([0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, -1, -1, -1, 1, 1, 1, 1])
OPP = np.array([0, 3, 4, 1, 2, 6, 5, 9, 10, 7, 8, 17, 18, 15, 16, 13, 14, 11, 12])

# cell flags (common.h:7-18)
FLUID, MOVING, CORNER, WALL = 0x1, 0x2, 0x4, 0x8
LEFT, RIGHT, TOP, BOTTOM, FRONT, BACK = 0x10, 0x20, 0x40, 0x80, 0x100, 0x200

STORE_NATIVE, STORE_FP16, STORE_BF16 = 0, 1, 2


class Params(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("threads", ctypes.c_int32),
        ("stride", ctypes.c_int64),
        ("viscosity", ctypes.c_double),
        ("velocity", ctypes.c_double),
        ("fast", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


def build() -> None:
    """Compile liblbm_oracle.so with the recipe in oracle/Makefile and, where the
    reference checkout is present, oracle/_ref with oracle/ref_build.py."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    from . import ref_build

    ref_build.build(quiet=True)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _declare(L):
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    P = ctypes.POINTER(Params)
    L.lbmo_cell_type.argtypes = [i32, i32, i32, i32]
    L.lbmo_cell_type.restype = ctypes.c_int
    L.lbmo_slot.argtypes = [i64, i32, i64]
    L.lbmo_slot.restype = i64
    L.lbmo_shift_scale.argtypes = [dbl]
    L.lbmo_shift_scale.restype = dbl
    for s in ("f32", "f64"):
        getattr(L, f"lbmo_constants_{s}").argtypes = [P, vp]
        getattr(L, f"lbmo_initialize_{s}").argtypes = [P, vp, vp, vp, vp, vp]
        getattr(L, f"lbmo_compute_{s}").argtypes = [P, vp, vp, vp, vp, vp, i32, vp]
        getattr(L, f"lbmo_diagnostics_{s}").argtypes = [P, vp, vp, vp]
    L.lbmo_pull_initialize_f32.argtypes = [P, i32, vp, vp, vp, vp, vp]
    L.lbmo_pull_initialize_f64.argtypes = [P, vp, vp, vp, vp, vp]
    L.lbmo_pull_compute_f32.argtypes = [P, i32, vp, vp, vp, vp, i32, vp, vp]
    L.lbmo_pull_compute_f64.argtypes = [P, vp, vp, vp, vp, i32, vp, vp]
    L.lbmo_pull_view_f32.argtypes = [P, i32, vp, vp]
    L.lbmo_pull_view_f64.argtypes = [P, vp, vp]
    L.lbmo_collide_pair.argtypes = [P, i32, vp, vp, vp, vp, vp]
    for name in ("lbmo_f32_to_f16", "lbmo_f16_to_f32", "lbmo_f32_to_bf16", "lbmo_bf16_to_f32"):
        getattr(L, name).argtypes = [vp, vp, i64]


# ---------------------------------------------------------------------------
# JIT literal rounding (lbmcl.hpp:140-141): the value is printed with the
# stream's default 6 significant digits and parsed back by the OpenCL compiler
# as a float literal in FP32 builds (-cl-single-precision-constant) or a double
# literal in FP64 builds.
# ---------------------------------------------------------------------------
_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
_libc.strtod.restype = ctypes.c_double
_libc.strtod.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def kernel_literal(value: float, precision: str) -> float:
    if precision == "f32":
        text = "%g" % float(np.float32(value))
        return float(_libc.strtof(text.encode(), None))
    text = "%g" % float(value)
    return float(_libc.strtod(text.encode(), None))


ARITH_EXACT, ARITH_FUSED, ARITH_CONTRACTED = 0, 1, 2


def params(dim, stride, viscosity, velocity, precision="f32", threads=None, fast=False) -> Params:
    """`fast`: False/True for the exact or fused collision, or ARITH_CONTRACTED
    for kernels.cl's expressions contracted as clang compiles OpenCL C."""
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    return Params(dim, threads, stride, kernel_literal(viscosity, precision), kernel_literal(velocity, precision),
                  int(fast), 0)


def constants(p: Params, precision="f32"):
    """(inv_tau, U, w[19]) exactly as the kernels use them."""
    dt = np.float32 if precision == "f32" else np.float64
    out = np.zeros(2 + Q, dtype=dt)
    getattr(lib(), f"lbmo_constants_{precision}")(ctypes.byref(p), _ptr(out))
    return out[0], out[1], out[2:]


def cell_type(N, x, y, z) -> int:
    return lib().lbmo_cell_type(N, x, y, z)


def slot(cell_id, q, stride) -> int:
    return lib().lbmo_slot(cell_id, q, stride)


def shift_scale(U: float) -> float:
    return lib().lbmo_shift_scale(U)


def f32_to_f16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    out = np.empty(a.shape, dtype=np.uint16)
    lib().lbmo_f32_to_f16(_ptr(a), _ptr(out), a.size)
    return out


def f16_to_f32(a):
    a = np.ascontiguousarray(a, dtype=np.uint16)
    out = np.empty(a.shape, dtype=np.float32)
    lib().lbmo_f16_to_f32(_ptr(a), _ptr(out), a.size)
    return out


def f32_to_bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    out = np.empty(a.shape, dtype=np.uint16)
    lib().lbmo_f32_to_bf16(_ptr(a), _ptr(out), a.size)
    return out


def bf16_to_f32(a):
    a = np.ascontiguousarray(a, dtype=np.uint16)
    out = np.empty(a.shape, dtype=np.float32)
    lib().lbmo_bf16_to_f32(_ptr(a), _ptr(out), a.size)
    return out


class PushSim:
    """The reference form, driven like LBMCL<T> (lbmcl.hpp:435-521)."""

    def __init__(self, dim, stride=None, viscosity=0.0089, velocity=0.05, precision="f32", threads=None, fast=False):
        self.N = dim
        self.stride = stride if stride is not None else dim
        self.precision = precision
        self.dt = np.float32 if precision == "f32" else np.float64
        self.p = params(dim, self.stride, viscosity, velocity, precision, threads, fast)
        n3 = dim ** 3
        self.f_stream = np.empty(n3 * Q, dtype=self.dt)
        self.f_collide = np.empty(n3 * Q, dtype=self.dt)
        self.rho = np.empty(n3, dtype=self.dt)
        self.u = np.empty(3 * n3, dtype=self.dt)
        self.map = np.empty(n3, dtype=np.int32)
        self.it = 0
        self.force = None

    def initialize(self):
        getattr(lib(), f"lbmo_initialize_{self.precision}")(
            ctypes.byref(self.p), _ptr(self.f_stream), _ptr(self.f_collide), _ptr(self.rho), _ptr(self.u),
            _ptr(self.map))
        self.it = 0

    def input_buffer(self, it):
        """The lattice that iteration `it` reads (lbmcl.hpp:437, 456-457)."""
        return self.f_stream if it % 2 == 0 else self.f_collide

    def step(self, update_macro=True):
        it = self.it + 1
        src = self.input_buffer(it)
        dst = self.f_collide if it % 2 == 0 else self.f_stream
        getattr(lib(), f"lbmo_compute_{self.precision}")(
            ctypes.byref(self.p), _ptr(dst), _ptr(src), _ptr(self.rho), _ptr(self.u), _ptr(self.map),
            int(bool(update_macro)), _ptr(self.force))
        self.it = it

    def macro(self):
        n3 = self.N ** 3
        return self.rho.copy(), self.u.reshape(3, n3).copy()


class PullSim:
    """The gather form (native, FP16- or BF16-shifted storage).

    fast selects the fused collision (or ARITH_CONTRACTED the contracted
    expressions); the shifted storage forms always use the deviation form."""

    def __init__(self, dim, stride=None, viscosity=0.0089, velocity=0.05, precision="f32",
                 storage=STORE_NATIVE, threads=None, fast=False):
        if storage != STORE_NATIVE and precision != "f32":
            raise ValueError("shifted storage computes in FP32")
        self.N = dim
        self.stride = stride if stride is not None else dim
        self.precision = precision
        self.storage = storage
        self.dt = np.float32 if precision == "f32" else np.float64
        self.st = self.dt if storage == STORE_NATIVE else np.uint16
        self.fast = ARITH_FUSED if storage != STORE_NATIVE else int(fast)
        self.p = params(dim, self.stride, viscosity, velocity, precision, threads, self.fast)
        n3 = dim ** 3
        self.A = np.empty(n3 * Q, dtype=self.st)   # read by odd iterations
        self.B = np.empty(n3 * Q, dtype=self.st)   # read by even iterations
        self.rho = np.empty(n3, dtype=self.dt)
        self.u = np.empty(3 * n3, dtype=self.dt)
        self.map = np.empty(n3, dtype=np.int32)
        self.it = 0
        self.force = None
        self.diag = np.zeros(2, dtype=np.float64)

    def initialize(self):
        L = lib()
        if self.precision == "f32":
            L.lbmo_pull_initialize_f32(ctypes.byref(self.p), self.storage, _ptr(self.A), _ptr(self.B),
                                       _ptr(self.rho), _ptr(self.u), _ptr(self.map))
        else:
            L.lbmo_pull_initialize_f64(ctypes.byref(self.p), _ptr(self.A), _ptr(self.B), _ptr(self.rho),
                                       _ptr(self.u), _ptr(self.map))
        self.it = 0

    def input_buffer(self, it):
        return self.B if it % 2 == 0 else self.A

    def step(self, update_macro=True):
        it = self.it + 1
        src = self.input_buffer(it)
        dst = self.A if it % 2 == 0 else self.B
        L = lib()
        if self.precision == "f32":
            L.lbmo_pull_compute_f32(ctypes.byref(self.p), self.storage, _ptr(src), _ptr(dst), _ptr(self.rho),
                                    _ptr(self.u), int(bool(update_macro)), _ptr(self.force), _ptr(self.diag))
        else:
            L.lbmo_pull_compute_f64(ctypes.byref(self.p), _ptr(src), _ptr(dst), _ptr(self.rho), _ptr(self.u),
                                    int(bool(update_macro)), _ptr(self.force), _ptr(self.diag))
        self.it = it

    def view(self, buf):
        """The reference's CSoA lattice corresponding to pull buffer `buf`."""
        out = np.empty(self.N ** 3 * Q, dtype=self.dt)
        if self.precision == "f32":
            lib().lbmo_pull_view_f32(ctypes.byref(self.p), self.storage, _ptr(buf), _ptr(out))
        else:
            lib().lbmo_pull_view_f64(ctypes.byref(self.p), _ptr(buf), _ptr(out))
        return out

    def macro(self):
        n3 = self.N ** 3
        return self.rho.copy(), self.u.reshape(3, n3).copy()


def diagnostics(p: Params, rho, u, precision="f32"):
    out = np.zeros(2, dtype=np.float64)
    getattr(lib(), f"lbmo_diagnostics_{precision}")(ctypes.byref(p), _ptr(rho), _ptr(u), _ptr(out))
    return out


# ---------------------------------------------------------------------------
# The reference's own kernels on the GPU (oracle/_ref, see ref_build.py).
# ---------------------------------------------------------------------------
def ref_unavailable_reason():
    """None when oracle/_ref is built and an OpenCL GPU device is present, else why not."""
    from . import ref_build

    if not os.path.exists(ref_build.RUNNER):
        return "oracle/_ref is not built (it needs the reference checkout at build time)"
    r = subprocess.run([ref_build.RUNNER, "--probe"], capture_output=True, text=True, timeout=120,
                       stdin=subprocess.DEVNULL)
    if r.returncode != 0:
        return f"no OpenCL GPU device ({r.stdout.strip()} {r.stderr.strip()})"
    return None


def ref_run(precision, dim, stride, lws, iterations, out_prefix=None, every=None, skip=0, fast=False,
            timeout=600):
    """Run the reference's initialize + `iterations` compute kernels (lbmcl.hpp:490-521).

    `every` defaults to `iterations` (rho/u stored by the last iteration).  Returns
    {"mlups", "kernels_ms", "device"} and, with `out_prefix`, the arrays "rho" (N^3),
    "u" (3 x N^3) and "f" (19 N^3, CSoA, the lattice iteration iterations + 1 reads)."""
    from . import ref_build

    co = ref_build.co_path(precision, dim, stride, fast)
    if not os.path.exists(co):
        raise FileNotFoundError(f"{co}: no reference build for this configuration (see ref_build.CONFIGS)")
    every = iterations if every is None else every
    cmd = [ref_build.RUNNER, co, precision, str(dim), str(stride), ",".join(map(str, lws)), str(iterations),
           str(every), str(skip), out_prefix or "-"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, stdin=subprocess.DEVNULL)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed ({r.returncode}): {r.stderr.strip()}")
    words = r.stdout.split()
    out = {"mlups": float(words[words.index("mlups") + 1]),
           "kernels_ms": float(words[words.index("kernels_ms") + 1]),
           "device": words[1]}
    if out_prefix:
        dt = np.float64 if precision == "f64" else np.float32
        n3 = dim ** 3
        out["rho"] = np.fromfile(out_prefix + ".rho", dtype=dt)
        out["u"] = np.fromfile(out_prefix + ".u", dtype=dt).reshape(3, n3)
        out["f"] = np.fromfile(out_prefix + ".f", dtype=dt)
    return out
