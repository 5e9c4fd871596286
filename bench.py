#!/usr/bin/env python3
"""Throughput of the D3Q19 lid-driven-cavity collide-stream step on one MI355X.

Workload: the 512^3 cavity (nu = 0.0089, lid u = 0.05), FP32 populations in
the CSoA layout with stride 4096 (the fastest stride measured on MI355X: 16 KB
runs of each population; stride 64 is ~5% slower), the vector row kernel
(8-byte lanes), the default arithmetic (lbm_config.optimize = 0): kernels.cl's
expressions as ROCm's OpenCL compiler evaluates them, i.e. the reference's
results bit for bit.  A "step" is one collide-stream iteration of the
whole lattice.  The timed region
holds exactly --steps iterations enqueued back to back on the library's stream,
bracketed by device synchronisation; the lattice is resident in HBM and already
initialised before it starts.

Printed: one JSON line with MLUPS (the reference's definition,
(dim-2)^3 * iterations / time, lbmcl.hpp:601-620), the HBM roofline of the
compute kernel (152 B per lattice update in FP32; average kernel duration from
HIP events recorded on the kernel's stream around every launch of the timed
region), the same figures for the 16-bit shifted-storage forms, FP64, the fused
collision (-o), the CPU oracle's exact arithmetic (LBM_ARITH_EXACT), the
128^3 configuration and a 1024^3 FP32 lattice, the cost of the adjoint on the
256^3 x 2000 configuration (forward and adjoint time, their ratio, peak device
memory), a copy kernel's sustained bandwidth for context, and the CPU oracle
timed on a bounded sample of the same workload (cpu_baseline).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = {("f32", "native"): 152, ("f64", "native"): 304, ("f32", "fp16"): 76, ("f32", "bf16"): 76}
VARIANT_KEYS = ("mlups", "mlups_with_events", "kernel_mlups", "avg_kernel_ms", "achieved_gbs", "frac",
                "bytes_per_update", "steps")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_lattice(dim, steps, warmup, precision="f32", storage="native", stride=None, kernel=0, xcd_remap=True,
                arith=0):
    """Two passes over the same iterations: with the reference's per-launch events
    (the compute kernels' own durations: kernel time, roofline) and without them
    (the wall time of the launches back to back: `mlups`, `ms_per_step`).
    `arith`: lbm_config.optimize (0 the reference's OpenCL rounding, 1 fused, 2 exact)."""
    from lbmcl_amd import capi
    stride = stride or dim ** 3
    walls = {}
    for profile in (True, False):
        with capi.Context(dim, stride=stride, precision=precision, storage=storage, kernel=kernel, profile=profile,
                          xcd_remap=xcd_remap, arith=arith) as g:
            g.initialize()
            it = 1
            if warmup:
                g.compute_range(it, warmup, 0)
                it += warmup
            g.finish()
            g.clear_events()
            t0 = time.perf_counter()
            g.compute_range(it, steps, 0)
            g.finish()
            walls[profile] = time.perf_counter() - t0
            if profile:
                kernel_ms = g.kernels_time_ms()
                events = g.events()
                total_ms = g.total_time_ms()
    wet = (dim - 2) ** 3
    wall_s = walls[False]
    avg_kernel_s = kernel_ms / 1e3 / steps
    b = BYTES_PER_UPDATE[(precision, storage)]
    achieved = wet * b / avg_kernel_s / 1e9
    return {
        "dim": dim, "precision": precision, "storage": storage, "stride": stride, "steps": steps,
        "wall_s": wall_s, "ms_per_step": wall_s * 1e3 / steps,
        "mlups": wet * steps / wall_s / 1e6,
        "mlups_with_events": wet * steps / walls[True] / 1e6,
        "kernel_mlups": wet * steps / (kernel_ms / 1e3) / 1e6,
        "avg_kernel_ms": kernel_ms / steps,
        "event_span_ms": total_ms,
        "n_events": len(events),
        "bytes_per_update": b,
        "achieved_gbs": achieved,
        "frac": achieved / HBM_PEAK_GBS,
    }


def copy_bandwidth(nbytes=4 << 30, reps=5):
    path = os.path.join(ROOT, "lbmcl_amd", "libmembw.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    L.membw_copy.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    ms = (ctypes.c_double * reps)()
    if L.membw_copy(0, nbytes, reps, ms) != 0:
        return None
    best = min(ms[1:]) if reps > 1 else ms[0]
    return {"bytes": nbytes, "best_ms": best, "gbs": 2 * nbytes / (best / 1e3) / 1e9}


def access_pattern_bandwidth(stride, reps=6):
    """The FP32 step's data movement at 512^3 without its arithmetic (membw.hip
    pattern_kernel): algorithmic GB/s (152 B per wet cell) at full occupancy (8
    waves per SIMD), at five -- the FP32 vector kernel's occupancy (88 VGPRs) --
    at four and at two (30, 40 and 64 KiB of LDS per 4-wave block; 160 KiB per
    CU)."""
    path = os.path.join(ROOT, "lbmcl_amd", "libmembw.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    L.membw_pattern.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_double)]
    out = {"dim": 512, "stride": stride, "bytes_per_cell": 152}
    for name, lds in (("full_occupancy", 0), ("five_waves_per_simd", 30720), ("four_waves_per_simd", 40960),
                      ("two_waves_per_simd", 65536)):
        ms = (ctypes.c_double * reps)()
        if L.membw_pattern(0, 512, stride, lds, reps, ms) != 0:
            return None
        best = min(ms[1:])
        out[name] = {"best_ms": best, "gbs": 510 ** 3 * 152 / (best / 1e3) / 1e9}
    return out


def cpu_baseline(dim, steps, threads, stride):
    """The CPU oracle (reference push form, OpenMP) on `steps` iterations of the same lattice."""
    import oracle as O
    O.build()
    sim = O.PushSim(dim, stride=stride, precision="f32", threads=threads)
    sim.initialize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step(update_macro=False)
    dt = time.perf_counter() - t0
    return {
        "value": (dim - 2) ** 3 * steps / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
        "sample": f"oracle push form, {dim}^3 FP32 stride {stride}, {steps} iterations, {threads} OpenMP threads, "
                  f"{dt:.1f} s",
    }


# the sources that define the compute kernels' code
KERNEL_SOURCES = ("lbmcl_amd/csrc/lbm_kernels.hpp", "lbmcl_amd/csrc/lbm_lattice.hpp",
                  "lbmcl_amd/csrc/lbm_opencl_arith.hip", "lbmcl_amd/csrc/lbm_ctx.hpp")


def kernel_sources_sha():
    """sha256 (first 16 hex digits) of the kernel sources, in KERNEL_SOURCES order."""
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def traffic_from_profiles(dim, stride, precision, storage):
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, as
    committed under profiles/ together with the kernel sources' hash they were
    measured on: (bytes, profile file, its hash).  The bytes are None when the
    profile is missing or was measured on other kernel sources."""
    name = f"collide_stream_vec_{dim}_s{stride}_{precision}_{storage}_pmc.md"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, name, None
    with open(path) as fh:
        text = fh.read()
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    h = re.search(r"kernel_sources_sha:\s*([0-9a-f]+)", text)
    sha = h.group(1) if h else None
    if not m or sha != kernel_sources_sha():
        return None, name, sha
    return float(m.group(1)), name, sha


def adjoint_cost(dim=256, steps=2000, every=45, seed=0):
    """The adjoint configuration: dim^3 cavity, FP32 forward and adjoint with FP64
    gradient sums, `steps` iterations, checkpoints every `every` steps, J = mean
    u_x on the mid-plane z = dim/2, a seeded Gaussian random body force
    (correlation 8 cells, amplitude 1e-6).  Times from the library's events."""
    from lbmcl_amd.adjoint import cavity_gradient, gaussian_random_field
    g = gaussian_random_field(dim, 8.0, 1e-6, seed=seed).astype("float32")
    t0 = time.perf_counter()
    r = cavity_gradient(dim, steps, force=g, objective="mean_ux_plane", plane_z=dim // 2, precision="f32",
                        checkpoint_every=every, grad_force=True)
    wall = time.perf_counter() - t0
    return {"dim": dim, "steps": steps, "checkpoint_every": r["segment"], "checkpoints": r["checkpoints"],
            "objective": "mean u_x on z = %d" % (dim // 2), "forward_ms": round(r["forward_ms"], 2),
            "adjoint_ms": round(r["adjoint_ms"], 2), "adjoint_over_forward": round(r["adjoint_ms"] / r["forward_ms"], 3),
            "device_bytes": r["device_bytes"], "device_gib": round(r["device_bytes"] / 2 ** 30, 2),
            "wall_s": round(wall, 2), "J": r["objective"], "dJ_dU": r["d_velocity"], "dJ_dnu": r["d_viscosity"]}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dim", type=int, default=512)
    ap.add_argument("--stride", type=int, default=4096, help="CSoA stride")
    ap.add_argument("--variant-steps", type=int, default=300, help="steps of the FP16/BF16/FP64 runs (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=12,
                    help="iterations of the CPU oracle sample, ~10 s on 16 threads (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-remap", action="store_true", help="disable the XCD-aware block order")
    ap.add_argument("--kernel", type=int, default=0, help="lbm_config.kernel: 0 auto, 1 generic, 2 row, 3 vector (16-byte lanes), 4 vector (8-byte), 5 tile, 6 row pipeline")
    ap.add_argument("--adjoint-steps", type=int, default=2000,
                    help="iterations of the 256^3 adjoint cost measurement (0: skip)")
    ap.add_argument("--no-calibration", action="store_true",
                    help="skip the copy and access-pattern legs (libmembw); used for runs under rocprofv3")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for {args.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        return 2
    rank = int(os.environ.get("RANK", "0"))
    dim = args.dim
    stride = args.stride

    log(f"[bench] main run: {dim}^3 f32 native stride {stride}, warmup {args.warmup}, steps {args.steps}")
    main_run = run_lattice(dim, args.steps, args.warmup, "f32", "native", stride, args.kernel, not args.no_remap)
    log(f"[bench] main: {main_run['mlups']:.0f} MLUPS, kernel {main_run['avg_kernel_ms']:.3f} ms, "
        f"{main_run['achieved_gbs']:.0f} GB/s")

    variants = {}
    if args.variant_steps > 0:
        for name, prec, st, arith in (("f32_fp16", "f32", "fp16", 0), ("f32_bf16", "f32", "bf16", 0),
                                      ("f64_native", "f64", "native", 0), ("f32_native_fused", "f32", "native", 1),
                                      ("f32_native_exact_arith", "f32", "native", 2),
                                      ("f64_native_exact_arith", "f64", "native", 2)):
            log(f"[bench] variant {name}")
            r = run_lattice(dim, args.variant_steps, max(5, args.warmup // 4), prec, st, stride, args.kernel,
                            not args.no_remap, arith)
            variants[name] = {k: r[k] for k in VARIANT_KEYS}
            log(f"[bench] {name}: {r['mlups']:.0f} MLUPS, {r['achieved_gbs']:.0f} GB/s")
        variants["fp16_speedup_vs_f32"] = variants["f32_fp16"]["mlups"] / main_run["mlups"]
        # the 128^3 configuration: MLUPS over iterations 100..1000 (99 untimed first)
        for prec in ("f32", "f64"):
            name = f"cavity_128_{prec}_it100_1000"
            log(f"[bench] variant {name}")
            r = run_lattice(128, 901, 99, prec, "native", 64, args.kernel, not args.no_remap)
            variants[name] = {k: r[k] for k in VARIANT_KEYS}
            log(f"[bench] {name}: {r['mlups']:.0f} MLUPS, {r['achieved_gbs']:.0f} GB/s")
        # 1024^3 FP32: 179 GB of HBM, a lattice the reference cannot index
        if dim == 512:
            from lbmcl_amd import capi
            name = "f32_native_1024"
            log(f"[bench] variant {name}")
            try:
                r = run_lattice(1024, 50, 5, "f32", "native", stride, args.kernel, not args.no_remap)
                variants[name] = {k: r[k] for k in VARIANT_KEYS}
                log(f"[bench] {name}: {r['mlups']:.0f} MLUPS, {r['achieved_gbs']:.0f} GB/s")
            except capi.LbmError as e:  # e.g. less free HBM than the 179 GB it needs
                variants[name] = {"skipped": str(e)}
                log(f"[bench] {name} skipped: {e}")

    adjoint = None
    if args.adjoint_steps > 0:
        log(f"[bench] adjoint: 256^3 x {args.adjoint_steps}, FP32, checkpoints every 45 steps")
        adjoint = adjoint_cost(256, args.adjoint_steps, 45)
        log(f"[bench] adjoint: forward {adjoint['forward_ms']:.0f} ms, adjoint {adjoint['adjoint_ms']:.0f} ms "
            f"({adjoint['adjoint_over_forward']:.2f}x), {adjoint['device_gib']:.1f} GiB")

    copy = pattern = None
    if not args.no_calibration:
        log("[bench] copy bandwidth")
        copy = copy_bandwidth()
        pattern = access_pattern_bandwidth(stride) if dim == 512 else None
    cpu = None
    if args.cpu_steps > 0 and rank == 0:
        log(f"[bench] cpu baseline: {args.cpu_steps} oracle iterations at {dim}^3")
        cpu = cpu_baseline(dim, args.cpu_steps, args.cpu_threads, stride)

    wet = (dim - 2) ** 3
    traffic = traffic_from_profiles(dim, stride, "f32", "native")
    out = {
        "metric": f"MLUPS (D3Q19 lid-driven cavity, {dim}^3, FP32)",
        "value": round(main_run["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (deterministic lid-driven cavity initial state; no dataset)",
        "config": {"workload": f"D3Q19 BGK lid-driven cavity {dim}^3, nu=0.0089, lid u=0.05, FP32 populations, "
                               f"CSoA stride {stride}, pull-form vector row kernel, the reference's arithmetic "
                               "(kernels.cl as ROCm's OpenCL compiler evaluates it; lbm_config.optimize = 0)",
                   "dim": dim, "stride": stride, "wet_cells": wet, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(main_run["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(main_run["frac"], 4),
                     "traffic": traffic[0],
                     "traffic_profile": {"file": "profiles/" + traffic[1], "kernel_sources_sha": traffic[2],
                                         "current_sources_sha": kernel_sources_sha()},
                     "kernel": "lbm_collide_stream_vec<float, native, 8-byte lanes, OpenCL arithmetic>",
                     "avg_kernel_ms": round(main_run["avg_kernel_ms"], 4),
                     "bytes_per_launch": wet * 152},
        "cpu_baseline": cpu,
        "kernel_mlups": round(main_run["kernel_mlups"], 1),
        "variants": variants,
        "adjoint": adjoint,
        "copy_bandwidth": copy,
        "access_pattern_bandwidth": pattern,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
