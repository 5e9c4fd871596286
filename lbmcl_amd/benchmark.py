"""Benchmark driver of the lbmcl command line (benchmark.sh:1-176 counterpart).

Runs `lbmcl_amd/bin/lbmcl` over a sweep of lattice sizes, work-group sizes and
CSoA strides, `--repeats` times each, appends every run's stderr statistics
line (lbmcl.hpp:648-669:
device;precision;dim;iterations;every;lws;stride;optimize;total_ms;kernels_ms;MLUPS;kernels_MLUPS)
to `<out>/stats.csv`, then aggregates them into `<out>/benchmark.csv` exactly
as benchmark.sh's awk program does:

* runs are grouped by device;precision;dim;lws;stride (fields 1, 2, 3, 6, 7);
* in each group the run with the smallest and the run with the largest total
  time (field 9; the first of equal values) are dropped, and the other runs'
  four timing fields are averaged (sums minus those two rows, divided by
  count - 2: at least 3 repeats are needed);
* each group prints `key;total_ms;kernels_ms;MLUPS;kernels_MLUPS`, numbers in
  awk's output format (integers as integers, otherwise %.6g).
Groups are printed in sorted order (awk's `for (i in names)` order is
unspecified).

The sweep of benchmark.sh: dims 8..128, lws x >= y >= z from {1..128} with
x*y*z <= 1024 and every component <= dim, strides {1, 8, 16, 32, 64, 128},
then stride dim^3 for every lws; 50 iterations, no output, fast math (-o).
The defaults here keep that shape; on the HIP build the work-group size only
reaches the generic kernel (dim < 64 for FP32), larger lattices run the vector
kernels, so `--lws` can be cut to one value for them.  The wall time of every
process (output files included) goes to `<out>/wall.csv`.

Usage: python -m lbmcl_amd.benchmark [--dims 8,16,32,64,128] [--repeats 10] [--out benchmarks] ...
"""
from __future__ import annotations

import argparse
import itertools
import os
import subprocess
import sys
import time

from . import build as native_build

STAT_FIELDS = 12


def awk_number(x: float) -> str:
    """awk's number-to-string conversion (CONVFMT "%.6g", integral values as integers)."""
    if x == int(x) and abs(x) < 9.2e18:
        return str(int(x))
    return "%.6g" % x


def parse_stat(line: str):
    """Fields of one statistics line, or None for any other line."""
    f = line.rstrip("\n").split(";")
    if len(f) != STAT_FIELDS:
        return None
    try:
        [float(v) for v in f[8:12]]
    except ValueError:
        return None
    return f


def aggregate(lines):
    """benchmark.sh:111-176 on the statistics lines; returns the output lines."""
    groups = {}
    for line in lines:
        f = parse_stat(line)
        if f is None:
            continue
        key = ";".join([f[0], f[1], f[2], f[5], f[6]])
        vals = [float(v) for v in f[8:12]]
        g = groups.get(key)
        if g is None:
            g = groups[key] = {"count": 0, "sum": [0.0] * 4, "min": vals, "max": vals}
        g["count"] += 1
        g["sum"] = [a + b for a, b in zip(g["sum"], vals)]
        if g["min"][0] > vals[0]:
            g["min"] = vals
        if g["max"][0] < vals[0]:
            g["max"] = vals
    out = []
    for key in sorted(groups):
        g = groups[key]
        n = g["count"] - 2
        if n <= 0:
            raise ValueError(f"{key}: {g['count']} runs; the trimmed mean needs at least 3")
        avg = [(s - lo - hi) / n for s, lo, hi in zip(g["sum"], g["min"], g["max"])]
        out.append(key + ";" + ";".join(awk_number(v) for v in avg))
    return out


def lws_sweep(values, dim):
    """(x, y, z) with x >= y >= z from `values`, x*y*z <= 1024, all <= dim (benchmark.sh:62-69)."""
    for x, y, z in itertools.product(values, repeat=3):
        if y <= x and z <= y and x * y * z <= 1024 and x <= dim and y <= dim and z <= dim:
            yield x, y, z


def run_one(cli, args, stats_fh, wall_fh, key):
    t0 = time.perf_counter()
    p = subprocess.run([cli, *args], stdin=subprocess.DEVNULL, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                       text=True)
    wall = time.perf_counter() - t0
    if p.returncode != 0:
        raise RuntimeError(f"lbmcl {' '.join(args)} failed ({p.returncode}): {p.stderr.strip()}")
    stats_fh.write(p.stderr)
    stats_fh.flush()
    wall_fh.write(f"{key};{wall:.6f}\n")
    wall_fh.flush()


def _ints(s):
    return [int(v) for v in s.split(",") if v]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default="benchmarks")
    ap.add_argument("--precision", choices=["single", "double"], default="single")
    ap.add_argument("--platform", type=int, default=0)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--viscosity", default="0.0089")
    ap.add_argument("--velocity", default="0.05")
    ap.add_argument("--iterations", type=int, default=50)
    ap.add_argument("--every", type=int, default=0)
    ap.add_argument("--dims", default="8,16,32,64,128")
    ap.add_argument("--lws", default="1,2,4,8,16,32,64,128")
    ap.add_argument("--strides", default="1,8,16,32,64,128")
    ap.add_argument("--no-soa", action="store_true", help="skip the second sweep with stride dim^3")
    ap.add_argument("--repeats", type=int, default=10)
    ap.add_argument("--storage", choices=["native", "fp16", "bf16"], default="native")
    ap.add_argument("--arith", choices=["exact", "opencl"], default=None,
                    help="pass --arith to lbmcl (opencl: the reference's own rounding; excludes -o)")
    ap.add_argument("--no-optimize", action="store_true", help="omit -o (benchmark.sh always passes it)")
    ap.add_argument("--vtk-path", default="", help="pass -v (with --every > 0, files are written)")
    ap.add_argument("--cli", default=native_build.CLI)
    a = ap.parse_args(argv)
    if a.repeats < 3:
        ap.error("--repeats must be at least 3 (two runs of each group are trimmed)")
    if not os.path.exists(a.cli):
        print(f"{a.cli} is not built: run python -m lbmcl_amd.build", file=sys.stderr)
        return 2
    os.makedirs(a.out, exist_ok=True)
    stats_path = os.path.join(a.out, "stats.csv")
    bench_path = os.path.join(a.out, "benchmark.csv")
    wall_path = os.path.join(a.out, "wall.csv")
    for p in (stats_path, bench_path, wall_path):
        if os.path.exists(p):
            os.remove(p)

    def args_for(d, lws, s):
        r = ["-P", str(a.platform), "-D", str(a.device), "-d", str(d), "-n", a.viscosity, "-u", a.velocity,
             "-i", str(a.iterations), "-e", str(a.every), "-w", ",".join(map(str, lws)), "-s", str(s)]
        if not a.no_optimize and a.arith is None:
            r.append("-o")
        if a.arith is not None:
            r += ["--arith", a.arith]
        if a.precision == "double":
            r.append("-F")
        if a.storage != "native":
            r += ["--storage", a.storage]
        if a.vtk_path:
            r += ["-v", a.vtk_path]
        return r

    dims, lws_values, strides = _ints(a.dims), _ints(a.lws), _ints(a.strides)
    runs = 0
    with open(stats_path, "a") as st, open(wall_path, "a") as wl:
        for sweep in ("csoa", "soa"):
            if sweep == "soa" and a.no_soa:
                continue
            for d in dims:
                for lws in lws_sweep(lws_values, d):
                    for s in (strides if sweep == "csoa" else [d ** 3]):
                        if s > d ** 3:
                            continue
                        key = f"{d};{lws[0]},{lws[1]},{lws[2]};{s}"
                        for _ in range(a.repeats):
                            run_one(a.cli, args_for(d, lws, s), st, wl, key)
                            runs += 1
                        print(f"[benchmark] dim {d} lws {lws} stride {s}: {a.repeats} runs", file=sys.stderr,
                              flush=True)
    with open(stats_path) as fh:
        out = aggregate(fh)
    with open(bench_path, "w") as fh:
        fh.write("".join(line + "\n" for line in out))
    print(f"[benchmark] {runs} runs, {len(out)} groups -> {bench_path}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
