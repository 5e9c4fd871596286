"""Summaries of rocprofv3 output for the profiles/ directory.

  python -m lbmcl_amd.rocprof_summary trace DIR
      kernel trace (rocprofv3 --kernel-trace ... --output-format csv -d DIR):
      per kernel name and grid size, launches and average / min / max / total
      duration in ms.
  python -m lbmcl_amd.rocprof_summary gaps DIR [--min-launches N]
      the same trace split into bursts -- maximal runs of consecutive
      dispatches of one kernel at one grid size, at least N (default 50) --
      with, per burst, the average kernel duration, the average / median /
      largest gap between one dispatch's end and the next one's start, and
      the share of the burst's span spent inside the kernel.
  python -m lbmcl_amd.rocprof_summary counters DIR [DIR ...]
      counter collection (rocprofv3 --pmc ... -d DIR): per kernel name and grid
      size, the average value of every counter over its dispatches.
  python -m lbmcl_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR [--calibration FETCH_DIR WRITE_DIR]
      HBM traffic per dispatch from FETCH_SIZE and WRITE_SIZE (KiB), scaled by
      what the same counters read for membw's copy kernel (4 GiB read and 4 GiB
      written per dispatch, 16-byte streaming loads and stores like the
      collide-stream kernels' accesses), from runs of `lbmcl_amd/bin/membw
      copy` given by --calibration, or else from the same runs; also prints
      the kernel sources' hash (`kernel_sources_sha`) that bench.py compares
      before it reports the traffic of its headline kernel.  (The counters'
      reading depends on the kind of request: torch's elementwise kernels,
      with ordinary loads, read half of what this copy reads per byte, so the
      calibration kernel has to access memory the way the measured kernels do.)

Markdown tables on stdout.  The CSV columns are found by name, so the files of
every rocprofv3 version that writes Kernel_Name / Grid_Size / Start_Timestamp /
End_Timestamp (trace) or Counter_Name / Counter_Value (counters) work.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import sys
from collections import defaultdict

COPY_BYTES = 4 << 30  # the calibration kernels: 4 GiB read and 4 GiB written per dispatch
CALIBRATION_KERNELS = ("copy_kernel",)  # membw.hip's copy


def _rows(directory, suffix):
    files = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {directory}")
    for path in files:
        with open(path, newline="") as fh:
            yield from csv.DictReader(fh)


def _grid(row):
    for key in ("Grid_Size", "Grid_Size_X"):
        if key in row and row[key]:
            g = int(float(row[key]))
            for extra in ("Grid_Size_Y", "Grid_Size_Z"):
                if extra in row and row[extra]:
                    g *= int(float(row[extra]))
            return g
    return 0


def _short(name):
    """The kernel's name without `void `, `(anonymous namespace)::` and its parameter list."""
    name = name.strip().replace("(anonymous namespace)::", "")
    if name.startswith("void "):
        name = name[5:]
    depth = 0
    for i, ch in enumerate(name):  # cut at the parameter list, not inside template arguments
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i]
    return name


def trace(directory):
    acc = defaultdict(list)
    for row in _rows(directory, "kernel_trace.csv"):
        ms = (int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e6
        acc[(_short(row["Kernel_Name"]), _grid(row))].append(ms)
    out = ["| kernel | grid | launches | avg ms | min ms | max ms | total ms |", "|---|---|---|---|---|---|---|"]
    for (name, grid), v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"| `{name}` | {grid:,} | {len(v)} | {sum(v) / len(v):.4f} | {min(v):.4f} | {max(v):.4f} | "
                   f"{sum(v):.1f} |")
    return "\n".join(out)


def gaps(directory, min_launches=50):
    rows = []
    for row in _rows(directory, "kernel_trace.csv"):
        rows.append((int(row["Start_Timestamp"]), int(row["End_Timestamp"]), _short(row["Kernel_Name"]), _grid(row)))
    rows.sort()
    bursts, cur = [], []
    for r in rows:
        if cur and (r[2], r[3]) != (cur[-1][2], cur[-1][3]):
            bursts.append(cur)
            cur = []
        cur.append(r)
    if cur:
        bursts.append(cur)
    out = ["| burst | kernel | grid | launches | avg kernel us | avg gap us | median gap us | max gap us | "
           "kernel share of span |", "|---|---|---|---|---|---|---|---|---|"]
    n = 0
    for b in bursts:
        if len(b) < min_launches:
            continue
        n += 1
        dur = [(e - s) / 1e3 for s, e, _, _ in b]
        gap = sorted((b[i + 1][0] - b[i][1]) / 1e3 for i in range(len(b) - 1))
        span = (b[-1][1] - b[0][0]) / 1e3
        out.append(f"| {n} | `{b[0][2]}` | {b[0][3]:,} | {len(b)} | {sum(dur) / len(dur):.2f} | "
                   f"{sum(gap) / len(gap):.2f} | {gap[len(gap) // 2]:.2f} | {gap[-1]:.2f} | "
                   f"{sum(dur) / span:.4f} |")
    return "\n".join(out)


def counters(directories):
    acc = defaultdict(lambda: defaultdict(list))
    for d in directories:
        for row in _rows(d, "counter_collection.csv"):
            acc[(_short(row["Kernel_Name"]), _grid(row))][row["Counter_Name"]].append(float(row["Counter_Value"]))
    names = sorted({c for v in acc.values() for c in v})
    out = ["| kernel | grid | " + " | ".join(names) + " |", "|---|---|" + "---|" * len(names)]
    for (kernel, grid), v in sorted(acc.items()):
        cells = [f"{sum(v[c]) / len(v[c]):,.0f}" if v.get(c) else "" for c in names]
        out.append(f"| `{kernel}` | {grid:,} | " + " | ".join(cells) + " |")
    return "\n".join(out), acc


def _calibration_value(acc, counter, what):
    keys = [k for k in acc if any(c in k[0] for c in CALIBRATION_KERNELS) and counter in acc[k]]
    if not keys:
        raise SystemExit(f"no calibration kernel ({' or '.join(CALIBRATION_KERNELS)}) in the {what} run; kernels: "
                         + ", ".join(sorted({k[0][:60] for k in acc})))
    key = max(keys, key=lambda k: k[1])  # the 4 GiB dispatch
    v = acc[key][counter]
    return key[0], sum(v) / len(v)


def traffic(fetch_dir, write_dir, headline=None, calibration=None):
    """headline: (kernel name, grid) whose traffic bench.py reports; printed as
    `traffic_bytes_per_launch:` next to the sources' hash.  calibration:
    (fetch dir, write dir) of the `calibrate` workload, else the copy kernel
    is taken from the runs themselves."""
    _, f = counters([fetch_dir])
    _, w = counters([write_dir])
    _, cfa = counters([calibration[0]]) if calibration else (None, f)
    _, cwa = counters([calibration[1]]) if calibration else (None, w)
    cname, cf = _calibration_value(cfa, "FETCH_SIZE", "fetch")
    _, cw = _calibration_value(cwa, "WRITE_SIZE", "write")
    sf, sw = COPY_BYTES / (cf * 1024), COPY_BYTES / (cw * 1024)
    out = [f"Calibration: `{cname[:80]}` (4 GiB read and 4 GiB written per dispatch) reads FETCH_SIZE {cf:,.0f} "
           f"KiB and WRITE_SIZE {cw:,.0f} KiB per dispatch, so bytes = FETCH_SIZE x 1024 x {sf:.4f} and "
           f"WRITE_SIZE x 1024 x {sw:.4f}.",
           "", "| kernel | grid | FETCH_SIZE KiB | WRITE_SIZE KiB | traffic (B) |", "|---|---|---|---|---|"]
    head = None
    for key in sorted(f):
        if key not in w or "FETCH_SIZE" not in f[key] or "WRITE_SIZE" not in w[key]:
            continue
        fk = sum(f[key]["FETCH_SIZE"]) / len(f[key]["FETCH_SIZE"])
        wk = sum(w[key]["WRITE_SIZE"]) / len(w[key]["WRITE_SIZE"])
        b = fk * 1024 * sf + wk * 1024 * sw
        out.append(f"| `{key[0]}` | {key[1]:,} | {fk:,.0f} | {wk:,.0f} | {b:.4e} |")
        if headline and key == headline:
            head = b
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    out += [""]
    if head is not None:
        out += [f"traffic_bytes_per_launch: {head:.4e}", ""]
    out += [f"kernel_sources_sha: {bench.kernel_sources_sha()}"]
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="what", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    g = sub.add_parser("gaps")
    g.add_argument("dir")
    g.add_argument("--min-launches", type=int, default=50)
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    r = sub.add_parser("traffic")
    r.add_argument("fetch_dir")
    r.add_argument("write_dir")
    r.add_argument("--calibration", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                   help="FETCH_SIZE and WRITE_SIZE runs of `lbmcl_amd/bin/membw copy`")
    r.add_argument("--headline", nargs=2, metavar=("KERNEL", "GRID"),
                   help="kernel name (as in the table) and grid of bench.py's headline kernel")
    a = ap.parse_args(argv)
    if a.what == "trace":
        print(trace(a.dir))
    elif a.what == "gaps":
        print(gaps(a.dir, a.min_launches))
    elif a.what == "counters":
        print(counters(a.dirs)[0])
    else:
        head = (a.headline[0], int(a.headline[1])) if a.headline else None
        print(traffic(a.fetch_dir, a.write_dir, head, a.calibration))
    return 0


if __name__ == "__main__":
    sys.exit(main())
