"""lbmcl_amd.rocprof_summary on synthetic rocprofv3 CSV files (CPU only)."""
import csv
import os

from lbmcl_amd import rocprof_summary as R


def _write(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(header)
        w.writerows(rows)


def test_trace_and_traffic(tmp_path):
    k = "void lbm::lbm_collide_stream_vec<float, 0, 8, 2>(lbm::StepArgs<float>)"
    _write(str(tmp_path / "t" / "host" / "1" / "run_kernel_trace.csv"),
           ["Kind", "Kernel_Name", "Grid_Size", "Start_Timestamp", "End_Timestamp"],
           [["KERNEL_DISPATCH", k, 67108864, 1000, 3001000], ["KERNEL_DISPATCH", k, 67108864, 5000, 3006000],
            ["KERNEL_DISPATCH", "copy_kernel(...)", 268435456, 0, 1300000]])
    t = R.trace(str(tmp_path / "t"))
    assert "| `lbm::lbm_collide_stream_vec<float, 0, 8, 2>` | 67,108,864 | 2 | 3.0005 | 3.0000 | 3.0010 | 6.0 |" in t
    hdr = ["Dispatch_Id", "Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"]
    _write(str(tmp_path / "f" / "h" / "2" / "run_counter_collection.csv"), hdr,
           [[1, "copy_kernel", 268435456, "FETCH_SIZE", 2097152], [2, k, 67108864, "FETCH_SIZE", 4942000]])
    _write(str(tmp_path / "w" / "h" / "3" / "run_counter_collection.csv"), hdr,
           [[1, "copy_kernel", 268435456, "WRITE_SIZE", 4194304], [2, k, 67108864, "WRITE_SIZE", 9884000]])
    out = R.traffic(str(tmp_path / "f"), str(tmp_path / "w"),
                    ("lbm::lbm_collide_stream_vec<float, 0, 8, 2>", 67108864))
    # copy: 4 GiB = 2,097,152 KiB x 2 (fetch) and 4,194,304 KiB x 1 (write)
    assert "x 2.0000" in out and "x 1.0000" in out
    want = 4942000 * 1024 * 2 + 9884000 * 1024
    assert f"{want:.4e}" in out
    assert "kernel_sources_sha: " in out and f"traffic_bytes_per_launch: {want:.4e}" in out


def test_gaps(tmp_path):
    k, c = "void lbm::k<float>(args)", "copy_kernel(...)"
    # burst 1: three launches of 10 us with gaps 2 and 4 us; one other kernel; burst 2: two launches, gap 1 us
    rows = [[k, 256, 0, 10000], [k, 256, 12000, 22000], [k, 256, 26000, 36000], [c, 64, 40000, 41000],
            [k, 256, 50000, 60000], [k, 256, 61000, 71000]]
    _write(str(tmp_path / "t" / "run_kernel_trace.csv"), ["Kernel_Name", "Grid_Size", "Start_Timestamp",
                                                           "End_Timestamp"], rows[::-1])
    out = R.gaps(str(tmp_path / "t"), min_launches=2)
    lines = [l for l in out.splitlines() if l.startswith("| 1") or l.startswith("| 2")]
    assert lines[0] == "| 1 | `lbm::k<float>` | 256 | 3 | 10.00 | 3.00 | 4.00 | 4.00 | 0.8333 |"
    assert lines[1] == "| 2 | `lbm::k<float>` | 256 | 2 | 10.00 | 1.00 | 1.00 | 1.00 | 0.9524 |"
    assert len(lines) == 2  # the single copy dispatch is below min_launches


def test_names_and_separate_calibration(tmp_path):
    assert R._short("void (anonymous namespace)::copy_kernel(unsigned int const*)") == "copy_kernel"
    assert R._short("void at::k<4, at::F<(bool)1> >(int, F)") == "at::k<4, at::F<(bool)1> >"
    k = "void lbm::lbm_collide_stream_vec<float, 0, 8, 2, false>(lbm::StepArgs<float>)"
    cp = "void (anonymous namespace)::copy_kernel(unsigned int const*, unsigned int*, long)"
    hdr = ["Dispatch_Id", "Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"]
    _write(str(tmp_path / "f" / "r.csv".replace("r.csv", "run_counter_collection.csv")), hdr,
           [[1, k, 67108864, "FETCH_SIZE", 1000]])
    _write(str(tmp_path / "w" / "run_counter_collection.csv"), hdr, [[1, k, 67108864, "WRITE_SIZE", 3000]])
    # the 4 GiB copy calibrates; the other kernels of the calibration runs do not
    _write(str(tmp_path / "cf" / "run_counter_collection.csv"), hdr,
           [[1, "void at::native::fill<float>(int)", 1024, "FETCH_SIZE", 7], [2, cp, 268435456, "FETCH_SIZE", 2097152]])
    _write(str(tmp_path / "cw" / "run_counter_collection.csv"), hdr,
           [[1, "void at::native::fill<float>(int)", 1024, "WRITE_SIZE", 4194304],
            [2, cp, 268435456, "WRITE_SIZE", 4194304]])
    out = R.traffic(str(tmp_path / "f"), str(tmp_path / "w"), ("lbm::lbm_collide_stream_vec<float, 0, 8, 2, false>",
                                                               67108864), (str(tmp_path / "cf"), str(tmp_path / "cw")))
    want = 1000 * 1024 * 2 + 3000 * 1024
    assert "x 2.0000" in out and "x 1.0000" in out and f"traffic_bytes_per_launch: {want:.4e}" in out
