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
