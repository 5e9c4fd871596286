"""bench.py's command-line contract.

CPU: more than one GPU is refused with a message and exit status 2 before
anything touches a GPU; the traffic figure is taken from the committed
FETCH_SIZE / WRITE_SIZE profile only while the kernel sources hash to the
value recorded there.  GPU: a short run prints one JSON line with the fields
the contract names (metric, value, unit, n_gpus, steps, warmup, ms_per_step,
higher_is_better, scaling, vs_baseline, dtype, data, config with workload,
roofline with bound/achieved/peak/unit/frac/traffic, cpu_baseline).
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_traffic_profile_is_bound_to_the_kernel_sources(tmp_path, monkeypatch):
    # the committed profile matches the current sources
    t, name, sha = bench.traffic_from_profiles(512, 4096, "f32", "native")
    assert sha == bench.kernel_sources_sha() and t is not None and 1.0 < t / (510 ** 3 * 152) < 1.02
    # a profile recorded on other sources gives no traffic figure
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / name).write_text("traffic_bytes_per_launch: 2.0e10\nkernel_sources_sha: 0123456789abcdef\n")
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    monkeypatch.setattr(bench, "kernel_sources_sha", lambda: "fedcba9876543210")
    assert bench.traffic_from_profiles(512, 4096, "f32", "native")[0] is None


@pytest.mark.gpu
def test_short_run_prints_the_contract_line():
    cmd = [sys.executable, "-u", os.path.join(ROOT, "bench.py"), "--dim", "128", "--steps", "5", "--warmup", "2",
           "--variant-steps", "0", "--cpu-steps", "0", "--adjoint-steps", "0", "--no-calibration"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 5 and d["warmup"] == 2 and d["unit"] == "MLUPS" and d["value"] > 0
    assert d["higher_is_better"] is True and d["dtype"] == "f32" and "workload" in d["config"]
    rf = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in rf, k
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3 and 0 < rf["frac"] < 1
