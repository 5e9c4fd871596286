"""Host code under sanitizers (CPU only).

The reference has no race or memory checking; its push streaming is race-free
because each slot (d, i) has exactly one writer, d - e_i (kernels.cl:423).
Here the native host code runs under GCC's sanitizers:

* AddressSanitizer + UndefinedBehaviorSanitizer (no recovery: the first
  report aborts): the CPU oracle's push and pull steps with per-slot write
  counters (each slot written exactly once per step where its source is a
  non-wall cell, never otherwise), push == pull, the 16-bit codecs; the dump
  writers (iostream reference and the threaded output path); the snapshot
  slot state machine.
* ThreadSanitizer: the threaded dump writers (several formatting threads per
  file) and the slot state machine under concurrent producers and consumers,
  including a failed wait that must not leave a producer blocked.

Each program exits non-zero on a failed check, and a sanitizer report fails
the test even if the exit status were zero.
"""
import os
import subprocess

import pytest

from tests.native import build_native

REPORT_MARKERS = ("ERROR: AddressSanitizer", "runtime error:", "WARNING: ThreadSanitizer", "ERROR: LeakSanitizer")


def _run(cmd, env_extra=None, timeout=600):
    env = dict(os.environ)
    env.update({"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1", "UBSAN_OPTIONS": "print_stacktrace=1",
                "TSAN_OPTIONS": "halt_on_error=1:second_deadlock_stack=1", "OMP_NUM_THREADS": "2"})
    env.update(env_extra or {})
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
    report = [m for m in REPORT_MARKERS if m in r.stderr]
    assert r.returncode == 0 and not report, f"{cmd[0]} rc={r.returncode} {report}\n{r.stderr[-4000:]}"
    return r


@pytest.mark.parametrize("sanitize", ["asan", None])
def test_oracle_selftest(sanitize):
    exe = build_native.build("oracle_selftest", sanitize)
    r = _run([exe])
    assert "all checks passed" in r.stdout


@pytest.mark.parametrize("sanitize", ["asan", "tsan"])
def test_slot_state_machine(sanitize):
    exe = build_native.build("slots_test", sanitize)
    r = _run([exe])
    assert "all checks passed" in r.stdout


@pytest.mark.parametrize("sanitize", ["asan", "tsan"])
def test_dump_writers(sanitize, tmp_path):
    exe = build_native.build("lbm_io_selftest", sanitize)
    for n, stride, prec, it, its, threads in ((16, 4, "f32", 7, 120, 5), (8, 512, "f64", 3, 3, 3)):
        d = tmp_path / f"{n}_{prec}"
        d.mkdir()
        _run([exe, str(d), str(n), str(stride), prec, str(it), str(its), str(threads)])
        # the threaded writers produce the same bytes as the iostream writers
        for name in os.listdir(d):
            if name.startswith("fast_"):
                assert (d / name).read_bytes() == (d / name[len("fast_"):]).read_bytes(), name
