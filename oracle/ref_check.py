"""Build-time check of oracle/_ref code objects (run in the build container only).

A compiled reference kernel is rejected if its disassembly holds a scalar-memory
store, atomic, or write-back/discard of the scalar data cache: GPU code here
writes through vector stores only.
"""
from __future__ import annotations

import os
import re
import subprocess

_SCALAR_WRITE = re.compile(r"^\s*s_(store|buffer_store|scratch_store|atomic|buffer_atomic|dcache_wb|dcache_discard)")


def check_no_scalar_writes(path, objdump, arch):
    r = subprocess.run([objdump, "-d", f"--mcpu={arch}", path], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{objdump} -d {path} failed: {r.stderr}")
    bad = [ln.strip() for ln in r.stdout.splitlines() if _SCALAR_WRITE.match(ln)]
    if bad:
        os.remove(path)
        raise RuntimeError(f"{path}: scalar-memory writes in the compiled reference kernels: {bad[:4]}")
