"""Build the CPU-only native test programs of tests/native (test infrastructure).

Each program is built with g++ into tests/native/_bin (git-ignored), optionally
with sanitizers:

  lbm_io_selftest   the dump formats: iostream reference writers
                    (lbm_io_reference.hpp) and the product's threaded writers
                    (lbmcl_amd/csrc/lbm_output.hpp)
  oracle_selftest   the CPU oracle (oracle/lbm_oracle.cpp): per-slot write
                    counts of the push and pull steps, push == pull, codecs
  slots_test        the snapshot slot state machine (lbmcl_amd/csrc/lbm_slots.hpp)

sanitize: None, "asan" (-fsanitize=address,undefined, no recovery) or "tsan"
(-fsanitize=thread).
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "_bin")
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(ROOT, "lbmcl_amd", "csrc")

SANITIZE = {
    None: [],
    "asan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"],
    "tsan": ["-fsanitize=thread"],
}

PROGRAMS = {
    "lbm_io_selftest": (["lbm_io_selftest.cpp"], ["-pthread"]),
    "oracle_selftest": (["oracle_selftest.cpp"], ["-fopenmp", "-ffp-contract=off", "-fno-fast-math"]),
    "slots_test": (["slots_test.cpp"], ["-pthread"]),
}


def build(name: str, sanitize: str | None = None) -> str:
    """Compile program `name` (a PROGRAMS key); returns the executable's path."""
    sources, extra = PROGRAMS[name]
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(BIN, name + (f"_{sanitize}" if sanitize else ""))
    cxx = shutil.which("g++") or "g++"
    opt = ["-O1", "-g"] if sanitize else ["-O2"]
    cmd = [cxx, "-std=c++17", *opt, "-Wall", "-Wextra", "-I", INCLUDE, "-I", CSRC, "-I", HERE,
           *[os.path.join(HERE, s) for s in sources], *extra, *SANITIZE[sanitize], "-o", out]
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    import sys
    san = sys.argv[1] if len(sys.argv) > 1 else None
    for prog in PROGRAMS:
        print(build(prog, san))
