"""Copy the reference's own fixture files into tests/golden (data, not code).

The fixtures are the VTI files of the reference's `make test8` and
`make test32` targets (Makefile:73-86): target_results/8/ldc.0.00..10.vti and
target_results/32/ldc.0.000..500.vti (every 20 iterations), written by
Sailfish in FP32.  They are copied byte for byte; tests read them from here
because the reference tree is not available where the GPU tests run.

usage: python tests/golden/copy_fixtures.py [reference_root]
"""
import filecmp
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(argv):
    ref = argv[1] if len(argv) > 1 else "/root/reference"
    for sub in ("8", "32"):
        src = os.path.join(ref, "target_results", sub)
        dst = os.path.join(HERE, "target_results", sub)
        os.makedirs(dst, exist_ok=True)
        names = sorted(n for n in os.listdir(src) if n.endswith(".vti"))
        for n in names:
            s, d = os.path.join(src, n), os.path.join(dst, n)
            if not (os.path.exists(d) and filecmp.cmp(s, d, shallow=False)):
                shutil.copyfile(s, d)
        print(f"{sub}: {len(names)} files")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
