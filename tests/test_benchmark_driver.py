"""The benchmark driver (benchmark.sh counterpart) on the CPU.

The aggregation is checked on hand-computed cases: the trimmed mean drops the
runs with the smallest and the largest total time (the first of equal values)
and removes those rows' other fields with them, whatever their own order; the
numbers print in awk's format.  A stand-in command line that prints a fixed
statistics line checks the sweep, the arguments and the files written.
"""
import os
import stat
import textwrap

import pytest

from lbmcl_amd import benchmark as B


def _line(total, kernels, mlups, kmlups, dim=32, lws="004,002,001", stride=8):
    return f"AMD Instinct MI355X;single;{dim};50;0;{lws};{stride};1;{total};{kernels};{mlups};{kmlups}\n"


def test_trimmed_mean_hand_computed():
    lines = [_line(10, 5, 100, 7), _line(40, 1, 400, 3), _line(20, 9, 200, 6), _line(30, 2, 300, 8),
             "beignet-opencl-icd: a line the awk script filters\n"]
    # min total 10 (row 1) and max total 40 (row 2) go; rows 3 and 4 remain
    assert B.aggregate(lines) == ["AMD Instinct MI355X;single;32;004,002,001;8;25;5.5;250;7"]


def test_ties_and_format():
    # equal totals: the first row is both min and max, subtracted twice
    lines = [_line(5, 1, 3, 1), _line(5, 2, 3, 2), _line(5, 4, 3, 4)]
    # kernels: (1 + 2 + 4) - 2 * 1 = 5
    assert B.aggregate(lines) == ["AMD Instinct MI355X;single;32;004,002,001;8;5;5;3;5"]
    assert B.awk_number(1234567.0) == "1234567"
    assert B.awk_number(1234567.5) == "1.23457e+06"
    assert B.awk_number(0.000123456789) == "0.000123457"
    with pytest.raises(ValueError):
        B.aggregate([_line(1, 1, 1, 1), _line(2, 2, 2, 2)])


def test_groups_by_key_fields():
    lines = [_line(t, 1, 1, 1, dim=d, stride=s) for d in (8, 16) for s in (1, 8) for t in (1, 2, 3)]
    # iterations and every are not key fields: this run joins the (8, stride 1) group
    lines.append(_line(9, 9, 9, 9, dim=8, stride=1).replace(";50;0;", ";60;5;"))
    out = B.aggregate(lines)
    assert len(out) == 4
    assert out[0] == "AMD Instinct MI355X;single;16;004,002,001;1;2;1;1;1"
    assert "AMD Instinct MI355X;single;8;004,002,001;1;2.5;1;1;1" in out


def test_lws_sweep():
    assert len(list(B.lws_sweep([1, 2, 4, 8, 16, 32, 64, 128], 8))) == 20
    s = set(B.lws_sweep([1, 2, 4, 8, 16, 32, 64, 128], 128))
    assert (128, 8, 1) in s and (128, 16, 1) not in s and (8, 16, 1) not in s


def test_driver_end_to_end(tmp_path):
    cli = tmp_path / "fake_lbmcl"
    cli.write_text(textwrap.dedent("""\
        #!/bin/sh
        echo "$@" >> "$(dirname "$0")/args.txt"
        echo "AMD Instinct MI355X;single;8;50;0;001,001,001;8;1;4;2;10;20" >&2
        """))
    cli.chmod(cli.stat().st_mode | stat.S_IEXEC)
    out = tmp_path / "b"
    rc = B.main(["--cli", str(cli), "--out", str(out), "--dims", "8", "--lws", "1,2", "--strides", "1,8",
                 "--repeats", "3"])
    assert rc == 0
    args = (tmp_path / "args.txt").read_text().splitlines()
    # lws (1,1,1), (2,1,1), (2,2,1), (2,2,2) x strides (1, 8) + stride 512, 3 repeats each
    assert len(args) == 4 * 3 * 3
    assert args[0] == "-P 0 -D 0 -d 8 -n 0.0089 -u 0.05 -i 50 -e 0 -w 1,1,1 -s 1 -o"
    assert sum("-s 512" in a for a in args) == 4 * 3
    assert (out / "benchmark.csv").read_text() == "AMD Instinct MI355X;single;8;001,001,001;8;4;2;10;20\n"
    assert len((out / "wall.csv").read_text().splitlines()) == 36


def test_driver_passes_arith(tmp_path):
    """--arith opencl reaches lbmcl in place of -o."""
    cli = tmp_path / "fake_lbmcl"
    cli.write_text(textwrap.dedent("""\
        #!/bin/sh
        echo "$@" >> "$(dirname "$0")/args.txt"
        echo "AMD Instinct MI355X;single;8;50;0;001,001,001;8;0;4;2;10;20" >&2
        """))
    cli.chmod(cli.stat().st_mode | stat.S_IEXEC)
    rc = B.main(["--cli", str(cli), "--out", str(tmp_path / "b"), "--dims", "8", "--lws", "1", "--strides", "8",
                 "--repeats", "3", "--no-soa", "--arith", "opencl"])
    assert rc == 0
    args = (tmp_path / "args.txt").read_text().splitlines()
    assert len(args) == 3 and all(a.endswith("-s 8 --arith opencl") for a in args)
