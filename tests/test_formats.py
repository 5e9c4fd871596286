"""The dump formats of the lbmcl host (CPU only).

lbmcl_amd/csrc/lbm_output.hpp writes map.dump, f_<it>.dump and lbmcl.<it>.vti
for the lbmcl host.  tests/native/lbm_io_selftest writes them from synthetic
arrays (NaNs included), with those writers and with an iostream statement of
the reference writers (tests/native/lbm_io_reference.hpp), and the test
formats the same arrays with an independent Python statement of the reference
writers (lbmcl.hpp:159-334): all three agree byte for byte.  The numpy VTI
reader must read the result back exactly.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from lbmcl_amd.vti import read_vti
from tests.native import build_native


def _c_float(fmt, x):
    """printf-style formatting as the reference's ostream does it: C prints the
    sign of a NaN ("-nan"), Python does not."""
    x = float(x)
    if math.isnan(x):
        s = "-nan" if math.copysign(1.0, x) < 0 else "nan"
        return s if fmt == "%.16e" else f"{s:>8}"
    return fmt % x


def _digits(v):  # lbmcl.hpp:13
    return int(math.log10(v)) + 1 if v > 0 else 1


def _map_dump(m, n):  # lbmcl.hpp:159-203
    out = ["# FLUID       1\n# MOVING      2\n# BOUNDARY    3\n# WALL        4\n# CORNER      5\n\n"]
    for z in range(n):
        for y in range(n):
            row = []
            for x in range(n):
                t = int(m[x + y * n + z * n * n])
                v = 0
                if t == 0x1:
                    v = 1
                if t & 0x2:
                    v = 2
                if t & 0x3F0:
                    v = 3
                if t == 0x8:
                    v = 4
                if t == 0x4:
                    v = 5
                row.append(f"{v} ")
            out.append("".join(row) + "\n")
        out.append("\n")
    out.append("\n")
    return "".join(out)


def _f_dump(f, n, stride):  # lbmcl.hpp:206-258
    dd = _digits(n)
    out = []
    for z in range(n):
        for y in range(n):
            out.append(" " * (dd * 3 + 5) + "".join(f"{q:>8} " for q in range(19)) + "\n")
            for x in range(n):
                c = x + y * n + z * n * n
                vals = "".join(_c_float("%8.6f", f[((c // stride) * 19 + q) * stride + (c & (stride - 1))]) + " "
                               for q in range(19))
                out.append(f"{'(':>{dd}}{x},{y},{z}) " + vals + "\n")
            out.append("\n")
        out.append("\n")
    out.append("\n")
    return "".join(out)


def _vti(rho, u, n, dtype_name):  # lbmcl.hpp:261-334
    e = n - 3
    n3 = n ** 3
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian" header_type="UInt64">\n',
           f'  <ImageData WholeExtent="0 {e} 0 {e} 0 {e}" Origin="0 0 0" Spacing="1 1 1">\n',
           f'    <Piece Extent="0 {e} 0 {e} 0 {e}">\n',
           '      <PointData Scalars="rho">\n',
           f'        <DataArray type="{dtype_name}" Name="rho" NumberOfComponents="1" format="ascii">\n']
    for z in range(1, n - 1):
        for y in range(1, n - 1):
            out.append("".join(_c_float("%.16e", rho[x + y * n + z * n * n]) + " " for x in range(1, n - 1)) + "\n")
    out.append("        </DataArray>\n")
    out.append(f'        <DataArray type="{dtype_name}" Name="v" NumberOfComponents="3" format="ascii">\n')
    for z in range(1, n - 1):
        for y in range(1, n - 1):
            row = []
            for x in range(1, n - 1):
                c = x + y * n + z * n * n
                row.append("".join(_c_float("%.16e", u[k * n3 + c]) + " " for k in range(3)))
            out.append("".join(row) + "\n")
    out.append("        </DataArray>\n      </PointData>\n    </Piece>\n  </ImageData>\n</VTKFile>\n")
    return "".join(out)


@pytest.fixture(scope="module")
def selftest_bin():
    return build_native.build("lbm_io_selftest")


@pytest.mark.parametrize("n,stride,prec,it,its,threads", [(8, 4, "f32", 7, 120, 3), (8, 512, "f64", 3, 3, 1),
                                                          (16, 1, "f32", 0, 0, 5), (4, 64, "f64", 12, 1000, 2),
                                                          (32, 8, "f64", 5, 10, 7)])
def test_dump_formats_byte_exact(selftest_bin, tmp_path, n, stride, prec, it, its, threads):
    """Both writers (iostream and the threaded to_chars output path) against an
    independent statement of the reference writers, byte for byte."""
    subprocess.run([selftest_bin, str(tmp_path), str(n), str(stride), prec, str(it), str(its), str(threads)],
                   check=True)
    dt = np.float32 if prec == "f32" else np.float64
    rho = np.fromfile(tmp_path / "rho.bin", dtype=dt)
    u = np.fromfile(tmp_path / "u.bin", dtype=dt)
    f = np.fromfile(tmp_path / "f.bin", dtype=dt)
    m = np.fromfile(tmp_path / "map.bin", dtype=np.int32)
    w = _digits(its)
    vti_path = tmp_path / f"lbmcl.{it:0{w}d}.vti"
    f_path = tmp_path / f"f_{it:0{w}d}.dump"
    want_map, want_f = _map_dump(m, n), _f_dump(f, n, stride)
    want_vti = _vti(rho, u, n, "Float32" if prec == "f32" else "Float64")
    if n >= 8:
        assert "-nan" in want_vti and "-inf" in want_vti and "-0.0000000000000000e+00" in want_vti
        assert "-nan" in want_f and "     inf" in want_f
    for prefix in ("", "fast_"):
        assert (tmp_path / f"{prefix}map.dump").read_text() == want_map, prefix
        assert (tmp_path / f"{prefix}f_{it:0{w}d}.dump").read_text() == want_f, prefix
        assert (tmp_path / f"{prefix}lbmcl.{it:0{w}d}.vti").read_text() == want_vti, prefix
    # and the VTI reads back exactly (values print with 17 significant digits)
    v = read_vti(str(vti_path))
    assert v.header_type == "UInt64" and v.dimensions == (n - 2,) * 3
    win = rho.reshape(n, n, n)[1:n - 1, 1:n - 1, 1:n - 1].reshape(-1)
    assert np.array_equal(v["rho"], win, equal_nan=True)
    uw = np.stack([u.reshape(3, n, n, n)[k][1:n - 1, 1:n - 1, 1:n - 1].reshape(-1) for k in range(3)], axis=1)
    assert np.array_equal(v["v"], uw, equal_nan=True)


def test_verify_reads_both_encodings(golden_dir, tmp_path):
    """verify.py counterpart: a fixture compared with itself has zero error."""
    from lbmcl_amd.verify import verify
    tdir = os.path.join(golden_dir, "target_results", "8")
    for i in range(0, 11):
        os.symlink(os.path.join(tdir, f"ldc.0.{i:02d}.vti"), tmp_path / f"lbmcl.{i:02d}.vti")
    rows = verify(10, 1, tdir, str(tmp_path), out=open(os.devnull, "w"))
    assert len(rows) == 11 and all(r[1:] == (0.0, 0.0, 0.0, 0.0) for r in rows)


def test_verify_reader_cross_check(golden_dir, tmp_path, capsys):
    """--pyvista: without pyvista the cross-check is reported as skipped; with a
    second reader the arrays must agree, and a reader that disagrees fails."""
    from lbmcl_amd import verify as V
    tdir = os.path.join(golden_dir, "target_results", "8")
    for i in (0, 5, 10):
        os.symlink(os.path.join(tdir, f"ldc.0.{i:02d}.vti"), tmp_path / f"lbmcl.{i:02d}.vti")
    if V.pyvista_reader() is None:
        assert V.main(["-i", "10", "-e", "5", "-t", tdir, "-p", str(tmp_path), "--pyvista"]) == 0
        assert "cross-check skipped" in capsys.readouterr().err
    same = lambda path: read_vti(path)  # noqa: E731
    assert len(V.verify(10, 5, tdir, str(tmp_path), out=open(os.devnull, "w"), other_reader=same)) == 3

    def off(path):
        d = read_vti(path)
        return {"rho": d["rho"] + 1e-7, "v": d["v"]}

    with pytest.raises(ValueError):
        V.verify(10, 5, tdir, str(tmp_path), out=open(os.devnull, "w"), other_reader=off)
