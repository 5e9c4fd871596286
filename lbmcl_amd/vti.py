"""numpy-only reader for the VTK ImageData (.vti) files of the lid-driven cavity.

Reads the two encodings the reference workflow meets:

* the files LBMCL writes (lbmcl.hpp:261-334): ``format="ascii"``,
  ``header_type="UInt64"``, arrays ``rho`` (1 component) and ``v`` (3
  components) of type Float32/Float64;
* the Sailfish fixtures the reference verifies against
  (target_results/*/ldc.0.*.vti): ``format="appended"``, base64 encoding,
  ``vtkZLibDataCompressor``, ``header_type="UInt32"``.

The reference reads both with pyvista (verify.py:44-45), which this package
does not depend on; this module replaces it for verify.py's purpose.
"""
from __future__ import annotations

import base64
import re
import zlib

import numpy as np

_DTYPES = {
    "Float32": np.float32, "Float64": np.float64, "Int32": np.int32, "Int64": np.int64,
    "UInt8": np.uint8, "UInt32": np.uint32, "UInt64": np.uint64,
}

_ATTR = re.compile(r'(\w+)="([^"]*)"')


def _attrs(tag: str) -> dict:
    return dict(_ATTR.findall(tag))


class VTIFile:
    """Parsed ImageData: ``extent`` (6 ints), ``arrays`` name -> ndarray.

    Arrays with several components have shape (npoints, ncomp); scalars have
    shape (npoints,).  Points are ordered x fastest, then y, then z.
    """

    def __init__(self, extent, arrays, header_type, dtype_names):
        self.extent = extent
        self.arrays = arrays
        self.header_type = header_type
        self.dtype_names = dtype_names

    @property
    def dimensions(self):
        e = self.extent
        return (e[1] - e[0] + 1, e[3] - e[2] + 1, e[5] - e[4] + 1)

    def __getitem__(self, name):
        return self.arrays[name]


def _b64len(nbytes: int) -> int:
    return ((nbytes + 2) // 3) * 4


def _read_appended(payload: str, offset: int, header_dtype, compressed: bool, dtype) -> np.ndarray:
    s = payload[offset:]
    hsize = np.dtype(header_dtype).itemsize
    if compressed:
        first = base64.b64decode(s[:_b64len(hsize)])
        nblocks = int(np.frombuffer(first[:hsize], dtype=header_dtype)[0])
        hbytes = (3 + nblocks) * hsize
        hlen = _b64len(hbytes)
        header = np.frombuffer(base64.b64decode(s[:hlen]), dtype=header_dtype)
        csizes = [int(c) for c in header[3:3 + nblocks]]
        dlen = _b64len(sum(csizes))
        data = base64.b64decode(s[hlen:hlen + dlen])
        out, pos = [], 0
        for c in csizes:
            out.append(zlib.decompress(data[pos:pos + c]))
            pos += c
        return np.frombuffer(b"".join(out), dtype=dtype)
    first = base64.b64decode(s[:_b64len(hsize)])
    nbytes = int(np.frombuffer(first[:hsize], dtype=header_dtype)[0])
    raw = base64.b64decode(s[:_b64len(hsize + nbytes)])
    return np.frombuffer(raw[hsize:hsize + nbytes], dtype=dtype)


def read_vti(path: str) -> VTIFile:
    with open(path, "r") as fh:
        text = fh.read()
    head = _attrs(re.search(r"<VTKFile[^>]*>", text).group(0))
    header_type = head.get("header_type", "UInt32")
    header_dtype = np.dtype(_DTYPES[header_type]).newbyteorder("<")
    compressed = "compressor" in head
    image = _attrs(re.search(r"<ImageData[^>]*>", text).group(0))
    extent = [int(v) for v in image["WholeExtent"].split()]
    payload = ""
    m = re.search(r"<AppendedData[^>]*>\s*_", text)
    if m:
        end = text.index("</AppendedData>", m.end())
        payload = text[m.end():end].strip()
    arrays, dtype_names = {}, {}
    for m in re.finditer(r"<DataArray([^>]*?)(/?)>", text):
        a = _attrs(m.group(1))
        name = a["Name"]
        dtype = np.dtype(_DTYPES[a["type"]]).newbyteorder("<")
        ncomp = int(a.get("NumberOfComponents", "1"))
        fmt = a.get("format", "ascii")
        if fmt == "ascii":
            end = text.index("</DataArray>", m.end())
            values = np.array(text[m.end():end].split(), dtype=np.float64).astype(dtype)
        elif fmt == "appended":
            values = _read_appended(payload, int(a["offset"]), header_dtype, compressed, dtype)
        else:  # binary inline
            end = text.index("</DataArray>", m.end())
            values = _read_appended(text[m.end():end].strip(), 0, header_dtype, compressed, dtype)
        values = values.astype(dtype.newbyteorder("="))
        if ncomp > 1:
            values = values.reshape(-1, ncomp)
        arrays[name] = values
        dtype_names[name] = a["type"]
    return VTIFile(extent, arrays, header_type, dtype_names)
