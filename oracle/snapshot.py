"""Test infrastructure: the reference's snapshot file format restated in numpy.

The layout follows output_data (src/input_output.f90:498-728) and the
reference's Python reader and writer (tools/file_io.py:6-80 read_field,
:141-212 save_field): a big-endian stream of

* int32 length + FP64 values of x, y, z, xm, ym, zm;
* int32[3] dims + FP64 values (Fortran order) of U, V, W, P;
* (body format only) the same for Hu_interior, Hv_interior, Hw_interior,
  then FP64 t, dt, dpdx, nu; int32 length + FP64 xb, yb, zb; int32 nxb,
  nzb; int32 length + FP64 fb, ub, sb, normals, tangents_1, tangents_2.

The fields-only format ("original", tools/file_io.py:8-19) stops after P.
Only tests/ use this module, to check the library's reader and writer.
"""
from __future__ import annotations

import numpy as np

I4, F8 = np.dtype(">i4"), np.dtype(">f8")
MESH = ("x", "y", "z", "xm", "ym", "zm")
FIELDS = ("U", "V", "W", "P")
MASKS = ("Hu_interior", "Hv_interior", "Hw_interior")
SCALARS = ("t", "dt", "dpdx", "nu")
BODY_VECS = ("fb", "ub", "sb", "normals", "tangents_1", "tangents_2")


def _vec(f):
    n = int(np.fromfile(f, dtype=I4, count=1)[0])
    a = np.fromfile(f, dtype=F8, count=n).astype(np.float64)
    if a.size != n:
        raise ValueError("truncated snapshot")
    return a


def _field(f):
    dims = np.fromfile(f, dtype=I4, count=3)
    n = int(np.prod(dims))
    a = np.fromfile(f, dtype=F8, count=n)
    if a.size != n:
        raise ValueError("truncated snapshot")
    return np.asfortranarray(a.astype(np.float64).reshape(tuple(int(d) for d in dims), order="F"))


def read(path, fields_only=False):
    """Snapshot -> dict of numpy arrays and scalars (tools/file_io.py:6-80)."""
    d = {}
    with open(path, "rb") as f:
        for k in MESH:
            d[k] = _vec(f)
        for k in FIELDS:
            d[k] = _field(f)
        if fields_only:
            return d
        for k in MASKS:
            d[k] = _field(f)
        for k in SCALARS:
            d[k] = float(np.fromfile(f, dtype=F8, count=1)[0])
        for k in ("xb", "yb", "zb"):
            d[k] = _vec(f)
        d["nxb"] = int(np.fromfile(f, dtype=I4, count=1)[0])
        d["nzb"] = int(np.fromfile(f, dtype=I4, count=1)[0])
        for k in BODY_VECS:
            d[k] = _vec(f)
    return d


def write(path, d, fields_only=False):
    """dict -> snapshot (tools/file_io.py:141-212)."""
    with open(path, "wb") as f:
        for k in MESH:
            a = np.asarray(d[k], dtype=np.float64)
            f.write(np.array([a.size], dtype=I4).tobytes())
            f.write(a.astype(F8).tobytes())
        keys = FIELDS if fields_only else FIELDS + MASKS
        for k in keys:
            a = np.asarray(d[k], dtype=np.float64)
            f.write(np.array(a.shape, dtype=I4).tobytes())
            f.write(a.astype(F8).flatten(order="F").tobytes())
        if fields_only:
            return
        f.write(np.array([d[k] for k in SCALARS], dtype=F8).tobytes())
        for k in ("xb", "yb", "zb"):
            a = np.asarray(d[k], dtype=np.float64)
            f.write(np.array([a.size], dtype=I4).tobytes())
            f.write(a.astype(F8).tobytes())
        f.write(np.array([d["nxb"], d["nzb"]], dtype=I4).tobytes())
        for k in BODY_VECS:
            a = np.asarray(d[k], dtype=np.float64)
            f.write(np.array([a.size], dtype=I4).tobytes())
            f.write(a.astype(F8).tobytes())


def expected_size(nx, ny, nz, nb, fields_only=False):
    """Byte length of a snapshot, counted item by item from the writer
    (src/input_output.f90:520-724)."""
    nxg, nyg, nzg = nx + 1, ny + 1, nz + 1
    mesh = sum(4 + 8 * n for n in (nx, ny, nz, nx - 1, ny - 1, nz - 1))
    shapes = [(nx, nyg, nzg), (nxg, ny, nzg), (nxg, nyg, nz), (nxg, nyg, nzg)]
    if not fields_only:
        shapes += [(nx, nyg, nzg), (nxg, ny, nzg), (nxg, nyg, nz)]
    fields = sum(12 + 8 * a * b * c for a, b, c in shapes)
    if fields_only:
        return mesh + fields
    body = 4 * 8 + 3 * (4 + 8 * nb) + 2 * 4
    body += sum(4 + 8 * n for n in (4 * nb, 3 * nb, nb, 3 * nb, 3 * nb, 3 * nb))
    return mesh + fields + body
