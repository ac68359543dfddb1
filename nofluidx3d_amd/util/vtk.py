"""Legacy-VTK unstructured grids without pyvista.

The reference reads meshes with ``pv.read`` and writes frames with ``mesh.save``
(``TetraCell.py:14``, ``util/IO.py:39``) and uses only a small part of pyvista's
``UnstructuredGrid``: ``points`` (get/set), ``cells_dict[10]`` (tetra connectivity),
``cell_data.set_array``, ``n_cells``, ``translate``, ``scale`` and ``save``.  This module provides
that surface for legacy ``.vtk`` files:

- reading: ASCII and BINARY (big-endian), file versions 2.0-4.2 (``CELLS n size`` with
  per-cell counts) and 5.1 (``CELLS`` + ``OFFSETS``/``CONNECTIVITY`` arrays); POINT_DATA and
  CELL_DATA ``SCALARS`` and ``FIELD`` arrays;
- writing: version 4.2, BINARY by default (as ``pyvista.DataSet.save``) or ASCII.
"""

from __future__ import annotations

import numpy as np

VTK_TRIANGLE = 5
VTK_TETRA = 10
_NPTS = {1: 1, 3: 2, 5: 3, 9: 4, 10: 4, 12: 8, 13: 6, 14: 5}
_VTK_TYPES = {
    "float": ">f4",
    "double": ">f8",
    "int": ">i4",
    "unsigned_int": ">u4",
    "long": ">i8",
    "unsigned_long": ">u8",
    "vtktypeint64": ">i8",
    "vtktypeint32": ">i4",
    "short": ">i2",
    "unsigned_short": ">u2",
    "char": ">i1",
    "unsigned_char": ">u1",
}
_NAME_OF = {np.dtype(np.float32): "float", np.dtype(np.float64): "double", np.dtype(np.int32): "int",
            np.dtype(np.int64): "vtktypeint64", np.dtype(np.uint32): "unsigned_int", np.dtype(np.uint8): "unsigned_char"}


class DataArrays(dict):
    """``mesh.cell_data`` / ``mesh.point_data``: name -> array, with pyvista's ``set_array``."""

    def set_array(self, array, name):
        self[name] = np.asarray(array)


class UnstructuredGrid:
    def __init__(self, points, cells: dict | None = None, cell_types=None, connectivity=None):
        self._points = np.array(points, dtype=np.float64).reshape(-1, 3)
        if cells is not None:
            types, conn = [], []
            for ctype, arr in cells.items():
                arr = np.asarray(arr, dtype=np.int64).reshape(-1, _NPTS[ctype])
                types.append(np.full(len(arr), ctype, dtype=np.uint8))
                conn.extend(list(arr))
            self.cell_types = np.concatenate(types) if types else np.zeros(0, np.uint8)
            self._conn = conn
        else:
            self.cell_types = np.asarray(cell_types, dtype=np.uint8)
            self._conn = list(connectivity)
        self.cell_data = DataArrays()
        self.point_data = DataArrays()

    # -- pyvista-like surface ------------------------------------------------------------------
    @property
    def points(self):
        return self._points

    @points.setter
    def points(self, value):
        self._points = np.array(value, dtype=np.float64).reshape(-1, 3)

    @property
    def n_points(self):
        return len(self._points)

    @property
    def n_cells(self):
        return len(self.cell_types)

    @property
    def cells_dict(self):
        out = {}
        for ctype in np.unique(self.cell_types):
            idx = np.nonzero(self.cell_types == ctype)[0]
            out[int(ctype)] = np.array([self._conn[i] for i in idx], dtype=np.int64).reshape(len(idx), -1)
        return out

    def translate(self, xyz, inplace=True):
        self._points = self._points + np.asarray(xyz, dtype=np.float64)
        return self

    def scale(self, xyz, inplace=True):
        self._points = self._points * np.asarray(xyz, dtype=np.float64)
        return self

    def copy(self):
        g = UnstructuredGrid(self._points.copy(), cell_types=self.cell_types.copy(), connectivity=list(self._conn))
        g.cell_data.update({k: v.copy() for k, v in self.cell_data.items()})
        g.point_data.update({k: v.copy() for k, v in self.point_data.items()})
        return g

    def save(self, path, binary=True):
        write(self, path, binary=binary)


def from_tetra(points, tets) -> UnstructuredGrid:
    return UnstructuredGrid(points, {VTK_TETRA: np.asarray(tets)})


# ------------------------------------------------------------------------------------------------
# writer
# ------------------------------------------------------------------------------------------------


def _emit(f, arr, binary, dtype_be, per_line=9):
    arr = np.asarray(arr)
    if binary:
        f.write(np.ascontiguousarray(arr.astype(dtype_be)).tobytes())
        f.write(b"\n")
    else:
        flat = arr.ravel()
        fmt = "%.17g" if arr.dtype.kind == "f" else "%d"
        for i in range(0, len(flat), per_line):
            f.write((" ".join(fmt % v for v in flat[i : i + per_line]) + "\n").encode())


def write(grid: UnstructuredGrid, path, binary=True):
    with open(path, "wb") as f:
        f.write(b"# vtk DataFile Version 4.2\nnofluidx3d_amd\n")
        f.write(b"BINARY\n" if binary else b"ASCII\n")
        f.write(b"DATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {grid.n_points} double\n".encode())
        _emit(f, grid.points, binary, ">f8", 3)
        n = grid.n_cells
        sizes = np.array([len(c) for c in grid._conn], dtype=np.int64)
        flat = np.empty(int(sizes.sum() + n), dtype=np.int64)
        pos = 0
        for c in grid._conn:
            flat[pos] = len(c)
            flat[pos + 1 : pos + 1 + len(c)] = c
            pos += 1 + len(c)
        f.write(f"CELLS {n} {len(flat)}\n".encode())
        _emit(f, flat, binary, ">i4", 9)
        f.write(f"CELL_TYPES {n}\n".encode())
        _emit(f, grid.cell_types.astype(np.int64), binary, ">i4", 9)
        for section, arrays, count in (("CELL_DATA", grid.cell_data, n), ("POINT_DATA", grid.point_data,
                                                                         grid.n_points)):
            if not arrays:
                continue
            f.write(f"{section} {count}\n".encode())
            f.write(f"FIELD FieldData {len(arrays)}\n".encode())
            for name, arr in arrays.items():
                arr = np.asarray(arr)
                ncomp = 1 if arr.ndim == 1 else arr.shape[1]
                dt = np.dtype(np.float64) if arr.dtype.kind == "f" else arr.dtype
                tname = _NAME_OF.get(np.dtype(dt), "double")
                f.write(f"{name} {ncomp} {len(arr)} {tname}\n".encode())
                _emit(f, arr.astype(dt), binary, np.dtype(_VTK_TYPES[tname]), 9)


# ------------------------------------------------------------------------------------------------
# reader
# ------------------------------------------------------------------------------------------------


class _Reader:
    def __init__(self, data: bytes):
        self.d = data
        self.p = 0

    def line(self) -> str:
        while True:
            e = self.d.find(b"\n", self.p)
            if e < 0:
                e = len(self.d)
            s = self.d[self.p : e].decode("latin-1").strip()
            self.p = e + 1
            if s or self.p >= len(self.d):
                return s

    def eof(self):
        return self.p >= len(self.d)

    def values(self, count: int, vtk_type: str, binary: bool):
        dt = np.dtype(_VTK_TYPES[vtk_type])
        if binary:
            nbytes = count * dt.itemsize
            arr = np.frombuffer(self.d[self.p : self.p + nbytes], dtype=dt).copy()
            self.p += nbytes
            if self.p < len(self.d) and self.d[self.p : self.p + 1] == b"\n":
                self.p += 1
            return arr.astype(dt.newbyteorder("="))
        out = []
        while len(out) < count:
            out.extend(self.line().split())
        return np.array(out[:count], dtype=dt.newbyteorder("=") if dt.kind != "f" else np.float64)


def read(path) -> UnstructuredGrid:
    with open(path, "rb") as fh:
        r = _Reader(fh.read())
    header = r.line()
    if not header.startswith("# vtk DataFile"):
        raise ValueError(f"{path}: not a legacy VTK file")
    r.line()  # title
    fmt = r.line().upper()
    binary = fmt == "BINARY"
    ds = r.line().split()
    if len(ds) < 2 or ds[1].upper() != "UNSTRUCTURED_GRID":
        raise ValueError(f"{path}: only UNSTRUCTURED_GRID is supported, got {' '.join(ds)}")
    points = None
    conn, types = None, None
    grid_cell, grid_point = {}, {}
    section, count = None, 0
    while not r.eof():
        s = r.line()
        if not s:
            continue
        tok = s.split()
        key = tok[0].upper()
        if key == "POINTS":
            n = int(tok[1])
            points = r.values(3 * n, tok[2].lower(), binary).astype(np.float64).reshape(n, 3)
        elif key == "CELLS":
            n, size = int(tok[1]), int(tok[2])
            nxt = r.p
            peek = r.line()
            if peek.upper().startswith("OFFSETS"):  # version 5.1
                offsets = r.values(n, peek.split()[1].lower(), binary).astype(np.int64)
                c = r.line().split()
                flat = r.values(size, c[1].lower(), binary).astype(np.int64)
                conn = [flat[offsets[i] : offsets[i + 1]] for i in range(n - 1)]
            else:
                r.p = nxt
                flat = r.values(size, "int", binary).astype(np.int64)
                conn, pos = [], 0
                for _ in range(n):
                    k = int(flat[pos])
                    conn.append(flat[pos + 1 : pos + 1 + k])
                    pos += 1 + k
        elif key == "CELL_TYPES":
            n = int(tok[1])
            types = r.values(n, "int", binary).astype(np.uint8)
        elif key in ("CELL_DATA", "POINT_DATA"):
            section, count = key, int(tok[1])
        elif key == "SCALARS":
            name, vt = tok[1], tok[2].lower()
            ncomp = int(tok[3]) if len(tok) > 3 else 1
            nxt = r.p
            if not r.line().upper().startswith("LOOKUP_TABLE"):
                r.p = nxt
            arr = r.values(count * ncomp, vt, binary)
            (grid_cell if section == "CELL_DATA" else grid_point)[name] = arr.reshape(count, ncomp).squeeze()
        elif key == "VECTORS" or key == "NORMALS":
            arr = r.values(count * 3, tok[2].lower(), binary)
            (grid_cell if section == "CELL_DATA" else grid_point)[tok[1]] = arr.reshape(count, 3)
        elif key == "FIELD":
            for _ in range(int(tok[2])):
                a = r.line().split()
                name, ncomp, ntup, vt = a[0], int(a[1]), int(a[2]), a[3].lower()
                arr = r.values(ncomp * ntup, vt, binary)
                (grid_cell if section == "CELL_DATA" else grid_point)[name] = arr.reshape(ntup, ncomp).squeeze()
        elif key == "METADATA":
            while not r.eof() and r.line():
                pass
    if points is None or conn is None or types is None:
        raise ValueError(f"{path}: missing POINTS/CELLS/CELL_TYPES")
    g = UnstructuredGrid(points, cell_types=types, connectivity=conn)
    g.cell_data.update(grid_cell)
    g.point_data.update(grid_point)
    return g
