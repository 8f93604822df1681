"""ParaView output of a solution: .vtu per load step and a .pvd collection.

Mirrors src/Interfaces/VTK.jl:
  write_vtk(sol, filename, time_index)  -- one step  (VTK.jl:215-236)
  write_vtk(sol, base_filename)         -- every step + `base_filename.pvd`
                                           (VTK.jl:250-269)
Cells: tetrahedra as VTK_TETRA (10), trusses as VTK_LINE (3) (VTK.jl:74-78).
Point data "Displacement" with components ux, uy, uz; cell data one scalar
array per tensor component, named as FIELD_NAMES of VTK.jl:139-146 and
picked through the INDEX_MAP of VTK.jl:113-116 (σxx ... τyx, ϵxx ... γyx).

Files are VTK XML UnstructuredGrid, written here directly (the reference goes
through WriteVTK.jl): binary (base64, UInt64 headers) by default, ASCII on
request.
"""
from __future__ import annotations

import base64
import os
from xml.sax.saxutils import quoteattr

import numpy as np

from .entities import Tetrahedra, Trusses

VTK_TETRA, VTK_LINE = 10, 3

# VTK.jl:113-116 -- substring of the label -> (row, column), 0-based here
INDEX_MAP = {"xx": (0, 0), "yy": (1, 1), "zz": (2, 2), "xy": (0, 1), "yz": (1, 2), "zx": (2, 0),
             "yx": (1, 0), "zy": (2, 1), "xz": (0, 2)}
# VTK.jl:139-146
FIELD_NAMES = {"u": ("Displacement", ["ux", "uy", "uz"]),
               "σ": ("Stress", ["σxx", "σyy", "σzz", "τyz", "τxz", "τxy", "τzy", "τzx", "τyx"]),
               "ϵ": ("Strain", ["ϵxx", "ϵyy", "ϵzz", "γyz", "γxz", "γxy", "γzy", "γzx", "γyx"])}
DEFAULT_FIELDS = ("u", "σ", "ϵ")


def _component_index(label):
    """The first INDEX_MAP key found in the label (VTK.jl:127-133 loops over the
    map and writes every match; the labels above match exactly one key)."""
    for key, ij in INDEX_MAP.items():
        if key in label:
            return ij
    raise ValueError(f"Unexpected label {label} didnt match INDEX_MAP")


def _cells(element_sets):
    conn, types = [], []
    for es in element_sets:
        if isinstance(es, Tetrahedra):
            conn.append(np.asarray(es.conn, dtype=np.int64).reshape(-1, 4))
            types.append(np.full(len(es), VTK_TETRA, dtype=np.uint8))
        elif isinstance(es, Trusses):
            conn.append(np.asarray(es.conn, dtype=np.int64).reshape(-1, 2))
            types.append(np.full(len(es), VTK_LINE, dtype=np.uint8))
        else:
            raise TypeError(f"no VTK cell for {type(es).__name__}")
    return conn, types


class _Writer:
    def __init__(self, ascii):
        self.ascii = ascii

    def array(self, name, data, ncomp=1, comp_names=None, vtype=None):
        data = np.ascontiguousarray(data)
        vtype = vtype or {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32",
                          np.dtype(np.int64): "Int64", np.dtype(np.uint8): "UInt8"}[data.dtype]
        attrs = f'type="{vtype}" Name={quoteattr(name)}'
        if ncomp > 1:
            attrs += f' NumberOfComponents="{ncomp}"'
            for k, cn in enumerate(comp_names or []):
                attrs += f" ComponentName{k}={quoteattr(cn)}"
        if self.ascii:
            body = " ".join(repr(float(v)) if data.dtype.kind == "f" else str(int(v))
                            for v in data.reshape(-1))
            return f'<DataArray {attrs} format="ascii">{body}</DataArray>'
        raw = data.tobytes()
        head = np.array([len(raw)], dtype=np.uint64).tobytes()
        body = base64.b64encode(head).decode() + base64.b64encode(raw).decode()
        return f'<DataArray {attrs} format="binary">{body}</DataArray>'


def _write_vtu(filename, coords, conn, types, point_data, cell_data, ascii=False):
    w = _Writer(ascii)
    ncell = sum(len(c) for c in conn)
    connectivity = np.concatenate([c.reshape(-1) for c in conn]) if conn else np.zeros(0, np.int64)
    offsets = np.cumsum(np.concatenate([np.full(len(c), c.shape[1]) for c in conn]) if conn
                        else np.zeros(0)).astype(np.int64)
    ctypes = np.concatenate(types) if types else np.zeros(0, np.uint8)
    out = ['<?xml version="1.0" encoding="utf-8"?>',
           '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
           'header_type="UInt64">',
           "<UnstructuredGrid>",
           f'<Piece NumberOfPoints="{coords.shape[0]}" NumberOfCells="{ncell}">',
           "<Points>", w.array("Points", coords.astype(np.float64), 3), "</Points>",
           "<Cells>", w.array("connectivity", connectivity), w.array("offsets", offsets),
           w.array("types", ctypes), "</Cells>"]
    if point_data:
        out.append("<PointData>")
        for name, (arr, comps) in point_data.items():
            out.append(w.array(name, arr, arr.shape[1] if arr.ndim > 1 else 1, comps))
        out.append("</PointData>")
    if cell_data:
        out.append("<CellData>")
        for name, arr in cell_data.items():
            if not np.all(np.isfinite(arr)):
                raise ValueError(f"non-finite values in cell data {name}")
            out.append(w.array(name, arr))
        out.append("</CellData>")
    out += ["</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
    with open(filename, "w", encoding="utf-8") as f:
        f.write("\n".join(out) + "\n")
    return filename


def _step_data(sol, k, fields):
    """Cells in the order of the structure's material pairs (the order of the
    per-group stresses the solution stores)."""
    s = sol.analysis.s
    sets = [es for _, es in s.materials]
    point_data, cell_data = {}, {}
    if "u" in fields:
        name, comps = FIELD_NAMES["u"]
        point_data[name] = (np.asarray(sol.U[k]).reshape(-1, 3), comps)
    for f, per_group in (("σ", sol.stress[k]), ("ϵ", sol.strain[k])):
        if f not in fields or per_group is None:
            continue
        T = np.concatenate([np.asarray(t).reshape(-1, 3, 3) for t in per_group])
        for label in FIELD_NAMES[f][1]:
            i, j = _component_index(label)
            cell_data[label] = T[:, i, j].copy()
    return s.mesh, sets, point_data, cell_data


def write_vtk(sol, filename, time_index=None, fields=DEFAULT_FIELDS, ascii=False):
    """write_vtk(sol, filename, time_index) writes load step `time_index`
    (1-based, as the reference) to `filename`; write_vtk(sol, base_filename)
    writes `base_filename_timestep_<k>.vtu` for every step and the collection
    `base_filename.pvd`, whose path it returns."""
    if time_index is not None:
        mesh, sets, pd, cd = _step_data(sol, time_index - 1, fields)
        conn, types = _cells(sets)
        return _write_vtu(filename, mesh.coords, conn, types, pd, cd, ascii)
    times = sol.analysis.load_factors
    entries = []
    for k, t in enumerate(times, start=1):
        if sol.U[k - 1] is None:
            continue
        fn = f"{filename}_timestep_{k}.vtu"
        write_vtk(sol, fn, k, fields, ascii)
        entries.append((float(t), os.path.basename(fn)))
    pvd = f"{filename}.pvd"
    with open(pvd, "w", encoding="utf-8") as f:
        f.write('<?xml version="1.0" encoding="utf-8"?>\n'
                '<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">\n'
                "<Collection>\n")
        for t, fn in entries:
            f.write(f'<DataSet timestep="{t!r}" part="0" file={quoteattr(fn)}/>\n')
        f.write("</Collection>\n</VTKFile>\n")
    return pvd


def read_vtu(filename):
    """Read back a file written by write_vtk (for tests and tooling):
    {"points", "connectivity", "offsets", "types", "point": {...}, "cell": {...}}."""
    import xml.etree.ElementTree as ET
    root = ET.parse(filename).getroot()
    header = np.uint64 if root.get("header_type") == "UInt64" else np.uint32
    dt = {"Float64": np.float64, "Float32": np.float32, "Int64": np.int64, "UInt8": np.uint8}

    def arr(el):
        t = dt[el.get("type")]
        if el.get("format") == "ascii":
            return np.array(el.text.split(), dtype=t)
        txt = el.text.strip()
        hlen = len(base64.b64encode(np.zeros(1, header).tobytes()))
        n = int(np.frombuffer(base64.b64decode(txt[:hlen]), header)[0])
        return np.frombuffer(base64.b64decode(txt[hlen:]), t)[: n // np.dtype(t).itemsize]

    piece = root.find("UnstructuredGrid/Piece")
    out = {"points": arr(piece.find("Points/DataArray")).reshape(-1, 3), "point": {}, "cell": {}}
    for el in piece.find("Cells"):
        out[el.get("Name")] = arr(el)
    for sec, key in (("PointData", "point"), ("CellData", "cell")):
        node = piece.find(sec)
        if node is not None:
            for el in node:
                a = arr(el)
                nc = int(el.get("NumberOfComponents", "1"))
                out[key][el.get("Name")] = a.reshape(-1, nc) if nc > 1 else a
    return out
