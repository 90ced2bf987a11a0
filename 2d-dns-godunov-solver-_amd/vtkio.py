"""Node fields on the mesh as a VTK XML unstructured grid (.vtu).

The reference writes its results with WriteVTK.jl: ``saveResults2VTK``,
``saveResults3VTK`` and ``saveResults4VTK`` (utilsIO.jl:2-53) open
``vtk_grid(filename, xNodes, yNodes, VTKCells)`` and attach node (point) data
such as "density", "Ux", "Uy", "pressure"; the cells are VTK triangles and
quads from ``mesh_connectivity`` (utilsIO.jl:188-225).  This module writes the
same grid and point data in the VTK XML format (``<UnstructuredGrid>``, points
with z = 0, connectivity/offsets/types, one ``DataArray`` per field), either
as ASCII or as base64 binary with a UInt64 size header, and reads such files
back.
"""
from __future__ import annotations

import base64
import xml.etree.ElementTree as ET

import numpy as np

from .mesh import mesh2d_Int32

VTK_TRIANGLE = 5
VTK_QUAD = 9
_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32", np.dtype(np.int64): "Int64",
          np.dtype(np.int32): "Int32", np.dtype(np.uint8): "UInt8"}


def _cells(m: mesh2d_Int32):
    nv = np.asarray(m.mesh_connectivity[:, 2])
    ids = np.asarray(m.mesh_connectivity[:, 3:7]) - 1  # 0-based node ids
    conn = np.concatenate([ids[i, :nv[i]] for i in range(m.nCells)]).astype(np.int64)
    offsets = np.cumsum(nv).astype(np.int64)
    types = np.where(nv == 4, VTK_QUAD, VTK_TRIANGLE).astype(np.uint8)
    return conn, offsets, types


def _array(parent, name, a, ncomp, binary):
    a = np.ascontiguousarray(a)
    el = ET.SubElement(parent, "DataArray", type=_TYPES[a.dtype], Name=name, NumberOfComponents=str(ncomp),
                       format="binary" if binary else "ascii")
    if binary:
        raw = a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
        el.text = base64.b64encode(np.uint64(len(raw)).tobytes() + raw).decode()
    else:
        el.text = " ".join(repr(float(x)) if a.dtype.kind == "f" else str(int(x)) for x in a.reshape(-1))


def save_results_vtk(filename: str, m: mesh2d_Int32, fields: dict, binary: bool = True) -> str:
    """Writes ``fields`` (name -> node array of length nNodes) on the mesh;
    returns the file name (".vtu" appended when missing, as WriteVTK does)."""
    if not filename.endswith(".vtu"):
        filename += ".vtu"
    conn, offsets, types = _cells(m)
    root = ET.Element("VTKFile", type="UnstructuredGrid", version="1.0", byte_order="LittleEndian",
                      header_type="UInt64")
    grid = ET.SubElement(root, "UnstructuredGrid")
    piece = ET.SubElement(grid, "Piece", NumberOfPoints=str(m.nNodes), NumberOfCells=str(m.nCells))
    pts = np.zeros((m.nNodes, 3))
    pts[:, 0], pts[:, 1] = np.asarray(m.xNodes).reshape(-1), np.asarray(m.yNodes).reshape(-1)
    _array(ET.SubElement(piece, "Points"), "Points", pts, 3, binary)
    cells = ET.SubElement(piece, "Cells")
    _array(cells, "connectivity", conn, 1, binary)
    _array(cells, "offsets", offsets, 1, binary)
    _array(cells, "types", types, 1, binary)
    pd = ET.SubElement(piece, "PointData")
    for name, v in fields.items():
        v = np.asarray(v, dtype=np.float64).reshape(-1)
        if v.size != m.nNodes:
            raise ValueError(f"field {name!r} has {v.size} values for {m.nNodes} nodes")
        _array(pd, name, v, 1, binary)
    ET.ElementTree(root).write(filename, xml_declaration=True, encoding="utf-8")
    return filename


def saveResults2VTK(filename: str, m: mesh2d_Int32, variable, variableName: str = "density") -> str:  # noqa: N802
    """utilsIO.jl:2-30: one node field."""
    return save_results_vtk(filename, m, {variableName: variable})


def saveResults4VTK(filename: str, m: mesh2d_Int32, density, Ux, Uy, pressure) -> str:  # noqa: N802,N803
    """utilsIO.jl:43-53: the node primitives."""
    return save_results_vtk(filename, m, {"density": density, "Ux": Ux, "Uy": Uy, "pressure": pressure})


def read_vtu(filename: str) -> dict:
    """Points, cells and point data of a file written by ``save_results_vtk``."""
    piece = ET.parse(filename).getroot().find("UnstructuredGrid/Piece")

    def arr(el):
        dt = np.dtype({v: k for k, v in _TYPES.items()}[el.get("type")]).newbyteorder("<")
        if el.get("format") == "binary":
            raw = base64.b64decode(el.text)
            n = int(np.frombuffer(raw[:8], np.uint64)[0])
            a = np.frombuffer(raw[8:8 + n], dt)
        else:
            a = np.array(el.text.split(), dtype=dt)
        nc = int(el.get("NumberOfComponents", "1"))
        return a.reshape(-1, nc) if nc > 1 else a

    out = {"points": arr(piece.find("Points/DataArray"))}
    for el in piece.find("Cells"):
        out[el.get("Name")] = arr(el)
    out["point_data"] = {el.get("Name"): arr(el) for el in piece.find("PointData")}
    return out
