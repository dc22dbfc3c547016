"""Writers for the PLIC surface (the polygons of ``SolveVofEqu.plic_surface``).

* ``write_vtp``: one VTK XML PolyData file, the format the reference writes
  through ``plicSurface::write(".../plicSurface.vtp")``
  (applications/utilities/setVofField/exportPlicSurface.H:27) and its surface
  function objects; points in Float64, one polygon per interface cell and the
  mesh cell of each polygon as cell data ``meshCells``.
* ``write_series``: the ``<name>.vtp.series`` JSON index of such files over
  time that ``writeVtkSeries`` produces
  (applications/utilities/writeVtkSeries/writeVtkSeries.C:60-146).
"""
from __future__ import annotations

import os

import numpy as np


def _ascii(a, fmt):
    return " ".join(fmt % v for v in np.asarray(a).reshape(-1))


def write_vtp(path, points, offsets, cells=None, cell_fields=None):
    """Write polygons to ``path``: polygon f is ``points[offsets[f]:offsets[f+1]]``.
    ``cells`` (n_faces) goes out as cell data ``meshCells``; ``cell_fields``
    maps further names to per-polygon arrays."""
    points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    offsets = np.asarray(offsets, dtype=np.int64)
    if offsets.ndim != 1 or offsets.size < 1 or offsets[0] != 0 or offsets[-1] != points.shape[0]:
        raise ValueError("offsets must start at 0 and end at the number of points")
    if np.any(np.diff(offsets) < 0):
        raise ValueError("offsets must not decrease")
    n_faces = offsets.size - 1
    fields = {}
    if cells is not None:
        fields["meshCells"] = np.asarray(cells, dtype=np.int64)
    for k, v in (cell_fields or {}).items():
        fields[k] = np.asarray(v)
    for k, v in fields.items():
        if v.shape[0] != n_faces:
            raise ValueError(f"cell field {k}: {v.shape[0]} values for {n_faces} polygons")
    lines = [
        '<?xml version="1.0"?>',
        '<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
        "<PolyData>",
        f'<Piece NumberOfPoints="{points.shape[0]}" NumberOfVerts="0" NumberOfLines="0" '
        f'NumberOfStrips="0" NumberOfPolys="{n_faces}">',
        "<Points>",
        '<DataArray type="Float64" NumberOfComponents="3" format="ascii">',
        _ascii(points, "%.17g"),
        "</DataArray>",
        "</Points>",
        "<Polys>",
        '<DataArray type="Int64" Name="connectivity" format="ascii">',
        _ascii(np.arange(points.shape[0]), "%d"),
        "</DataArray>",
        '<DataArray type="Int64" Name="offsets" format="ascii">',
        _ascii(offsets[1:], "%d"),
        "</DataArray>",
        "</Polys>",
    ]
    if fields:
        lines.append("<CellData>")
        for k, v in fields.items():
            if np.issubdtype(v.dtype, np.integer):
                lines += [f'<DataArray type="Int64" Name="{k}" format="ascii">', _ascii(v, "%d"), "</DataArray>"]
            else:
                nc = 1 if v.ndim == 1 else v.shape[1]
                lines += [f'<DataArray type="Float64" Name="{k}" NumberOfComponents="{nc}" format="ascii">',
                          _ascii(v.astype(np.float64), "%.17g"), "</DataArray>"]
        lines.append("</CellData>")
    lines += ["</Piece>", "</PolyData>", "</VTKFile>", ""]
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write("\n".join(lines))


def read_vtp(path):
    """Read back what ``write_vtp`` wrote: dict(points, offsets, fields)."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    piece = root.find("PolyData/Piece")

    def arr(node, dtype):
        text = (node.text or "").split()
        return np.array([dtype(t) for t in text])

    pts = arr(piece.find("Points/DataArray"), float).reshape(-1, 3)
    polys = {d.get("Name"): d for d in piece.findall("Polys/DataArray")}
    offs = np.concatenate([[0], arr(polys["offsets"], int)]).astype(np.int64)
    conn = arr(polys["connectivity"], int).astype(np.int64)
    fields = {}
    for d in piece.findall("CellData/DataArray"):
        fields[d.get("Name")] = arr(d, int if d.get("type").startswith("Int") else float)
    return dict(points=pts[conn] if conn.size else pts, offsets=offs, fields=fields)


def write_series(path, entries):
    """``<name>.vtp.series`` index: entries are (file name, time) pairs in time
    order, written in the layout of writeVtkSeries."""
    out = ['{\n  "file-series-version" : "1.0",\n  "files" : [\n']
    for name, t in entries:
        out.append(f'    {{ "name" : "{name}", "time" : {t} }},\n')
    out.append("  ]\n}\n")
    with open(path, "w") as f:
        f.write("".join(out))
