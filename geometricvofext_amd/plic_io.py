"""Writers for the PLIC surface (the polygons of ``SolveVofEqu.plic_surface``).

* ``write_vtp``: one VTK XML PolyData file, the format the reference writes
  through ``plicSurface::write(".../plicSurface.vtp")``
  (applications/utilities/setVofField/exportPlicSurface.H:27) and its surface
  function objects; points in Float64, one polygon per interface cell and the
  mesh cell of each polygon as cell data ``meshCells``.
* ``write_series``: the ``<name>.vtp.series`` JSON index of such files over
  time that ``writeVtkSeries`` produces
  (applications/utilities/writeVtkSeries/writeVtkSeries.C:60-146).
* ``write_plic_fields``: the fields the reference writes into each time
  directory when ``writePlicFields`` is on (interfaceN, interfaceD,
  interfaceC, interfaceS as volume fields and the face flatness zeta as a
  surface field, src/SimPLIC/reconstruction/reconstruction.C:448-471,
  520-584), as OpenFOAM ASCII field files.
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


# name -> (OpenFOAM class, dimensions, value type) of the fields the
# reference registers with writePlicFields (reconstruction.C:520-584)
PLIC_FIELDS = {
    "interfaceN": ("volVectorField", "[0 -1 0 0 0 0 0]", "vector"),
    "interfaceD": ("volScalarField", "[0 1 0 0 0 0 0]", "scalar"),
    "interfaceC": ("volVectorField", "[0 1 0 0 0 0 0]", "vector"),
    "interfaceS": ("volVectorField", "[0 2 0 0 0 0 0]", "vector"),
    "zeta": ("surfaceScalarField", "[0 0 0 0 0 0 0]", "scalar"),
}


def _foam_list(values, kind):
    v = np.asarray(values, dtype=np.float64)
    if kind == "vector":
        v = v.reshape(-1, 3)
        body = "\n".join("(%.17g %.17g %.17g)" % tuple(r) for r in v)
        return f"nonuniform List<vector>\n{v.shape[0]}\n(\n{body}\n)\n"
    body = "\n".join("%.17g" % x for x in v.reshape(-1))
    return f"nonuniform List<scalar>\n{v.size}\n(\n{body}\n)\n"


def write_foam_field(path, name, cls, dims, kind, internal, patches):
    """One OpenFOAM ASCII field file; ``patches`` is a list of (patch name,
    values or None), None giving a calculated patch of uniform zero."""
    zero = "(0 0 0)" if kind == "vector" else "0"
    lines = ["FoamFile", "{", "    version     2.0;", "    format      ascii;", f"    class       {cls};",
             f'    location    "{os.path.basename(os.path.dirname(os.path.abspath(path)))}";',
             f"    object      {name};", "}", "", f"dimensions      {dims};", "",
             "internalField   " + _foam_list(internal, kind) + ";", "", "boundaryField", "{"]
    for pname, vals in patches:
        lines += [f"    {pname}", "    {", "        type            calculated;"]
        if vals is None or np.asarray(vals).size == 0:
            lines.append(f"        value           uniform {zero};")
        else:
            lines.append("        value           " + _foam_list(vals, kind) + ";")
        lines.append("    }")
    lines += ["}", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


def write_plic_fields(case_dir, time_name, mesh, fields):
    """Write ``fields`` (``SolveVofEqu.plic_fields()``) into
    ``<case_dir>/<time_name>/`` as the reference does with writePlicFields:
    the four interface fields over the cells with calculated, zero boundary
    patches, and zeta over the internal faces with its boundary face values
    per patch.  Returns the paths written."""
    d = os.path.join(case_dir, str(time_name))
    os.makedirs(d, exist_ok=True)
    out = []
    nif = mesh.n_internal_faces
    for name, (cls, dims, kind) in PLIC_FIELDS.items():
        if name not in fields:
            continue
        v = np.asarray(fields[name], dtype=np.float64)
        if name == "zeta":
            internal = v[:nif]
            patches = [(pn, v[int(st):int(st) + int(sz)])
                       for pn, st, sz in zip(mesh.patch_names, mesh.patch_start, mesh.patch_size)]
        else:
            internal = v
            patches = [(pn, None) for pn in mesh.patch_names]
        path = os.path.join(d, name)
        write_foam_field(path, name, cls, dims, kind, internal, patches)
        out.append(path)
    return out


def read_foam_field(path):
    """Internal field of an ASCII field file written by ``write_foam_field``
    (n or n x 3 values)."""
    text = open(path).read()
    i = text.index("internalField")
    head, rest = text[i:].split("(", 1)
    n = int(head.split()[-1])
    if "List<vector>" in head:
        rows = rest.split(")", n)[:n]
        return np.array([[float(t) for t in r.replace("(", " ").split()] for r in rows])
    return np.array([float(t) for t in rest.split(")", 1)[0].split()[:n]])
