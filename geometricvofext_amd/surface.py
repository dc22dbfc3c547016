"""Closed triangulated surfaces for the exact-cut initial alpha
(``SolveVofEqu.init_alpha_surface``, the setVofField utility's shapes,
``applications/utilities/setVofField/setAlphaField.H:41-62``).

A surface is an (n, 3, 3) float64 array: n triangles of 3 points, oriented
counter-clockwise seen from outside (outward normals).
"""
from __future__ import annotations

import struct

import numpy as np


def read_stl(path: str) -> np.ndarray:
    """Triangles of an ASCII or binary STL file (the facet normals are not
    used; the vertex order gives the orientation)."""
    with open(path, "rb") as fh:
        data = fh.read()
    if len(data) >= 84:
        n = struct.unpack_from("<I", data, 80)[0]
        if 84 + 50 * n == len(data):
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]),
                                count=n, offset=84)
            return rec["v"].astype(np.float64)
    pts = []
    for line in data.decode("ascii", errors="replace").splitlines():
        w = line.split()
        if len(w) == 4 and w[0] == "vertex":
            pts.append([float(w[1]), float(w[2]), float(w[3])])
    arr = np.asarray(pts, dtype=np.float64)
    if arr.size == 0 or len(arr) % 3:
        raise ValueError(f"{path}: no triangles read")
    return arr.reshape(-1, 3, 3)


def write_stl(path: str, tris: np.ndarray, name: str = "surface") -> None:
    """ASCII STL of the triangles (facet normals from the vertex order)."""
    tris = np.asarray(tris, dtype=np.float64)
    with open(path, "w") as fh:
        fh.write(f"solid {name}\n")
        for t in tris:
            n = np.cross(t[1] - t[0], t[2] - t[0])
            ln = np.linalg.norm(n)
            n = n / ln if ln > 0 else n
            fh.write(f"  facet normal {n[0]:.17g} {n[1]:.17g} {n[2]:.17g}\n    outer loop\n")
            for p in t:
                fh.write(f"      vertex {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
            fh.write("    endloop\n  endfacet\n")
        fh.write(f"endsolid {name}\n")


def enclosed_volume(tris: np.ndarray) -> float:
    """Volume enclosed by a closed outward-oriented surface: the signed sum
    of the tets (origin, triangle)."""
    t = np.asarray(tris, dtype=np.float64)
    return float(np.sum(np.einsum("ij,ij->i", t[:, 0], np.cross(t[:, 1], t[:, 2])))) / 6.0


def box_surface(lo, hi) -> np.ndarray:
    """The 12 outward triangles of the axis-aligned box [lo, hi]."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    c = np.array([[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1], lo[2] if i & 4 == 0 else hi[2]]
                  for i in range(8)])
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tris = []
    for a, b, cc, d in quads:
        tris.append([c[a], c[b], c[cc]])
        tris.append([c[a], c[cc], c[d]])
    return np.asarray(tris)


def icosphere(centre, radius: float, subdivisions: int = 2) -> np.ndarray:
    """Triangulated sphere: an icosahedron refined ``subdivisions`` times
    (20 * 4^s triangles), vertices on the sphere."""
    p = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p), (0, -1, -p), (0, 1, -p),
         (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    verts = [np.asarray(x, float) / np.linalg.norm(x) for x in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.asarray(verts) * float(radius) + np.asarray(centre, float)
    tris = V[np.asarray(faces)]
    # outward orientation
    if enclosed_volume(tris) < 0.0:
        tris = tris[:, ::-1]
    return tris


def torus_surface(centre, R: float, r: float, n_major: int = 24, n_minor: int = 12) -> np.ndarray:
    """Triangulated torus about the z axis through ``centre`` (a non-convex
    closed surface)."""
    u = 2.0 * np.pi * np.arange(n_major) / n_major
    v = 2.0 * np.pi * np.arange(n_minor) / n_minor
    P = np.zeros((n_major, n_minor, 3))
    P[..., 0] = (R + r * np.cos(v)[None, :]) * np.cos(u)[:, None]
    P[..., 1] = (R + r * np.cos(v)[None, :]) * np.sin(u)[:, None]
    P[..., 2] = r * np.sin(v)[None, :]
    P += np.asarray(centre, float)
    tris = []
    for i in range(n_major):
        for j in range(n_minor):
            a, b = P[i, j], P[(i + 1) % n_major, j]
            c, d = P[(i + 1) % n_major, (j + 1) % n_minor], P[i, (j + 1) % n_minor]
            tris.append([a, b, c])
            tris.append([a, c, d])
    tris = np.asarray(tris)
    if enclosed_volume(tris) < 0.0:
        tris = tris[:, ::-1]
    return tris
