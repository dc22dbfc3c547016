"""CPU restatement of the exact-cut initial alpha from a closed triangulated
surface (geometricvofext_amd/csrc/harness_kernels.hip, k_surface_classify /
k_surface_cut): test infrastructure only, imported by tests/ as the checker
of the HIP kernels; never imported by the product package.

The reference computes this fraction with CGAL booleans
(applications/utilities/setVofField/setAlphaField.H:41-62); CGAL is not
here, so the restatement is of this library's own method, and it is pinned
by what the mathematics fixes: the exact box-cell overlap of an axis-aligned
box, the enclosed volume of a closed polyhedron summed over a mesh, and the
winding number of points inside and outside.

* ``clip_tet_volume(cone, T)``: volume of tet ``cone`` inside tet ``T``,
  clipped plane by plane (a tet cut by a plane leaves a tet or a prism of
  three tets).
* ``cell_fraction(mesh, C, Cf, c, tris)``: the signed sum over the surface
  triangles of the cone from the cell centre clipped to each of the cell's
  (centre, face centre, edge) tets, over the sum of those tets.
* ``winding(p, tris)``: generalized winding number (solid angles, van
  Oosterom & Strackee).
"""
from __future__ import annotations

import math

import numpy as np


def _vol(a, b, c, d):
    return abs(float(np.dot(b - a, np.cross(c - a, d - a)))) / 6.0


def clip_tet_volume(cone, T):
    cone = [np.asarray(p, float) for p in cone]
    T = [np.asarray(p, float) for p in T]
    planes = []
    for k in range(4):
        a, b, c = T[(k + 1) & 3], T[(k + 2) & 3], T[(k + 3) & 3]
        n = np.cross(b - a, c - a)
        if np.dot(n, T[k] - a) < 0.0:
            n = -n
        planes.append((n, float(np.dot(n, a))))
    pieces = [cone]
    for n, d in planes:
        nxt = []
        for t in pieces:
            sd = [float(np.dot(n, p)) - d for p in t]
            ins = [i for i in range(4) if sd[i] >= 0.0]
            out = [i for i in range(4) if sd[i] < 0.0]

            def cut(i, o):
                w = sd[i] / (sd[i] - sd[o])
                return t[i] + w * (t[o] - t[i])

            if not ins:
                continue
            if len(ins) == 4:
                nxt.append(t)
            elif len(ins) == 1:
                a = ins[0]
                nxt.append([t[a], cut(a, out[0]), cut(a, out[1]), cut(a, out[2])])
            else:
                if len(ins) == 3:
                    P = [t[i] for i in ins]
                    Q = [cut(i, out[0]) for i in ins]
                else:
                    a, b = ins
                    P = [t[a], cut(a, out[0]), cut(a, out[1])]
                    Q = [t[b], cut(b, out[0]), cut(b, out[1])]
                nxt += [[P[0], P[1], P[2], Q[0]], [P[1], P[2], Q[0], Q[1]], [P[2], Q[0], Q[1], Q[2]]]
        pieces = nxt
    return sum(_vol(*t) for t in pieces)


def cell_tets(mesh, C, Cf, c):
    """The (centre, face centre, edge) tets of cell c, faces in cell order."""
    faces = [f for f in range(mesh.n_faces) if mesh.owner[f] == c]
    faces += [f for f in range(mesh.n_internal_faces) if mesh.neighbour[f] == c]
    out = []
    for f in faces:
        pts = mesh.points[mesh.face(f)]
        for e in range(len(pts)):
            out.append([C[c], Cf[f], pts[e], pts[(e + 1) % len(pts)]])
    return out


def cell_fraction(mesh, C, Cf, c, tris):
    tets = cell_tets(mesh, C, Cf, c)
    cc = np.asarray(C[c], float)
    acc = 0.0
    for T in tets:
        for a, b, d in tris:
            s = float(np.dot(a - cc, np.cross(b - cc, d - cc)))
            if s == 0.0:
                continue
            v = clip_tet_volume([cc, a, b, d], T)
            acc += v if s > 0.0 else -v
    vc = sum(_vol(*T) for T in tets)
    return acc / vc if vc > 0.0 else 0.0


def winding(p, tris):
    p = np.asarray(p, float)
    w = 0.0
    for a, b, c in tris:
        ra, rb, rc = a - p, b - p, c - p
        la, lb, lc = np.linalg.norm(ra), np.linalg.norm(rb), np.linalg.norm(rc)
        num = float(np.dot(ra, np.cross(rb, rc)))
        den = la * lb * lc + float(np.dot(ra, rb)) * lc + float(np.dot(ra, rc)) * lb + float(np.dot(rb, rc)) * la
        w += 2.0 * math.atan2(num, den)
    return w / (4.0 * math.pi)
