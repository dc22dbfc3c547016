"""The CPU restatement of the exact-cut initial alpha from a closed
triangulated surface, pinned by closed forms: an axis-aligned box gives each
hexahedral cell the product of its axis overlaps; the fractions of any
closed polyhedron summed over a mesh that contains it give its enclosed
volume; the winding number is 1 inside and 0 outside.  A planted mistake
(the cone signs dropped) fails the volume sum."""
import numpy as np

from geometricvofext_amd import hex_box
from geometricvofext_amd import surface as S
from oracle import surface_oracle as SO


def _centres(mesh):
    """Cell and face centres of a mesh of planar faces (face: point average
    of a parallelogram; cell: point average of a parallelepiped)."""
    Cf = np.array([mesh.points[mesh.face(f)].mean(axis=0) for f in range(mesh.n_faces)])
    pts = [[] for _ in range(mesh.n_cells)]
    for f in range(mesh.n_faces):
        for c in ([mesh.owner[f]] + ([mesh.neighbour[f]] if f < mesh.n_internal_faces else [])):
            pts[c].extend(mesh.face(f).tolist())
    C = np.array([mesh.points[sorted(set(p))].mean(axis=0) for p in pts])
    return C, Cf


def test_box_overlaps():
    mesh = hex_box(3)
    C, Cf = _centres(mesh)
    lo, hi = np.array([0.1, 0.25, 0.4]), np.array([0.55, 0.9, 0.7])
    tris = S.box_surface(lo, hi)
    h = 1.0 / 3.0
    for c in range(mesh.n_cells):
        clo = C[c] - 0.5 * h
        ov = np.prod(np.clip(np.minimum(hi, clo + h) - np.maximum(lo, clo), 0.0, None)) / h ** 3
        assert abs(SO.cell_fraction(mesh, C, Cf, c, tris) - ov) < 1e-13, c


def test_enclosed_volume_and_winding():
    mesh = hex_box(2)
    C, Cf = _centres(mesh)
    for tris in (S.icosphere((0.47, 0.52, 0.5), 0.3, 1), S.torus_surface((0.5, 0.5, 0.5), 0.25, 0.1, 10, 6)):
        fr = np.array([SO.cell_fraction(mesh, C, Cf, c, tris) for c in range(mesh.n_cells)])
        assert abs(np.sum(fr) / 8.0 - S.enclosed_volume(tris)) < 1e-13
        # planted mistake: every cone counted positive
        bad = 0.0
        for c in range(mesh.n_cells):
            for T in SO.cell_tets(mesh, C, Cf, c):
                for a, b, d in tris:
                    bad += SO.clip_tet_volume([C[c], a, b, d], T)
        assert abs(bad - S.enclosed_volume(tris)) > 1e-3
    ico = S.icosphere((0.5, 0.5, 0.5), 0.3, 1)
    assert abs(SO.winding((0.5, 0.5, 0.5), ico) - 1.0) < 1e-12
    assert abs(SO.winding((0.95, 0.1, 0.5), ico)) < 1e-12


def test_clip_tet_volume_cases():
    """Tet inside, outside, and cut so that one, two and three vertices stay."""
    T = [np.array(p, float) for p in ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1))]
    assert abs(SO.clip_tet_volume(T, T) - 1.0 / 6.0) < 1e-15
    far = [p + np.array([5.0, 0, 0]) for p in T]
    assert SO.clip_tet_volume(far, T) == 0.0
    # the unit cube's corner tets against T: a scaled copy inside T
    half = [0.5 * p for p in T]
    assert abs(SO.clip_tet_volume(half, T) - 1.0 / 48.0) < 1e-15
    # T cut by a translated copy: the overlap is the tet scaled by (1 - s)
    for s in (0.2, 0.5, 0.8):
        sh = [p + np.array([s, 0, 0]) for p in T]
        assert abs(SO.clip_tet_volume(sh, T) - (1 - s) ** 3 / 6.0) < 1e-14
