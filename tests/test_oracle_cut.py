"""Oracle pinning for the cut geometry (cutFace.C, cutCell.C) against closed
forms, complementarity and round trips.  CPU only."""
import itertools
import math

import numpy as np
import pytest

from geometricvofext_amd import mesh as M

SQUARE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], dtype=float)


def polygon_below(m, beta):
    """Area fraction of the unit square where m.u < beta (m_i > 0): 2-D
    inclusion-exclusion of the corner simplices."""
    m1, m2 = m
    s = 0.0
    for S in itertools.product((0, 1), repeat=2):
        sh = beta - S[0] * m1 - S[1] * m2
        s += (-1) ** sum(S) * max(sh, 0.0) ** 2
    return min(max(s / (2 * m1 * m2), 0.0), 1.0)


def cube_below(m, beta):
    """Volume fraction of the unit cube where m.u < beta (m_i > 0)
    (Scardovelli & Zaleski 2000, inclusion-exclusion form)."""
    m1, m2, m3 = m
    if beta > 0.5 * (m1 + m2 + m3):
        # evaluate the smaller side: the sums cancel less
        return 1.0 - cube_below(m, m1 + m2 + m3 - beta)
    s = 0.0
    for S in itertools.product((0, 1), repeat=3):
        sh = beta - S[0] * m1 - S[1] * m2 - S[2] * m3
        s += (-1) ** sum(S) * max(sh, 0.0) ** 3
    return min(max(s / (6 * m1 * m2 * m3), 0.0), 1.0)


def cube_fraction(n, d, h):
    """Fraction of [0,h]^3 with n.x + d < 0, by reflection to positive m."""
    n = np.asarray(n, float)
    beta = -d / h - sum(ni for ni in n if ni < 0)
    return cube_below(np.abs(n), beta)


def square_fraction(n2, d):
    n2 = np.asarray(n2, float)
    beta = -d - sum(ni for ni in n2 if ni < 0)
    return polygon_below(np.abs(n2), beta)


def test_sub_face_closed_forms(oracle_lib):
    rng = np.random.default_rng(0)
    for _ in range(200):
        a, b = rng.uniform(-1, 1, 2)
        if min(abs(a), abs(b)) < 1e-3:
            continue
        n = np.array([a, b, rng.uniform(-1, 1)])
        d = rng.uniform(-1.5, 1.5)
        st, area, centre, nsub, ncut = oracle_lib.calc_sub_face(SQUARE, n, d)
        exact = square_fraction(n[:2], d)
        assert abs(np.linalg.norm(area) - exact) < 1e-13
        if 1e-12 < exact < 1 - 1e-12:
            assert st == 0 and ncut == 2
        # complement: the two sides tile the face
        st2, area2, *_ = oracle_lib.calc_sub_face(SQUARE, -n, -d)
        assert abs(np.linalg.norm(area) + np.linalg.norm(area2) - 1.0) < 1e-13


def test_sub_face_triangle_corner_and_centroid(oracle_lib):
    # plane x + y = 0.5 cuts the corner triangle (0,0),(0.5,0),(0,0.5)
    n = np.array([1.0, 1.0, 0.0]) / math.sqrt(2)
    d = -0.5 / math.sqrt(2)
    st, area, centre, nsub, ncut = oracle_lib.calc_sub_face(SQUARE, n, d)
    assert st == 0 and nsub == 3 and ncut == 2
    assert abs(area[2] - 0.125) < 1e-15
    assert np.allclose(centre, [1 / 6, 1 / 6, 0.0], atol=1e-15)
    # pentagon: complement of the opposite corner
    st, area, centre, nsub, ncut = oracle_lib.calc_sub_face(SQUARE, n, -1.5 / math.sqrt(2))
    assert nsub == 5 and abs(area[2] - (1 - 0.125)) < 1e-14


def test_sub_face_vertex_lift(oracle_lib):
    """|s| < 1e-14 lifts the vertex by sign(s)*1e-14 with sign(0)=+1
    (cutFace.C:161-165): a vertex on the plane counts as not submerged."""
    n = np.array([1.0, 0.0, 0.0])
    st, area, *_ = oracle_lib.calc_sub_face(SQUARE, n, 0.0)      # plane on the x=0 edge
    assert st == 1 and np.all(area == 0)
    # plane on the x=1 edge: those vertices are lifted out, the face is cut
    # along them and the sub-face is the whole face
    st, area, centre, nsub, ncut = oracle_lib.calc_sub_face(SQUARE, n, -1.0)
    assert st == 0 and ncut == 2 and abs(area[2] - 1.0) < 1e-13
    # a tiny positive distance is pushed further out: still cut along x=1
    st, area, *_ = oracle_lib.calc_sub_face(SQUARE, n, -1.0 + 5e-15)
    assert st == 0 and abs(area[2] - 1.0) < 1e-13
    # a tiny negative distance is pushed further in: the face is submerged
    st, area, *_ = oracle_lib.calc_sub_face(SQUARE, n, -1.0 - 5e-15)
    assert st == -1 and abs(area[2] - 1.0) < 1e-15


@pytest.fixture(scope="module")
def cube_mesh(oracle_lib):
    return oracle_lib.OracleMesh(M.hex_box(4))


def test_sub_cell_scardovelli_zaleski(cube_mesh):
    rng = np.random.default_rng(1)
    h = 0.25
    cell = 21   # interior cell (1,1,1): origin at (0.25,0.25,0.25)
    o = np.array([0.25, 0.25, 0.25])
    for _ in range(300):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        if np.min(np.abs(n)) < 1e-3:
            continue
        dloc = rng.uniform(-1.8 * h, 1.8 * h)      # plane n.(x-o) + dloc = 0
        d = dloc - n @ o
        st, out = cube_mesh.calc_sub_cell(cell, n, d)
        exact = cube_fraction(n, dloc, h)
        # the reference's 1e-14 thresholds (vertex lift, zero-area interface)
        # make sub-1e-12 slivers round to empty/full
        assert abs(out[0] - exact) < 1e-12, (n, dloc, out[0], exact)
        st2, out2 = cube_mesh.calc_sub_cell(cell, -n, -d)
        assert abs(out[0] + out2[0] - 1.0) < 1e-12


def test_sub_cell_tet_closed_form(oracle_lib):
    """Plane cutting one vertex off a tetrahedron: volume = V * prod of edge fractions."""
    om = oracle_lib.OracleMesh(M.tet_box(2, perturb=0.0))
    rng = np.random.default_rng(2)
    mesh = om.mesh
    for cell in rng.choice(mesh.n_cells, 10, replace=False):
        faces = [f for f in range(mesh.n_faces) if mesh.owner[f] == cell or
                 (f < mesh.n_internal_faces and mesh.neighbour[f] == cell)]
        pts = np.unique(np.concatenate([mesh.face(f) for f in faces]))
        P = mesh.points[pts]
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        s = P @ n
        order = np.argsort(s)
        # plane just above the lowest vertex: isolates it
        lo, nxt = s[order[0]], s[order[1]]
        level = lo + 0.37 * (nxt - lo)
        st, out = om.calc_sub_cell(int(cell), n, -level)
        frac = np.prod([(level - lo) / (s[j] - lo) for j in order[1:]])
        assert abs(out[0] - frac) < 1e-12


@pytest.mark.parametrize("kind", ["hex", "tet"])
def test_find_signed_distance_round_trip(oracle_lib, kind):
    """alpha* = VOF(n, d*) -> findSignedDistance recovers d* on planar-faced
    cells, where the volume is one cubic between vertex levels."""
    mesh = M.hex_box(4) if kind == "hex" else M.tet_box(3)
    om = oracle_lib.OracleMesh(mesh)
    rng = np.random.default_rng(3)
    h = 0.25 if kind == "hex" else 1 / 3
    n_checked = 0
    for _ in range(200):
        cell = int(rng.integers(mesh.n_cells))
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        c = om.C[cell]
        dstar = -(n @ c) + rng.uniform(-0.4, 0.4) * h
        _, out = om.calc_sub_cell(cell, n, dstar)
        a = out[0]
        if not (1e-6 < a < 1 - 1e-6):
            continue
        st, res = om.find_signed_distance(cell, a, n)
        assert st == 0
        assert abs(res[0] - dstar) < 1e-12 * h + 1e-14, (res[0], dstar)
        _, chk = om.calc_sub_cell(cell, n, res[0])
        assert abs(chk[0] - a) < 1e-13
        n_checked += 1
    assert n_checked > 50


def test_find_signed_distance_special_cases(oracle_lib):
    om = oracle_lib.OracleMesh(M.hex_box(4))
    cell = 21
    # zero normal: status sign(0.5 - alpha) and nothing written (cutCell.C:625-628)
    st, out = om.find_signed_distance(cell, 0.3, np.zeros(3))
    assert st == 1 and out[7] == 0
    st, out = om.find_signed_distance(cell, 0.7, np.zeros(3))
    assert st == -1
    # axis-aligned normal: tied vertex levels take the coincident-level exit
    # or the cubic; alpha = 0.5 puts the plane through the cell centre
    for ax in range(3):
        n = np.zeros(3)
        n[ax] = 1.0
        st, out = om.find_signed_distance(cell, 0.5, n)
        assert st == 0
        assert abs(-out[0] - om.C[cell][ax]) < 1e-15
        assert abs(out[1 + ax] - om.C[cell][ax]) < 1e-15       # interface centre
        assert abs(abs(out[4 + ax]) - 0.0625) < 1e-15          # |S| = h^2
    # symmetric diagonal, alpha 0.5: plane through the centre
    n = np.ones(3) / math.sqrt(3)
    st, out = om.find_signed_distance(cell, 0.5, n)
    assert abs(out[0] + n @ om.C[cell]) < 1e-15


def test_interface_area_orientation_quirk(oracle_lib):
    """interfaceS is oriented against the origin, because subCellCentre_ is
    still zero when calcInterfaceCentreAndArea checks it (cutCell.C:96)."""
    om = oracle_lib.OracleMesh(M.hex_box(4))
    cell = 42
    c = om.C[cell]
    for n in (np.array([1.0, 0.2, 0.1]), np.array([-1.0, 0.3, -0.2])):
        n = n / np.linalg.norm(n)
        st, out = om.calc_sub_cell(cell, n, -(n @ c))
        ifc, ifs = out[5:8], out[8:11]
        assert ifs @ ifc >= 0.0
        assert abs(abs(ifs @ n) - np.linalg.norm(ifs)) < 1e-14


def _cell_points(mesh, cell):
    faces = [f for f in range(mesh.n_faces) if mesh.owner[f] == cell or
             (f < mesh.n_internal_faces and mesh.neighbour[f] == cell)]
    return faces, np.unique(np.concatenate([mesh.face(f) for f in faces]))


def test_sub_cell_closed_form_on_refined_and_brick_cells(oracle_lib):
    """Cells of a 2:1-refined hex mesh (a coarse cube whose faces carry the
    edge midpoints, up to 9 faces and 20 points) and of a brick mesh (boxes
    2h x 2h x h with split sides, up to 16 faces) cut by random planes: the
    volume fraction is the Scardovelli-Zaleski closed form of the box, so the
    collinear midpoints and split faces change nothing."""
    rng = np.random.default_rng(7)
    for mesh in (M.refined_hex_box(6), M.brick_box(4, seed=3)):
        om = oracle_lib.OracleMesh(mesh)
        counts = np.bincount(mesh.owner, minlength=mesh.n_cells) + \
            np.bincount(mesh.neighbour, minlength=mesh.n_cells)
        big = np.nonzero(counts > 6)[0]
        assert big.size > 10
        cut = 0
        for cell in rng.choice(big, 12, replace=False):
            _, pts = _cell_points(mesh, int(cell))
            P = mesh.points[pts]
            lo, hi = P.min(axis=0), P.max(axis=0)
            ext = hi - lo
            for _ in range(20):
                n = rng.normal(size=3)
                n /= np.linalg.norm(n)
                if np.min(np.abs(n)) < 1e-3:
                    continue
                # plane n.(x - lo) + dloc = 0 through the box; scaled to the unit cube
                lo_s, hi_s = float(np.minimum(n, 0) @ ext), float(np.maximum(n, 0) @ ext)
                dloc = -(lo_s + rng.uniform(0.05, 0.95) * (hi_s - lo_s))
                d = dloc - n @ lo
                st, out = om.calc_sub_cell(int(cell), n, d)
                m = n * ext
                beta = -dloc - sum(mi for mi in m if mi < 0)
                exact = cube_below(np.abs(m), beta)
                assert abs(out[0] - exact) < 1e-12, (mesh.name, int(cell), out[0], exact)
                cut += 1e-6 < exact < 1 - 1e-6
        assert cut > 150


def test_find_signed_distance_round_trip_refined_cells(oracle_lib):
    """The round trip d* -> alpha* -> d on the coarse cells of a 2:1-refined
    mesh: planar faces, so one cubic per vertex-level bracket, exact."""
    mesh = M.refined_hex_box(6)
    om = oracle_lib.OracleMesh(mesh)
    counts = np.bincount(mesh.owner, minlength=mesh.n_cells) + np.bincount(mesh.neighbour, minlength=mesh.n_cells)
    big = np.nonzero(counts > 6)[0]
    rng = np.random.default_rng(9)
    h = 1.0 / 6
    n_checked = 0
    for _ in range(150):
        cell = int(rng.choice(big))
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        dstar = -(n @ om.C[cell]) + rng.uniform(-0.4, 0.4) * h
        _, out = om.calc_sub_cell(cell, n, dstar)
        a = out[0]
        if not (1e-6 < a < 1 - 1e-6):
            continue
        st, res = om.find_signed_distance(cell, a, n)
        assert st == 0
        assert abs(res[0] - dstar) < 1e-12 * h + 1e-14
        n_checked += 1
    assert n_checked > 50
