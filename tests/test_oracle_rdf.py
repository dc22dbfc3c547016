"""Pins the oracle's isoRDF normal (reconstruction.C:194-405 with OpenFOAM's
reconstructedDistanceFunction, a dependency that is not in the reference)
before it checks the GPU kernels:

* a plane is a fixed point: with exact volume fractions of a plane, the
  iteration (planes -> distance function -> least-squares gradient ->
  normals) converges to the exact plane normal, from a least-squares start
  that is off by about 0.1;
* one iteration returns the least-squares gradient normalised, which is the
  LS normal up to its + SMALL in the denominator;
* on a sphere the RDF normal is more accurate than the LS normal;
* the iteration count follows the reference's stopping rule (at least two
  iterations unless one is asked for, never more than asked for).
"""
import numpy as np

from oracle import oracle as O
from tests.test_oracle_cut import cube_fraction


def _plane_case(N=16, n=(0.3, 0.5, 0.81), offset=-0.013):
    import geometricvofext_amd as g
    h = 1.0 / N
    mesh = g.hex_box(N)
    om = O.OracleMesh(mesh)
    n = np.asarray(n, dtype=np.float64)
    n = n / np.linalg.norm(n)
    d = -np.dot(n, [0.5, 0.5, 0.5]) + offset
    a = np.array([cube_fraction(n, d + np.dot(n, om.C[c] - h / 2), h) for c in range(mesh.n_cells)])
    return om, n, a


def _normals(om, a, **kw):
    s = O.OracleSolver(om, **kw)
    s.set_alpha(a)
    s.reconstruct()
    mx = s.mixed()
    return s, mx


def test_plane_is_a_fixed_point():
    om, n, a = _plane_case()
    _, ls = _normals(om, a, orientationMethod="LS")
    s, rdf = _normals(om, a, orientationMethod="RDF", iterations=30, tol=1e-15, relTol=1e-15)
    C = om.C[rdf["cells"]]
    inner = np.all((C > 0.2) & (C < 0.8), axis=1)
    assert inner.sum() > 150
    assert np.max(np.linalg.norm(ls["n"][inner] - n, axis=1)) > 0.05
    assert np.max(np.linalg.norm(rdf["n"][inner] - n, axis=1)) < 1e-7
    assert np.max(np.linalg.norm(rdf["n"] - n, axis=1)) < 1e-5
    assert s.rdf_iterations_run() > 2


def test_one_iteration_is_the_normalised_ls_gradient():
    om, n, a = _plane_case(N=10, n=(0.7, -0.2, 0.4), offset=0.02)
    _, ls = _normals(om, a, orientationMethod="LS")
    s, r1 = _normals(om, a, orientationMethod="RDF", iterations=1)
    assert s.rdf_iterations_run() == 1
    assert np.array_equal(ls["cells"], r1["cells"])
    assert np.max(np.abs(ls["n"] - r1["n"])) < 1e-13


def _sphere_alpha(om, mesh, centre, r, N, sub=12):
    h = 1.0 / N
    u = (np.arange(sub) + 0.5) / sub - 0.5
    X, Y, Z = np.meshgrid(u, u, u, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) * h
    dist = np.linalg.norm(om.C - centre, axis=1)
    a = (dist < r).astype(np.float64)
    near = np.nonzero(np.abs(dist - r) < h)[0]
    for c in near:
        a[c] = np.mean(np.linalg.norm(om.C[c] + P - centre, axis=1) < r)
    return a


def test_sphere_rdf_beats_ls():
    import geometricvofext_amd as g
    N = 16
    mesh = g.hex_box(N)
    om = O.OracleMesh(mesh)
    centre = np.array([0.52, 0.49, 0.5])
    a = _sphere_alpha(om, mesh, centre, 0.3, N)

    def err(mx):
        r = om.C[mx["cells"]] - centre
        r /= np.linalg.norm(r, axis=1)[:, None]
        return np.degrees(np.arccos(np.clip(np.sum(mx["n"] * r, axis=1), -1, 1)))

    _, ls = _normals(om, a, orientationMethod="LS")
    _, rdf = _normals(om, a, orientationMethod="RDF")
    assert np.array_equal(ls["cells"], rdf["cells"])
    assert err(rdf).mean() < 0.8 * err(ls).mean()


def test_iteration_control():
    om, n, a = _plane_case(N=10)
    for it in (2, 3, 5):
        s, _ = _normals(om, a, orientationMethod="RDF", iterations=it)
        assert 2 <= s.rdf_iterations_run() <= it
    s, _ = _normals(om, a, orientationMethod="RDF", iterations=4, tol=1.0)
    assert s.rdf_iterations_run() == 2          # converged at the first iteration allowed to stop
