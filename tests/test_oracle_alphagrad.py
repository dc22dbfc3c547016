"""Pins the oracle's alphaGrad normal (reconstruction.C:74-82: minus the
Gauss-linear gradient of alpha, normalised) before it checks the GPU kernel:
Gauss linear is exact for a linear field in the interior of a uniform mesh,
the complement field turns the normal around on closed interior cells, and
on a perturbed mesh the normal equals the face sum written out independently
here in numpy (face loop, linear weights from face and cell centres)."""
import numpy as np
import pytest

from oracle import oracle as O


def _normals(mesh, alpha):
    om = O.OracleMesh(mesh)
    s = O.OracleSolver(om, orientationMethod="alphaGrad")
    s.set_alpha(alpha)
    s.reconstruct()
    mx = s.mixed()
    return om, mx["cells"], np.asarray(mx["n"]).reshape(-1, 3)


def _interior(mesh_n, C, h):
    return np.all((C > 1.5 * h) & (C < 1.0 - 1.5 * h), axis=1)


def test_linear_field_gives_its_gradient():
    import geometricvofext_amd as g
    N = 10
    mesh = g.hex_box(N)
    om = O.OracleMesh(mesh)
    grad = np.array([0.3, -0.2, 0.1])
    alpha = 0.5 + (om.C - 0.5) @ grad
    _, cells, n = _normals(mesh, alpha)
    assert len(cells) == mesh.n_cells           # every cell is mixed
    inner = _interior(N, om.C[cells], 1.0 / N)
    expect = -grad / np.linalg.norm(grad)
    assert inner.sum() > 100
    assert np.max(np.abs(n[inner] - expect)) < 1e-12


def test_complement_flips_the_normal():
    """alpha -> 1 - alpha negates the gradient on closed interior cells (the
    face areas sum to zero), so the normal turns around."""
    import geometricvofext_amd as g
    mesh = g.hex_box(8, perturb=0.2, seed=4)
    om = O.OracleMesh(mesh)
    alpha = 0.5 + 0.3 * np.sin(3.0 * om.C[:, 0]) * np.cos(2.0 * om.C[:, 1]) + 0.1 * om.C[:, 2]
    _, c1, n1 = _normals(mesh, alpha)
    _, c2, n2 = _normals(mesh, 1.0 - alpha)
    assert np.array_equal(c1, c2)
    inner = _interior(8, om.C[c1], 1.0 / 8)
    assert inner.sum() > 50
    assert np.max(np.abs(n1[inner] + n2[inner])) < 1e-9


def test_perturbed_mesh_matches_face_sum():
    import geometricvofext_amd as g
    mesh = g.hex_box(6, perturb=0.2, seed=9)
    om = O.OracleMesh(mesh)
    rng = np.random.default_rng(3)
    alpha = rng.uniform(0.05, 0.95, mesh.n_cells)
    _, cells, n = _normals(mesh, alpha)
    own, nbr = np.asarray(mesh.owner), np.asarray(mesh.neighbour)
    nif = mesh.n_internal_faces
    Cf, Sf, C, V = om.Cf, om.Sf, om.C, om.V
    gsum = np.zeros((mesh.n_cells, 3))
    for f in range(nif):
        o, k = own[f], nbr[f]
        wo = abs(Sf[f] @ (Cf[f] - C[o]))
        wn = abs(Sf[f] @ (C[k] - Cf[f]))
        a_f = wn / (wo + wn) * alpha[o] + wo / (wo + wn) * alpha[k]
        gsum[o] += Sf[f] * a_f
        gsum[k] -= Sf[f] * a_f
    from geometricvofext_amd.mesh import PATCH_EMPTY
    for start, size, ptype in zip(mesh.patch_start, mesh.patch_size, mesh.patch_type):
        if ptype == PATCH_EMPTY:
            continue
        for f in range(int(start), int(start + size)):
            gsum[own[f]] += Sf[f] * alpha[own[f]]
    grad = gsum / V[:, None]
    expect = -grad / (np.linalg.norm(grad, axis=1)[:, None] + 1e-15)
    assert np.max(np.abs(n - expect[cells])) < 1e-10
