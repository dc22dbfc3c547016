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


# --- Gauss pointLinear (fvSchemes grad(alpha1) of the reference's NAG case) ----

def _uniform_point_linear_normals(N, alpha):
    """Closed form of the pointLinear alphaGrad normal on a uniform N^3 mesh for
    cells whose faces have interior points only: each square face splits into
    four equal triangles about its centre, so its value is
    lin / 3 + (2 / 3) mean(point values), every interior point value is the
    mean of its 8 cells (equal inverse distances) and lin is the mean of the
    two cells.  Returns (cell ids, normals) in natural i + N j + N^2 k order."""
    a = alpha.reshape(N, N, N)                    # [k, j, i]
    P = np.full((N + 1, N + 1, N + 1), np.nan)    # interior point values
    P[1:N, 1:N, 1:N] = 0.125 * sum(a[dk:N - 1 + dk, dj:N - 1 + dj, di:N - 1 + di]
                                   for dk in (0, 1) for dj in (0, 1) for di in (0, 1))
    h = 1.0 / N
    out, ids = [], []
    for k in range(2, N - 2):
        for j in range(2, N - 2):
            for i in range(2, N - 2):
                g = np.zeros(3)
                for axis in range(3):
                    for side in (0, 1):
                        ijk = [i, j, k]
                        nb = list(ijk)
                        nb[axis] += 1 if side else -1
                        lin = 0.5 * (a[ijk[2], ijk[1], ijk[0]] + a[nb[2], nb[1], nb[0]])
                        # the face's 4 points: index along axis = cell index + side
                        pts = []
                        for u in (0, 1):
                            for v in (0, 1):
                                q = list(ijk)
                                q[axis] += side
                                o1, o2 = [x for x in range(3) if x != axis]
                                q[o1] += u
                                q[o2] += v
                                pts.append(P[q[2], q[1], q[0]])
                        val = lin / 3.0 + (2.0 / 3.0) * np.mean(pts)
                        g[axis] += (1.0 if side else -1.0) * val * h * h
                g /= h ** 3
                n = -g
                out.append(n / (np.linalg.norm(n) + 1e-15))
                ids.append(i + N * j + N * N * k)
    return np.array(ids), np.array(out)


def _oracle_grad_normals(mesh, alpha, scheme):
    om = O.OracleMesh(mesh)
    s = O.OracleSolver(om, orientationMethod="alphaGrad", gradScheme=scheme)
    return s.grad_normals(alpha)


def test_point_linear_exact_for_linear_fields():
    import geometricvofext_amd as g
    N = 10
    mesh = g.hex_box(N)
    om = O.OracleMesh(mesh)
    grad = np.array([0.3, -0.2, 0.1])
    n = _oracle_grad_normals(mesh, 0.5 + (om.C - 0.5) @ grad, "Gauss pointLinear")
    inner = np.all((om.C > 2.5 / N) & (om.C < 1.0 - 2.5 / N), axis=1)
    assert np.max(np.abs(n[inner] + grad / np.linalg.norm(grad))) < 1e-12


def test_point_linear_closed_form_uniform_mesh():
    """On a uniform mesh with a random field the normal equals the closed form
    above, and differs from Gauss linear (the correction is not zero)."""
    import geometricvofext_amd as g
    N = 8
    mesh = g.hex_box(N)
    alpha = np.random.default_rng(4).uniform(0.0, 1.0, mesh.n_cells)
    ids, ref = _uniform_point_linear_normals(N, alpha)
    got = _oracle_grad_normals(mesh, alpha, "Gauss pointLinear")[ids]
    assert np.max(np.abs(got - ref)) < 1e-12
    lin = _oracle_grad_normals(mesh, alpha, "Gauss linear")[ids]
    assert np.max(np.abs(lin - ref)) > 1e-2


def test_point_linear_pin_rejects_open_fan():
    """A fan without its closing triangle (pi, p_0, p_last) breaks the closed
    form.  (Which face weight pi takes, the owner cell's index or the
    face's, cannot show on a uniform mesh, where all weights are 1/2: that
    detail of OpenFOAM's pointLinear is unpinned here.)"""
    import geometricvofext_amd as g
    N = 8
    mesh = g.hex_box(N)
    alpha = np.random.default_rng(4).uniform(0.0, 1.0, mesh.n_cells)
    ids, ref = _uniform_point_linear_normals(N, alpha)
    with O.mutation(5):
        got = _oracle_grad_normals(mesh, alpha, "Gauss pointLinear")[ids]
    assert np.max(np.abs(got - ref)) > 1e-3
