"""Initial alpha by adaptive tetrahedral subdivision (the harness's exact-cut
initializer): the liquid volume of an ellipsoid against 4/3 pi abc, cells
well inside and well outside the shape exactly 1 and 0, and the union of
two overlapping shapes against the inclusion-exclusion volume measured on
the same mesh."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DICT = dict(orientationMethod="LS", nAlphaBounds=3, snapTol=0.0, clip=False)


def _rot(seed):
    from geometricvofext_amd import cases
    return cases.random_rotation(np.random.default_rng(seed))


@pytest.mark.parametrize("mesh_kind", ["hex", "tet"])
def test_ellipsoid_volume(mesh_kind):
    import geometricvofext_amd as g
    n = 40 if mesh_kind == "hex" else 20
    mesh = g.hex_box(n) if mesh_kind == "hex" else g.tet_box(n, perturb=0.1, seed=3)
    s = g.SolveVofEqu(mesh, DICT)
    R = _rot(5)
    a = np.array([0.21, 0.11, 0.07])
    c = np.array([0.48, 0.52, 0.5])
    s.init_alpha([(c, a, R)], max_depth=4)
    alpha = s.alpha()
    geo = s.geometry()
    V, C = geo["V"], geo["C"]
    assert np.all(alpha >= 0.0) and np.all(alpha <= 1.0)
    exact = 4.0 / 3.0 * np.pi * a.prod()
    got = float(np.sum(alpha * V))
    print(f"{mesh_kind}: ellipsoid volume {got:.9f} vs {exact:.9f} (rel {got / exact - 1:.2e})")
    assert abs(got / exact - 1.0) < 1e-3
    # cells whose centre lies far inside / outside in the shape's frame (a
    # cell's points lie within 2.5 lattice spacings of its centre)
    q = ((C - c) @ R.T) / a
    r = np.linalg.norm(q, axis=1)
    far = 2.5 / n / a.min()
    assert np.all(alpha[r < 1.0 - far] == 1.0)
    assert np.all(alpha[r > 1.0 + far] == 0.0)


def test_union_of_overlapping_shapes():
    import geometricvofext_amd as g
    mesh = g.hex_box(32)
    s = g.SolveVofEqu(mesh, DICT)
    V = s.geometry()["V"]
    A = (np.array([0.45, 0.5, 0.5]), 0.2)
    B = (np.array([0.6, 0.5, 0.52]), np.array([0.18, 0.09, 0.12]), _rot(9))
    vols = {}
    for name, shapes in (("A", [A]), ("B", [B]), ("AB", [A, B])):
        s.init_alpha(shapes, max_depth=4)
        vols[name] = float(np.sum(s.alpha() * V))
    # |A u B| = |A| + |B| - |A n B| with |A n B| > 0 for these shapes
    inter = vols["A"] + vols["B"] - vols["AB"]
    print(f"union {vols['AB']:.6f}, parts {vols['A']:.6f} + {vols['B']:.6f}, overlap {inter:.6f}")
    assert 0.0 < inter < min(vols["A"], vols["B"])
    assert abs(vols["A"] / (4.0 / 3.0 * np.pi * 0.2 ** 3) - 1.0) < 1e-3
