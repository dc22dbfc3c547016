"""Pins the oracle's mapAlphaField (reconstruction.C:725-784) before it
checks the GPU entry point: with the exact plane of a slanted interface in
every cell the mapped fractions equal the Scardovelli-Zaleski closed form,
cells outside the refine levels keep their alpha, and a zero normal gives 0
or 1 by the sign of d (the uncut branches of calcSubCell)."""
import numpy as np

from oracle import oracle as O
from tests.test_oracle_cut import cube_fraction


def test_exact_plane_fractions_and_levels():
    import geometricvofext_amd as g
    N = 8
    h = 1.0 / N
    mesh = g.hex_box(N)
    om = O.OracleMesh(mesh)
    n = np.array([0.6, -0.3, 0.74])
    n /= np.linalg.norm(n)
    d0 = -np.dot(n, [0.5, 0.5, 0.5])
    exact = np.array([cube_fraction(n, d0 + np.dot(n, om.C[c] - h / 2), h) for c in range(mesh.n_cells)])
    rng = np.random.default_rng(5)
    alpha = rng.uniform(0.0, 1.0, mesh.n_cells)
    s = O.OracleSolver(om)
    s.set_alpha(alpha)
    nn = np.tile(n, (mesh.n_cells, 1))
    dd = np.full(mesh.n_cells, d0)
    s.map_alpha_field(nn, dd, 0.2, 0.8)
    out = s.alpha()
    inside = (alpha >= 0.2) & (alpha <= 0.8)
    assert inside.sum() > 100 and (~inside).sum() > 100
    assert np.array_equal(out[~inside], alpha[~inside])
    assert np.max(np.abs(out[inside] - exact[inside])) < 1e-13


def test_zero_normal_branch():
    import geometricvofext_amd as g
    mesh = g.hex_box(3)
    om = O.OracleMesh(mesh)
    s = O.OracleSolver(om)
    s.set_alpha(np.full(mesh.n_cells, 0.5))
    d = np.where(np.arange(mesh.n_cells) % 2 == 0, -0.3, 0.3)
    s.map_alpha_field(np.zeros((mesh.n_cells, 3)), d, 0.0, 1.0)
    assert np.array_equal(s.alpha(), np.where(d < 0, 1.0, 0.0))
