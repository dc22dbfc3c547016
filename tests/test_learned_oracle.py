"""Pins the CPU restatement of the learned-normal pieces (oracle/learned_oracle.py)
before it is used to check the HIP kernels: the cube-fraction closed form
against the oracle's exact polyhedral cut, the symmetry canonicalisation by
invariance, and the network backward by finite differences."""
import itertools

import numpy as np
import pytest

from oracle import learned_oracle as LO


@pytest.fixture(scope="module")
def unit_cube(oracle_lib):
    from geometricvofext_amd.mesh import hex_box
    m = hex_box(1)
    return oracle_lib.OracleMesh(m)


def _random_planes(rng, k):
    out = []
    for i in range(k):
        n = rng.normal(size=3)
        if i % 5 == 1:
            n[rng.integers(3)] = 0.0          # plane parallel to an axis
        if i % 5 == 2:
            n[:2] = 0.0                       # axis-aligned plane
        if i % 5 == 3:
            n[rng.integers(3)] *= 1e-13       # nearly parallel
        n /= np.linalg.norm(n)
        # offset spread over the whole cut range of the unit cube
        lo = -np.sum(np.maximum(n, 0.0))
        hi = -np.sum(np.minimum(n, 0.0))
        d = lo + (hi - lo) * rng.uniform(-0.05, 1.05)
        out.append((n, d))
    return out


def test_cube_fraction_matches_exact_cut(unit_cube):
    rng = np.random.default_rng(11)
    worst = 0.0
    for n, d in _random_planes(rng, 400):
        st, out = unit_cube.calc_sub_cell(0, n, d)
        exact = out[0] if st == 0 else (1.0 if st == -1 else 0.0)
        f = LO.plane_cube_fraction(n, d, (0.0, 0.0, 0.0), 1.0)
        worst = max(worst, abs(f - exact))
    assert worst < 1e-12, worst


def test_cube_fraction_symmetries():
    rng = np.random.default_rng(3)
    for _ in range(200):
        m = rng.uniform(0, 1, 3)
        a = rng.uniform(0, 1)
        f = LO.cube_fraction(m[0], m[1], m[2], a * m.sum())
        # complement: the other side of the plane in the point-reflected cube
        g = LO.cube_fraction(m[0], m[1], m[2], (1 - a) * m.sum())
        assert abs(f + g - 1.0) < 1e-13
        # permutation invariance
        for p in itertools.permutations(range(3)):
            assert abs(LO.cube_fraction(m[p[0]], m[p[1]], m[p[2]], a * m.sum()) - f) < 1e-13
    # closed cases: a plane cutting off one corner, one half
    assert abs(LO.cube_fraction(1, 1, 1, 1.0) - 1.0 / 6.0) < 1e-15
    assert abs(LO.cube_fraction(1, 0, 0, 0.5) - 0.5) < 1e-15
    assert abs(LO.cube_fraction(1, 1, 0, 1.0) - 0.5) < 1e-15


def _plane_stencil(n, d):
    """Exact stencil alphas of unit cubes centred at {-1,0,1}^3 below n.x + d = 0."""
    a = np.zeros(27)
    for dx, dy, dz in itertools.product((-1, 0, 1), repeat=3):
        a[LO.lat_index(dx, dy, dz)] = LO.plane_cube_fraction(n, d, (dx - 0.5, dy - 0.5, dz - 0.5), 1.0)
    return a


def _symmetries():
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            yield perm, signs


def _apply(perm, signs, a):
    """Stencil of the configuration mapped by x'[k] = signs[k] x[perm[k]]."""
    b = np.zeros(27)
    for dx, dy, dz in itertools.product((-1, 0, 1), repeat=3):
        e = (dx, dy, dz)
        src = [0, 0, 0]
        for k in range(3):
            src[perm[k]] = signs[k] * e[k]
        b[LO.lat_index(*e)] = a[LO.lat_index(*src)]
    return b


def test_canonical_input_is_symmetry_invariant():
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(40):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        a = _plane_stencil(n, rng.uniform(-0.3, 0.3))
        est = np.abs(LO.youngs_normal(a))
        if min(est) < 1e-6 or min(abs(est[0] - est[1]), abs(est[1] - est[2]), abs(est[0] - est[2])) < 1e-6:
            continue
        x0 = LO.canonical_input(a)
        perm0, sgn0 = LO.canon_of(a)
        t0 = LO.to_canon(perm0, sgn0, n)
        for perm, signs in _symmetries():
            b = _apply(perm, signs, a)
            nb = [signs[k] * n[perm[k]] for k in range(3)]
            assert np.array_equal(LO.canonical_input(b), x0)
            pb, sb = LO.canon_of(b)
            assert np.allclose(LO.to_canon(pb, sb, nb), t0, atol=1e-15)
            assert np.allclose(LO.from_canon(pb, sb, LO.to_canon(pb, sb, nb)), nb, atol=1e-15)
        checked += 1
    assert checked > 20
    # the canonical estimate is non-negative and ordered
    est_c = LO.youngs_normal(_apply(*next(_symmetries()), a))
    assert len(est_c) == 3


def test_mlp_backward_matches_finite_differences():
    rng = np.random.default_rng(9)
    p = rng.normal(scale=0.2, size=LO.N_PARAMS)
    x = rng.uniform(-1, 1, size=(4, 27))
    y = rng.normal(size=(4, 3))
    loss, g = LO.mlp_loss_grad(p, x, y, rounded=False)
    idx = np.concatenate([rng.choice(np.arange(LO.P_W1, LO.P_B1), 8), rng.choice(np.arange(LO.P_B1, LO.P_W2), 4),
                          rng.choice(np.arange(LO.P_W2, LO.P_B2), 8), rng.choice(np.arange(LO.P_B2, LO.P_W3), 4),
                          rng.choice(np.arange(LO.P_W3, LO.P_B3), 6), np.arange(LO.P_B3, LO.N_PARAMS)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (LO.mlp_loss_grad(pp, x, y, rounded=False)[0] - LO.mlp_loss_grad(pm, x, y, rounded=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(g[i]), (i, fd, g[i])


def test_rounded_network_close_to_fp64():
    rng = np.random.default_rng(2)
    p = rng.normal(scale=0.1, size=LO.N_PARAMS)
    x = rng.uniform(-1, 1, size=(64, 27))
    xb = np.zeros((64, 32), np.uint16)
    xb[:, :27] = LO.f32_to_bf16(x.astype(np.float32))
    y = rng.normal(size=(64, 3))
    _, g64 = LO.mlp_loss_grad(p, LO.bf16_to_f(xb[:, :27]), y, rounded=False)
    _, gbf = LO.mlp_loss_grad(p, xb, y, rounded=True)
    assert np.linalg.norm(gbf - g64) < 2e-2 * np.linalg.norm(g64)


def test_residual_of_exact_plane_is_zero():
    rng = np.random.default_rng(4)
    for _ in range(20):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        d = rng.uniform(-0.4, 0.4)
        a = _plane_stencil(n, d)
        assert LO.residual(n, d, (-0.5, -0.5, -0.5), 1.0, a) < 1e-14
        # a tilted plane leaves a positive residual
        n2 = n + 0.3 * rng.normal(size=3)
        n2 /= np.linalg.norm(n2)
        assert LO.residual(n2, d, (-0.5, -0.5, -0.5), 1.0, a) > 1e-4


def test_cube_plane_offset_matches_find_signed_distance(unit_cube):
    """The plane of a cube at a given alpha (inverse of the cube fraction) is
    the plane the oracle's findSignedDistance places in the same cube
    (cutCell.C:611-799)."""
    rng = np.random.default_rng(8)
    for _ in range(100):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        a = rng.uniform(0.01, 0.99)
        d = LO.cube_plane_offset(n, a, (0.0, 0.0, 0.0), 1.0)
        assert abs(LO.plane_cube_fraction(n, d, (0.0, 0.0, 0.0), 1.0) - a) < 1e-9
        st, out = unit_cube.find_signed_distance(0, a, n)
        assert st == 0 and abs(out[0] - d) < 1e-10, (out[0], d)


def test_cube_plane_offset_round_trip():
    """Every piece of the inverse (both closed forms, the linear piece, both
    cubic pieces) and the degenerate normals (one or two zero components,
    equal components): the fraction of the returned plane is alpha."""
    rng = np.random.default_rng(12)
    pieces = set()
    for t in range(3000):
        n = rng.normal(size=3)
        if t % 5 == 1:
            n[t % 3] = 0.0
        if t % 5 == 2:
            n[:] = 0.0
            n[t % 3] = 1.0
        if t % 5 == 3:
            n[:] = rng.choice([-1.0, 1.0], 3)
        n /= np.linalg.norm(n)
        h = rng.uniform(0.1, 2.0)
        lo = rng.uniform(-1, 1, 3)
        a = rng.uniform(0.0, 1.0) if t % 7 else rng.choice([1e-9, 0.5, 1.0 - 1e-9])
        d = LO.cube_plane_offset(n, a, lo, h)
        assert abs(LO.plane_cube_fraction(n, d, lo, h) - a) < 1e-13, (n, a)
        m = sorted(abs(n) / np.abs(n).sum())
        v = min(a, 1.0 - a)
        lev = LO._level_sorted(*m, v)
        if lev < m[0]:
            pieces.add("cube root")
        elif lev < m[1]:
            pieces.add("square root")
        elif lev < min(m[0] + m[1], m[2]):
            pieces.add("first cubic")
        elif m[2] < m[0] + m[1]:
            pieces.add("second cubic")
        else:
            pieces.add("linear")
    assert pieces == {"cube root", "square root", "first cubic", "second cubic", "linear"}, pieces
