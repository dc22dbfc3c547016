"""Learned interface normal on the GPU against the CPU restatement
(oracle/learned_oracle.py) and the oracle's exact cut and plane placement.

Bars: cube fractions, plane training stencils, canonical inputs and the
plane placement fed by learned normals are bit-exact (same FP64 operations);
the BF16/FP32 network agrees with the FP64 network that applies the same BF16
rounding points to within FP32 accumulation (tolerances in each test);
curved training stencils agree with an independent midpoint-rule estimate.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import learned_oracle as LO  # noqa: E402


@pytest.fixture(scope="module")
def g():
    import geometricvofext_amd as g
    g.lib()
    return g


def test_cube_fractions_bit_exact(g):
    rng = np.random.default_rng(1)
    rows = []
    for i in range(2000):
        n = rng.normal(size=3)
        if i % 4 == 1:
            n[rng.integers(3)] = 0.0
        if i % 4 == 2:
            n[rng.integers(3)] *= 1e-12
        n /= np.linalg.norm(n)
        lo = rng.uniform(-2, 2, 3)
        h = rng.uniform(0.1, 2.0)
        d = -(n @ (lo + rng.uniform(-0.2, 1.2, 3) * h))
        rows.append([*n, d, *lo, h])
    rows = np.array(rows)
    got = g.cube_fractions(rows)
    ref = np.array([LO.plane_cube_fraction(r[:3], r[3], r[4:7], r[7]) for r in rows])
    assert np.array_equal(got, ref), np.abs(got - ref).max()


@pytest.fixture(scope="module")
def small_set(g):
    m = g.NormalModel(seed=3)
    m.generate(3000, seed=17, depth=5)
    return m, m.dataset(0, 3000)


def _cell_centres():
    return np.array([(L // 9 - 1, (L // 3) % 3 - 1, L % 3 - 1) for L in range(27)], dtype=np.float64)


def test_plane_samples_exact(small_set):
    _, ds = small_set
    cen = _cell_centres()
    k = np.where(ds["shape"][:, 0] == 0)[0][:300]
    assert k.size > 200
    for i in k:
        sh = ds["shape"][i]
        n, c = sh[16:19], sh[1:4]
        D = -((n[0] * c[0] + n[1] * c[1]) + n[2] * c[2])
        ref = [LO.plane_cube_fraction(n, D, cen[L] - 0.5, 1.0) for L in range(27)]
        assert np.array_equal(ds["alpha27"][i], np.array(ref))


def _implicit(sh, P):
    """Negative inside for the curved shapes of the 19-double record."""
    c, a, Q = sh[1:4], sh[4:7], sh[7:16].reshape(3, 3)
    e = (P - c) @ Q / a
    return np.linalg.norm(e, axis=-1) - 1.0


def test_curved_samples_against_midpoint_rule(small_set):
    _, ds = small_set
    cen = _cell_centres()
    N = 48
    u = (np.arange(N) + 0.5) / N - 0.5
    grid = np.stack(np.meshgrid(u, u, u, indexing="ij"), -1).reshape(-1, 3)
    worst = 0.0
    for kind in (1, 2):
        for i in np.where(ds["shape"][:, 0] == kind)[0][:6]:
            sh = ds["shape"][i]
            for L in range(27):
                frac = np.mean(_implicit(sh, grid + cen[L]) < 0.0)
                worst = max(worst, abs(frac - ds["alpha27"][i][L]))
    assert worst < 4e-3, worst


def test_targets_and_inputs(small_set):
    _, ds = small_set
    n = 400
    for i in range(n):
        a, sh = ds["alpha27"][i], ds["shape"][i]
        assert np.array_equal(ds["x"][i], LO.canonical_input(a))
        perm, sgn = LO.canon_of(a)
        t = LO.to_canon(perm, sgn, sh[16:19])
        assert np.allclose(ds["y"][i][:3], t, atol=1e-7)
        assert abs(np.linalg.norm(sh[16:19]) - 1.0) < 1e-12
        kind = int(sh[0])
        if kind == 1:
            c = sh[1:4]
            assert np.allclose(sh[16:19], -c / np.linalg.norm(c), atol=1e-14)
        if kind == 2:
            # nearest surface point x* to the cell centre: on the surface, and
            # the origin lies on the normal line through it
            c, ax, Q = sh[1:4], sh[4:7], sh[7:16].reshape(3, 3)
            y = -(c @ Q)
            f = lambda t: np.sum((ax * y / (ax * ax + t)) ** 2) - 1.0  # noqa: E731
            lo, hi = -ax.min() ** 2 * (1 - 1e-15), ax.max() * np.linalg.norm(y)
            from scipy.optimize import brentq
            t = brentq(f, lo + 1e-12 * abs(lo), hi, xtol=1e-15, rtol=1e-15, maxiter=500)
            xe = ax * ax * y / (ax * ax + t)
            xw = c + Q @ xe
            assert abs(_implicit(sh, xw[None])[0]) < 1e-9
            nrm = sh[16:19]
            assert np.linalg.norm(np.cross(-xw, nrm)) < 1e-7 * max(1.0, np.linalg.norm(xw))


def _random_inputs(rng, n):
    x = np.zeros((n, 32), np.uint16)
    x[:, :27] = LO.f32_to_bf16(rng.uniform(-1, 1, (n, 27)).astype(np.float32))
    return x


def test_forward_matches_fp64_network(g):
    rng = np.random.default_rng(4)
    m = g.NormalModel(seed=5)
    p = m.params
    x = _random_inputs(rng, 1000)
    got = m.forward(x)
    ref, _ = LO.mlp_forward(p, x, rounded=True)
    err = np.abs(got - ref)
    scale = np.abs(ref).max()
    # FP32 accumulation against FP64 sums; a BF16 rounding of a hidden value
    # may land one ulp apart, which moves an output by far less than 1e-2
    assert np.median(err) < 1e-5 * scale, np.median(err)
    assert err.max() < 1e-2 * scale, err.max()


def test_gradient_matches_fp64_network(g):
    rng = np.random.default_rng(6)
    m = g.NormalModel(seed=8)
    p = m.params
    x = _random_inputs(rng, 512)
    y = rng.normal(size=(512, 4)).astype(np.float32)
    gg, loss = m.gradient(x, y)
    ref_loss, ref = LO.mlp_loss_grad(p, x, y, rounded=True)
    assert abs(loss - ref_loss) < 1e-5 * ref_loss
    blocks = [(LO.P_W1, LO.P_B1), (LO.P_B1, LO.P_W2), (LO.P_W2, LO.P_B2), (LO.P_B2, LO.P_W3),
              (LO.P_W3, LO.P_B3), (LO.P_B3, LO.N_PARAMS)]
    for a, b in blocks:
        rel = np.linalg.norm(gg[a:b] - ref[a:b]) / np.linalg.norm(ref[a:b])
        assert rel < 5e-3, (a, b, rel)
    # and the FP64 gradient without rounding points is within BF16 distance
    _, ref64 = LO.mlp_loss_grad(p, x, y, rounded=False)
    assert np.linalg.norm(gg - ref64) < 3e-2 * np.linalg.norm(ref64)


def _angles(a, b):
    a = a / np.linalg.norm(a, axis=1, keepdims=True)
    b = b / np.linalg.norm(b, axis=1, keepdims=True)
    return np.degrees(np.arccos(np.clip(np.sum(a * b, 1), -1, 1)))


@pytest.fixture(scope="module")
def trained(g):
    m = g.NormalModel(seed=11)
    m.generate(1 << 19, seed=23, depth=5)
    losses, sec = m.train(epochs=6, batch=2048, lr0=3e-3, lr1=1e-4, seed=5)
    return m, losses, sec


def test_training_reduces_loss(trained):
    _, losses, sec = trained
    print(f"losses {losses} in {sec:.2f}s")
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.25 * losses[0]


def test_learned_beats_least_squares_on_held_out_shapes(g, trained, oracle_lib):
    """Mean angle between the predicted and the exact normal on held-out
    training-style stencils: network against the least-squares gradient over
    the 27-cell stencil (the plain method's normal on a uniform mesh)."""
    m, _, _ = trained
    held = g.NormalModel(seed=99)
    held.generate(3000, seed=1234, depth=5)
    ds = held.dataset(0, 3000)
    out = m.forward(ds["x"]).astype(np.float64)
    cen = _cell_centres()
    pred, ls, exact = [], [], []
    for i in range(3000):
        a = ds["alpha27"][i]
        perm, sgn = LO.canon_of(a)
        pred.append(LO.from_canon(perm, sgn, out[i]))
        ls.append(-oracle_lib.ls_grad(cen, a))
        exact.append(ds["shape"][i][16:19])
    e_net = _angles(np.array(pred), np.array(exact))
    e_ls = _angles(np.array(ls), np.array(exact))
    print(f"mean angle error: network {e_net.mean():.3f} deg, least squares {e_ls.mean():.3f} deg")
    assert e_net.mean() < e_ls.mean()


def test_learned_reconstruction_against_oracle(g, trained, oracle_lib):
    """Learned handle on a uniform 20^3 mesh: network normals where the
    stencil is a full lattice, plane placement bit-exact with the oracle's
    findSignedDistance on those normals, residuals as restated, fallback
    cells carrying the least-squares normal of the plain path."""
    from geometricvofext_amd import harness  # noqa: F401
    m, _, _ = trained
    mesh = g.hex_box(20)
    tol = 0.1
    d = dict(orientationMethod="learned", nAlphaBounds=3, snapTol=0.0, clip=False, learnedResidualTol=tol)
    s = g.SolveVofEqu(mesh, d)
    s.set_normal_model(m)
    # a sphere touching the wall region, so some interface cells lack a full stencil
    s.init_alpha([((0.45, 0.5, 0.3), 0.26)], max_depth=5)
    alpha = s.alpha()
    s.reconstruct()
    nmix, kept, fb, lsonly = s.learned_stats()
    print(f"interface cells {nmix}: learned {kept}, residual fallback {fb}, least squares only {lsonly}")
    assert nmix == kept + fb + lsonly and kept > 0
    itf = s.interface()
    res = s.learned_residuals()
    om = oracle_lib.OracleMesh(mesh)
    geo = s.geometry()
    C, V = geo["C"], geo["V"]
    h = V[0] ** (1.0 / 3.0)
    # plain-path least-squares normals for comparison
    plain = g.SolveVofEqu(mesh, dict(orientationMethod="LS", nAlphaBounds=3, snapTol=0.0, clip=False))
    plain.set_alpha(alpha)
    plain.reconstruct()
    pitf = plain.interface()
    assert np.array_equal(pitf["cells"], itf["cells"])
    n_kept = 0
    for k, c in enumerate(itf["cells"]):
        n = itf["n"][k]
        st, out = om.find_signed_distance(int(c), alpha[c], n)
        assert st == itf["status"][k]
        assert out[0] == itf["d"][k] or (st != 0 and itf["d"][k] == 0.0)
        if res[k] >= 0.0:
            # the cell had a full lattice stencil and a network normal
            off = np.rint((C - C[c]) / h).astype(int)
            near = np.where(np.all(np.abs(off) <= 1, axis=1))[0]
            a27 = np.zeros(27)
            c27 = np.zeros((27, 3))
            for j in near:
                L = LO.lat_index(*off[j])
                a27[L] = alpha[j]
                c27[L] = C[j]
            perm, sgn = LO.canon_of(a27)
            o = m.forward(LO.canonical_input(a27)[None]).astype(np.float64)[0]
            v = LO.from_canon(perm, sgn, o)
            mg = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
            v = np.array([v[0] / mg, v[1] / mg, v[2] / mg])
            # the residual of the network normal's cube plane
            lo = C[c] - 0.5 * h
            dv = LO.cube_plane_offset(v, alpha[c], lo, h)
            assert abs(LO.residual(v, dv, lo, h, a27) - res[k]) < 1e-12
            if np.allclose(n, v, atol=1e-12, rtol=0):
                n_kept += 1
                assert res[k] <= tol
                continue
            assert res[k] > tol
        # least-squares normal (no full stencil, or fallback)
        assert np.array_equal(n, pitf["n"][k])
    assert n_kept == kept


def test_learned_ellipsoid_deformation_to_end_time(g, trained, oracle_lib):
    """The learned-normal workload (3 seeded ellipsoids, deformation to
    t = 3) at 64^3: plain and learned paths to the end time.  Every 100 steps
    the learned path's plane placement is checked against the oracle's
    findSignedDistance on the GPU's normals (status and d bit-exact); at
    t = T the learned shape error is at most the plain one and volume is
    conserved."""
    import bench
    from geometricvofext_amd import cases, harness
    m, _, _ = trained
    mesh = g.hex_box(64)
    shapes = cases.random_ellipsoids(bench.LEARNED_WORKLOAD["shapes_seed"])
    om = oracle_lib.OracleMesh(mesh)
    res = {}
    seq = None
    for method in ("LS", "learned"):
        s = g.SolveVofEqu(mesh, dict(bench.ALPHA_DICT, orientationMethod=method))
        if method == "learned":
            s.set_normal_model(m)
        s.init_alpha(shapes, max_depth=5)
        a0 = s.alpha()
        U0, phi0 = s.deformation_field()
        s.upload_base_flow(U0, phi0)
        if seq is None:
            seq = bench.time_sequence(s, a0, 3.0)
        s.set_alpha(a0)
        checked = 0
        for k, (_, dt, uf) in enumerate(seq):
            s.scale_flow(uf)
            s.reconstruct()
            if method == "learned" and k % 100 == 0:
                alpha = s.alpha()
                itf = s.interface()
                for j, c in enumerate(itf["cells"]):
                    st, out = om.find_signed_distance(int(c), alpha[c], itf["n"][j])
                    assert st == itf["status"][j], (k, int(c))
                    assert out[0] == itf["d"][j] or (st != 0 and itf["d"][j] == 0.0), (k, int(c))
                checked += len(itf["cells"])
            s.advect(dt)
        V = s.geometry()["V"]
        a1 = s.alpha()
        res[method] = (harness.shape_error(a1, a0, V), harness.volume_error(a1, V, float(np.sum(a0 * V))))
        print(f"{method}: {len(seq)} steps, E_s {res[method][0]:.4f}, E_v {res[method][1]:.1e}"
              + (f", {checked} plane placements checked against the oracle" if checked else ""), flush=True)
        s.close()
    assert abs(res["LS"][1]) < 1e-12 and abs(res["learned"][1]) < 1e-12
    assert res["learned"][0] <= res["LS"][0]


def test_learned_on_non_lattice_mesh_is_the_plain_path(g, trained):
    """On a perturbed hexahedral mesh no cell has a uniform 3x3x3 lattice
    stencil, so every interface cell takes the least-squares normal: the
    learned handle must then step exactly as the plain one, bit for bit."""
    from geometricvofext_amd import harness
    m, _, _ = trained
    mesh = g.hex_box(14, perturb=0.15, seed=4)
    base = dict(nAlphaBounds=3, snapTol=0.0, clip=False)
    plain = g.SolveVofEqu(mesh, dict(base, orientationMethod="LS"))
    learned = g.SolveVofEqu(mesh, dict(base, orientationMethod="learned"))
    learned.set_normal_model(m)
    shapes = [((0.45, 0.5, 0.45), 0.22)]
    plain.init_alpha(shapes, max_depth=4)
    learned.set_alpha(plain.alpha())
    U0, phi0 = plain.deformation_field()
    for s in (plain, learned):
        s.upload_base_flow(U0, phi0)
    rt = harness.RunTime(end_time=0.2)
    n = 0
    while rt.run():
        co, aco = plain.courant()
        rt.set_delta_t(0.5 * co * rt.delta_t, 0.5 * aco * rt.delta_t)
        rt.increment()
        uf = harness.u_factor(rt.value, rt.delta_t, 6.0)
        for s in (plain, learned):
            s.scale_flow(uf)
            s.reconstruct()
        nmix, kept, fb, lsonly = learned.learned_stats()
        assert kept == 0 and fb == 0 and lsonly == nmix
        for s in (plain, learned):
            s.advect(rt.delta_t)
        assert np.array_equal(plain.alpha(), learned.alpha()), f"step {n}"
        n += 1
    assert n > 10
    print(f"non-lattice mesh: {n} steps, learned handle bit-identical to the plain path")
