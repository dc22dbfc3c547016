"""The benchmark configurations run to their end time on the GPU:

* the reference's own 64^3 deformation case (its t=0 exact field, LS,
  nAlphaBounds 3, no clipping: tutorials/test/plicVofAdvectionFoam) to t = 3
  in lockstep with the oracle, alpha within 1e-12 after every step, with the
  shape error against the 13 exact fields at the write times;
* the tetrahedral workload at full size (96^3 x 6 = 5.3M cells, 8 spheres,
  solid-body rotation) through size-independent properties, and at 24^3 x 6
  in lockstep with the oracle for a quarter revolution;
* the 128^3 perturbed-hex workload to t = 3 in lockstep with the oracle
  (minutes of CPU time: only with SVP_FULL_RUNS=1).
"""
import os
import sys
import time

import numpy as np
import pytest

from tests.helpers import bits_equal, max_abs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-12
THREADS = 16   # the host CPU share of one GPU on the box


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _lockstep(gpu, ors, om, a0, U0, phi0, end_time, period, check_every=1, on_write=None, label=""):
    """Harness loop of plicVofAdvectionFoam (plicVof.H:13-57) driving both
    solvers: Courant number and time step from the oracle's field, flow
    scaled on the GPU and set on the oracle, reconstruct + advect, alpha
    compared.  Returns (steps, bit-exact steps, worst difference)."""
    from geometricvofext_amd import harness as H
    rt = H.RunTime(end_time=end_time)
    gpu.set_alpha(a0)
    ors.set_alpha(a0)
    phi = np.zeros_like(phi0)
    n = exact = 0
    worst = 0.0
    t0 = time.time()
    while rt.run():
        co = om.courant(phi, ors.alpha())
        rt.set_delta_t(0.5 * co[0] * rt.delta_t, 0.5 * co[1] * rt.delta_t)
        rt.increment()
        uf = H.u_factor(rt.value, rt.delta_t, period)
        phi = phi0 * uf
        gpu.scale_flow(uf)
        ors.set_flow(phi, U0 * uf)
        gpu.reconstruct()
        ors.reconstruct()
        gpu.advect(rt.delta_t)
        ors.advect(rt.delta_t)
        n += 1
        if n % check_every == 0 or rt.write_time or not rt.run():
            ga, oa = gpu.alpha(), ors.alpha()
            d = max_abs(ga, oa)
            worst = max(worst, d)
            exact += bits_equal(ga, oa)
            assert d <= TOL, f"{label} step {n} (t={rt.value:.4f}): alpha off by {d:.2e}"
            if rt.write_time and on_write:
                on_write(rt.value, ga, oa)
        if n % 100 == 0:
            _say(f"{label}: step {n}, t={rt.value:.4f}, {time.time() - t0:.0f}s, worst |diff| {worst:.1e}")
    return n, exact, worst


def test_config0_reference_case_to_end_time(oracle_lib):
    import geometricvofext_amd as g
    from geometricvofext_amd import harness as H
    d = np.load(os.path.join(HERE, "golden", "exact_solutions_64.npz"))
    times, ex = d["times"], d["alpha"]
    mesh = g.hex_box(64)
    om = oracle_lib.OracleMesh(mesh)
    dct = dict(orientationMethod="LS", nAlphaBounds=3, snapTol=0.0, clip=False, mixedCellTol=1e-8)
    gpu = g.SolveVofEqu(mesh, dct)
    ors = oracle_lib.OracleSolver(om, nAlphaBounds=3, snapTol=0.0, clip=False, threads=THREADS)
    U0, phi0 = gpu.deformation_field()
    U0o, phi0o = om.deformation_field()
    assert bits_equal(U0, U0o) and bits_equal(phi0, phi0o)
    gpu.upload_base_flow(U0, phi0)
    V = om.V
    v0 = float(np.sum(ex[0] * V))
    table = []

    def on_write(t, ga, oa):
        k = int(round(t / 0.25))
        es_g, es_o = H.shape_error(ga, ex[k], V), H.shape_error(oa, ex[k], V)
        ev = H.volume_error(ga, V, v0)
        table.append((t, es_g, ev))
        assert es_g == es_o
        assert abs(ev) < 1e-12, f"t={t}: volume error {ev:.2e}"
        _say(f"config 0: t={t:.2f}  E_s vs exact field {es_g:.4f}  E_v {ev:.1e}")

    _say("config 0: 64^3, fixture sphere, deformation to t=3, lockstep with the oracle")
    n, exact, worst = _lockstep(gpu, ors, om, ex[0], U0, phi0, 3.0, 6.0, on_write=on_write, label="config 0")
    a = gpu.alpha()
    es_return = H.shape_error(a, ex[0], V)
    _say(f"config 0: {n} steps, alpha bit-identical on {exact} of {n} steps (worst {worst:.1e}); "
         f"E_s(3) vs t=0 field {es_return:.4f}")
    assert [round(t, 6) for t, _, _ in table] == [round(0.25 * k, 6) for k in range(1, 13)]
    assert exact == n
    # the restatement's measured accuracy on this case (E_s(3) = 0.381)
    assert 0.3 < es_return < 0.45


def test_config2_tets_full_size_properties():
    """5.3M tetrahedra, 8 spheres, rotation: volume conserved to 1e-12
    relative over 60 steps, mixed counts equal the host count of the field,
    and a second run is bit-identical."""
    import bench
    wl = bench.WORKLOADS["tet96"]
    mesh, gpu, a0, U0, phi0 = bench.build_case(wl)
    _say("config 2:", mesh.n_cells, "tetrahedra")
    assert mesh.n_cells == 6 * 96 ** 3
    V = gpu.geometry()["V"]
    seq = bench.time_sequence(gpu, a0, wl["end_time"], max_steps=60, period=bench.flow_period(wl))[:60]
    v0 = float(np.sum(a0 * V))

    def run(check):
        gpu.set_alpha(a0)
        for k, (_, dt, uf) in enumerate(seq):
            if check and k % 20 == 0:
                a = gpu.alpha()
                expect = int(np.count_nonzero((a > 1e-8) & (a < 1.0 - 1e-8)))
            gpu.scale_flow(uf)
            gpu.reconstruct()
            if check and k % 20 == 0:
                assert gpu.n_mixed() == expect
            gpu.advect(dt)
        return gpu.alpha()

    a1 = run(True)
    drift = abs(float(np.sum(a1 * V)) - v0) / v0
    a2 = run(False)
    _say(f"config 2: {len(seq)} steps, relative volume change {drift:.1e}, rerun bit-identical "
         f"{bits_equal(a1, a2)}")
    assert drift < 1e-12
    assert bits_equal(a1, a2)


def test_config2_tets_reduced_lockstep(oracle_lib):
    """The same workload at 24^3 x 6 for a quarter revolution, every step
    against the oracle."""
    import geometricvofext_amd as g
    from geometricvofext_amd import cases
    mesh = g.tet_box(24, perturb=0.1, seed=2)
    om = oracle_lib.OracleMesh(mesh)
    dct = dict(orientationMethod="LS", nAlphaBounds=3, snapTol=0.0, clip=False)
    gpu = g.SolveVofEqu(mesh, dct)
    gpu.init_alpha(cases.random_spheres(7), max_depth=4)
    a0 = gpu.alpha()
    omega = cases.rotation_omega(3, 1.0)
    U0, phi0 = gpu.rotation_field(omega, (0.5, 0.5, 0.5))
    gpu.upload_base_flow(U0, phi0)
    ors = oracle_lib.OracleSolver(om, nAlphaBounds=3, snapTol=0.0, clip=False, threads=THREADS)
    _say("config 2 reduced:", mesh.n_cells, "tetrahedra, a quarter revolution")
    n, exact, worst = _lockstep(gpu, ors, om, a0, U0, phi0, 0.25, 0.0, label="config 2 reduced")
    _say(f"config 2 reduced: {n} steps, bit-identical on {exact} (worst {worst:.1e})")
    assert n > 50 and exact == n


@pytest.mark.skipif(not os.environ.get("SVP_FULL_RUNS"),
                    reason="the 128^3 run to t=3 with the oracle in lockstep takes minutes of CPU time; "
                           "run with SVP_FULL_RUNS=1")
def test_config1_perturbed_hex_to_end_time(oracle_lib):
    import bench
    from geometricvofext_amd import harness as H
    wl = bench.WORKLOADS["hex128p"]
    mesh, gpu, a0, U0, phi0 = bench.build_case(wl)
    om = oracle_lib.OracleMesh(mesh)
    ors = oracle_lib.OracleSolver(om, nAlphaBounds=3, snapTol=0.0, clip=False, threads=THREADS)
    _say("config 1: 128^3 perturbed hex to t=3, lockstep with the oracle")
    n, exact, worst = _lockstep(gpu, ors, om, a0, U0, phi0, 3.0, 6.0, label="config 1")
    a = gpu.alpha()
    es = H.shape_error(a, a0, om.V)
    ev = H.volume_error(a, om.V, float(np.sum(a0 * om.V)))
    _say(f"config 1: {n} steps, alpha bit-identical on {exact} of {n} steps (worst {worst:.1e}); "
         f"E_s(3) {es:.4f}, E_v {ev:.1e}")
    assert exact == n and abs(ev) < 1e-12
