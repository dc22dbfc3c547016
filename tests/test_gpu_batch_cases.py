"""Batched cases (svp_batch_*): independent handles advanced together, one
launch per stage over all cases, each case with its own mesh size, shapes,
flow, Courant limit and end time (the harness loop of plicVof.H:13-57 per
case).  Bar: every case bit-identical, every step, to the same case stepped
alone through its own handle, and, for a subset, to the oracle; cases that
reach their end time drop out while the others continue."""
import sys

import numpy as np
import pytest

from tests.helpers import bits_equal, max_abs

pytestmark = pytest.mark.gpu

DICT = dict(orientationMethod="LS", nAlphaBounds=3, snapTol=0.0, clip=False, mixedCellTol=1e-8)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _make(specs):
    from geometricvofext_amd import cases
    return [cases.make_batch_case(sp, DICT) for sp in specs]


def _alone_step(c):
    """One harness step of a case on its own handle; returns its dt (0 when done)."""
    from geometricvofext_amd import harness
    s, rt = c["solver"], c["rt"]
    if not rt.run():
        return 0.0
    co, aco = s.courant()
    rt.set_delta_t(0.5 * co * rt.delta_t, 0.5 * aco * rt.delta_t)
    rt.increment()
    s.scale_flow(harness.u_factor(rt.value, rt.delta_t, c["period"]))
    s.reconstruct()
    s.advect(rt.delta_t)
    return rt.delta_t


def test_batch_matches_separate_handles():
    import geometricvofext_amd as g
    from geometricvofext_amd import cases
    specs = cases.batch_cases(7, seed=11, nmin=9, nmax=15, tmin=0.04, tmax=0.12)
    batched, alone = _make(specs), _make(specs)
    b = g.CaseBatch([c["solver"] for c in batched])
    _say("batch of", len(specs), "cases:", [sp["N"] for sp in specs], [sp["flow"][0] for sp in specs])
    n_steps = 0
    finished = set()
    while any(c["rt"].run() for c in batched):
        co, aco = b.courant()
        dt, fac = cases.batch_time_step(batched, co, aco)
        b.scale_flow(fac)
        b.step(dt)
        for k, c in enumerate(alone):
            dk = _alone_step(c)
            assert dk == dt[k], (n_steps, k)
            if dk == 0.0:
                finished.add(k)
            ab, aa = batched[k]["solver"].alpha(), c["solver"].alpha()
            assert bits_equal(ab, aa), f"step {n_steps} case {k}: alpha off by {max_abs(ab, aa)}"
        if n_steps % 10 == 0:
            for k in range(len(specs)):
                if k in finished:
                    continue
                bi, ai = batched[k]["solver"].interface(), alone[k]["solver"].interface()
                assert np.array_equal(bi["cells"], ai["cells"]) and bits_equal(bi["d"], ai["d"])
        n_steps += 1
    assert 0 < len(finished) and n_steps > 20
    sec, nic = b.timers()
    _say(f"batch: {n_steps} steps, {nic} interface-cell updates in {sec * 1e3:.1f} ms of batch steps, "
         f"every case bit-identical to its handle alone")
    b.close()


@pytest.mark.parametrize("kind", ["tet", "perturbed_hex"])
def test_batch_other_cells_match_separate_handles(kind):
    """Batches of tetrahedral cases and of perturbed-hexahedron cases (warped
    faces: the handle alone integrates them with the four-lane fan kernel, the
    batch with its one-lane face kernel) against the same cases stepped alone,
    bit for bit every step."""
    import geometricvofext_amd as g
    from geometricvofext_amd import cases
    specs = cases.batch_cases(4, seed=5, nmin=6, nmax=9, tmin=0.03, tmax=0.06)
    for k, sp in enumerate(specs):
        if kind == "tet":
            sp.update(kind="tet", N=sp["N"] - 2, perturb=0.1, seed=k)
        else:
            sp.update(perturb=0.2, seed=k)
    batched, alone = _make(specs), _make(specs)
    b = g.CaseBatch([c["solver"] for c in batched])
    _say(f"{kind} batch of", len(specs), "cases:", [sp["N"] for sp in specs])
    n_steps = 0
    while any(c["rt"].run() for c in batched):
        co, aco = b.courant()
        dt, fac = cases.batch_time_step(batched, co, aco)
        b.scale_flow(fac)
        b.step(dt)
        for k, c in enumerate(alone):
            dk = _alone_step(c)
            assert dk == dt[k], (n_steps, k)
            ab, aa = batched[k]["solver"].alpha(), c["solver"].alpha()
            assert bits_equal(ab, aa), f"{kind} step {n_steps} case {k}: alpha off by {max_abs(ab, aa)}"
        n_steps += 1
    assert n_steps > 5
    _say(f"{kind} batch: {n_steps} steps, every case bit-identical to its handle alone")
    b.close()


def test_batch_subset_against_oracle(oracle_lib):
    """Three batched cases (rotation and deformation) against the oracle on
    the same meshes, fields and time steps."""
    import geometricvofext_amd as g
    from geometricvofext_amd import cases
    specs = cases.batch_cases(3, seed=23, nmin=10, nmax=14, tmin=0.05, tmax=0.08)
    batched = _make(specs)
    ors = []
    for c in batched:
        om = oracle_lib.OracleMesh(c["mesh"])
        o = oracle_lib.OracleSolver(om, nAlphaBounds=3, snapTol=0.0, clip=False)
        o.set_alpha(c["alpha0"])
        ors.append(o)
    b = g.CaseBatch([c["solver"] for c in batched])
    steps = 0
    while any(c["rt"].run() for c in batched):
        co, aco = b.courant()
        dt, fac = cases.batch_time_step(batched, co, aco)
        b.scale_flow(fac)
        b.step(dt)
        for k, c in enumerate(batched):
            if dt[k] == 0.0:
                continue
            ors[k].set_flow(c["phi0"] * fac[k], c["U0"] * fac[k])
            ors[k].reconstruct()
            ors[k].advect(dt[k])
            ga, oa = c["solver"].alpha(), ors[k].alpha()
            assert max_abs(ga, oa) <= 1e-12, f"step {steps} case {k}: {max_abs(ga, oa)}"
            assert bits_equal(ga, oa)
        steps += 1
    _say(f"batch vs oracle: {steps} steps of {len(specs)} cases, bit-identical")
    b.close()


def test_batch_tuner_keeps_results():
    """The launch-shape tuner restores the fields after its trials: a tuned
    batch and an untuned one give the same fields."""
    import geometricvofext_amd as g
    from geometricvofext_amd import cases
    specs = cases.batch_cases(5, seed=31, nmin=10, nmax=16, tmin=0.05, tmax=0.06)
    one, two = _make(specs), _make(specs)
    b1 = g.CaseBatch([c["solver"] for c in one])
    b2 = g.CaseBatch([c["solver"] for c in two])
    first = True
    while any(c["rt"].run() for c in one):
        co, aco = b1.courant()
        dt, fac = cases.batch_time_step(one, co, aco)
        co2, aco2 = b2.courant()
        dt2, fac2 = cases.batch_time_step(two, co2, aco2)
        assert np.array_equal(dt, dt2)
        b1.scale_flow(fac)
        b2.scale_flow(fac2)
        if first:
            shapes = b1.tune(dt)
            _say("tuned launch shapes:", shapes)
            first = False
        b1.step(dt)
        b2.step(dt2)
        for k in range(len(specs)):
            assert bits_equal(one[k]["solver"].alpha(), two[k]["solver"].alpha())
    b1.close()
    b2.close()


def test_batch_rejects_mixed_inputs():
    import geometricvofext_amd as g
    hexes = [g.SolveVofEqu(g.hex_box(6), DICT), g.SolveVofEqu(g.hex_box(7), DICT)]
    tet = g.SolveVofEqu(g.tet_box(4), DICT)
    rdf = g.SolveVofEqu(g.hex_box(6), dict(DICT, orientationMethod="RDF"))
    other = g.SolveVofEqu(g.hex_box(6), dict(DICT, nAlphaBounds=2))
    for bad, code in (([hexes[0], tet], -1), ([hexes[0], rdf], -3), ([hexes[0], hexes[0]], -1),
                      ([hexes[0], other], -1)):
        with pytest.raises(g.SvpError) as e:
            g.CaseBatch(bad)
        assert e.value.code == code
    # a handle can be read between batch steps and used alone after the batch
    b = g.CaseBatch(hexes)
    for s in hexes:
        s.set_alpha(np.full(s.mesh.n_cells, 0.5))
    b.step(np.array([0.0, 0.0]))
    assert np.all(hexes[0].alpha() == 0.5)
    b.close()
    hexes[0].reconstruct()
    assert hexes[0].n_mixed() == hexes[0].mesh.n_cells
