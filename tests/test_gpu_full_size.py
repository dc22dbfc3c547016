"""The 128^3 perturbed-hex workload at full size (2,097,152 cells): parity
with the oracle over its first steps, then size-independent properties over
a longer run on the GPU alone:

* volume conservation (clip and snap off: the update and the bounding only
  move volume between cells);
* the mixed-cell count of every reconstruct equals the host count of
  mixedCellTol < alpha < 1 - mixedCellTol on the field it starts from;
* determinism: the same run twice gives bit-identical fields (the
  look-back lists, work lists and dataflow bounding take their order from
  hardware timing, so this pins that the results do not);
* the PLIC polygon of every cell whose cut faces each hold exactly two plane
  crossings (one straight segment per face, so the section is one loop) has
  the shoelace area vector S of the cell.
"""
import numpy as np
import pytest

from tests.helpers import bits_equal, max_abs

pytestmark = pytest.mark.gpu

DICT = dict(orientationMethod="LS", nAlphaBounds=3, snapTol=0.0, clip=False, mixedCellTol=1e-8)


def _say(*a):
    print(*a, flush=True)


def _case():
    import geometricvofext_amd as g
    mesh = g.hex_box(128, perturb=0.2, seed=1)
    gpu = g.SolveVofEqu(mesh, DICT)
    gpu.init_alpha([((0.35, 0.35, 0.35), 0.15)], max_depth=5)
    a0 = gpu.alpha()
    U0, phi0 = gpu.deformation_field()
    gpu.upload_base_flow(U0, phi0)
    return mesh, gpu, a0, U0, phi0


def test_full_size_parity_first_steps(oracle_lib):
    import bench
    mesh, gpu, a0, U0, phi0 = _case()
    _say("full size:", mesh.n_cells, "cells")
    seq = bench.time_sequence(gpu, a0, 3.0, max_steps=4)[:4]
    om = oracle_lib.OracleMesh(mesh)
    ors = oracle_lib.OracleSolver(om, nAlphaBounds=3, clip=False, threads=16)
    ors.set_alpha(a0)
    gpu.set_alpha(a0)
    for k, (_, dt, uf) in enumerate(seq):
        gpu.scale_flow(uf)
        ors.set_flow(phi0 * uf, U0 * uf)
        gpu.reconstruct()
        ors.reconstruct()
        gi, oi = gpu.interface(), ors.mixed()
        assert np.array_equal(gi["cells"], oi["cells"]), f"step {k}: mixed cells differ"
        assert max_abs(gi["d"], oi["d"]) <= 1e-12 and max_abs(gi["n"], oi["n"]) <= 1e-12
        gpu.advect(dt)
        ors.advect(dt)
        ga, oa = gpu.alpha(), ors.alpha()
        assert max_abs(ga, oa) <= 1e-12, f"step {k}: alpha off by {max_abs(ga, oa)}"
        _say(f"full size step {k}: {len(gi['cells'])} mixed cells, alpha bit-exact={bits_equal(ga, oa)}")


def test_full_size_properties():
    import bench
    mesh, gpu, a0, U0, phi0 = _case()
    V = gpu.geometry()["V"]
    seq = bench.time_sequence(gpu, a0, 3.0, max_steps=150)[:150]
    v0 = float(np.sum(a0 * V))
    tol = DICT["mixedCellTol"]

    def run(check):
        gpu.set_alpha(a0)
        for k, (_, dt, uf) in enumerate(seq):
            if check and k % 25 == 0:
                a = gpu.alpha()
                expect = int(np.count_nonzero((a > tol) & (a < 1.0 - tol)))
            gpu.scale_flow(uf)
            gpu.reconstruct()
            if check and k % 25 == 0:
                assert gpu.n_mixed() == expect, f"step {k}: {gpu.n_mixed()} mixed cells, host count {expect}"
            gpu.advect(dt)
        return gpu.alpha()

    a1 = run(True)
    drift = abs(float(np.sum(a1 * V)) - v0) / v0
    _say(f"full size, {len(seq)} steps: relative volume change {drift:.2e}")
    assert drift < 1e-12
    a2 = run(False)
    assert bits_equal(a1, a2), f"two runs differ by {max_abs(a1, a2)}"
    # PLIC polygons of the last reconstruct
    gpu.reconstruct()
    ps, gi = gpu.plic_surface(), gpu.interface()
    sel = np.isin(gi["cells"], ps["cells"])
    S = gi["S"][sel]
    # cut faces per cell: 2-bit fields (status + 1) equal to 1
    fc = gi["faceclass"][sel].astype(np.uint64)
    nCutFaces = sum(((fc >> np.uint64(2 * k)) & np.uint64(3)) == 1 for k in range(8)).astype(int)
    oneLoop = gi["ncutpts"][sel] == 2 * nCutFaces
    offs, P = ps["offsets"], ps["points"]
    A = np.zeros((len(ps["cells"]), 3))
    for f in range(len(ps["cells"])):
        Q = P[offs[f]:offs[f + 1]]
        A[f] = 0.5 * np.sum(np.cross(Q, np.roll(Q, -1, axis=0)), axis=0)
    assert len(ps["cells"]) > 1000 and oneLoop.mean() > 0.99
    err = np.max(np.abs(A[oneLoop] - S[oneLoop]))
    assert err <= 1e-12 * max(1.0, float(np.max(np.abs(S))))
    _say(f"full size PLIC: {len(ps['cells'])} polygons ({int((~oneLoop).sum())} with a face crossed more than "
         f"twice), area vectors of the others match S to {err:.1e}")
