"""Batched cases on the GPU: one handle over the disjoint union of several
meshes, each case with its own time step, against separate handles per case
and the oracle.  Bar: bit-identical alpha, interface data and face volumes;
an idle case (step 0) stays untouched."""
import numpy as np
import pytest

from tests.helpers import bits_equal, ramp_alpha

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="the union-mesh batch path this test exercised stalled without an error on "
                                      "its first GPU run and the cause was not found; that path is removed "
                                      "(svp_batch_define no longer exists).  Batches now run as case-major "
                                      "launches over separate handles: tests/test_gpu_batch_cases.py")]


def _cases():
    import geometricvofext_amd as g
    meshes = [g.hex_box(10), g.hex_box(8, perturb=0.2, seed=5), g.tet_box(6, perturb=0.1, seed=2), g.hex_box(7)]
    shapes = [[((0.45, 0.5, 0.5), 0.25)], [((0.5, 0.45, 0.55), 0.3)], [((0.5, 0.5, 0.45), 0.28)],
              [((0.5, 0.5, 0.5), 0.3)]]
    return meshes, shapes


def test_batch_matches_separate_handles(oracle_lib):
    import geometricvofext_amd as g
    from geometricvofext_amd.mesh import union_mesh
    meshes, shapes = _cases()
    u, off = union_mesh(meshes)
    d = dict(orientationMethod="LS", nAlphaBounds=3, snapTol=0.0, clip=False)
    b = g.SolveVofEqu(u, d)
    b.batch_define(off)
    sep = [g.SolveVofEqu(m, d) for m in meshes]
    a0 = []
    for k, m in enumerate(meshes):
        om = oracle_lib.OracleMesh(m)
        a0.append(ramp_alpha(om.C, shapes[k], 1.5 / 8))
    b.set_alpha(np.concatenate(a0))
    for k, s in enumerate(sep):
        s.set_alpha(a0[k])
    U, phi = b.deformation_field()
    b.upload_base_flow(U, phi)
    for k, s in enumerate(sep):
        Uk, phik = s.deformation_field()
        s.upload_base_flow(Uk, phik)
    rng = np.random.default_rng(1)
    for step in range(6):
        factors = rng.uniform(0.5, 1.0, len(meshes))
        dts = rng.uniform(0.004, 0.012, len(meshes))
        dts[3] = 0.0 if step % 2 == 0 else dts[3]    # case 3 idle every other step
        b.scale_flow_batch(factors)
        co, aco = b.courant_batch()
        b.batch_set_dt(dts)
        before = b.alpha()
        b.reconstruct()
        b.advect(0.0)
        ab = b.alpha()
        bi = b.interface()
        for k, s in enumerate(sep):
            sl = slice(off[k], off[k + 1])
            if dts[k] == 0.0:
                assert bits_equal(ab[sl], before[sl]), (step, k)
                assert not np.any((bi["cells"] >= off[k]) & (bi["cells"] < off[k + 1]))
                continue
            s.scale_flow(factors[k])
            c1, c2 = s.courant()
            assert c1 == co[k] and c2 == aco[k]
            s.reconstruct()
            s.advect(dts[k])
            assert bits_equal(ab[sl], s.alpha()), (step, k)
            si = s.interface()
            sel = (bi["cells"] >= off[k]) & (bi["cells"] < off[k + 1])
            assert np.array_equal(bi["cells"][sel] - off[k], si["cells"])
            assert bits_equal(bi["d"][sel], si["d"]) and bits_equal(bi["n"][sel], si["n"])
        # face volumes of active cases
        ap = b.alpha_phi()
        for k, s in enumerate(sep):
            if dts[k] == 0.0:
                continue
            m = meshes[k]
            nif_off = sum(mm.n_internal_faces for mm in meshes[:k])
            assert bits_equal(ap[nif_off:nif_off + m.n_internal_faces], s.alpha_phi()[:m.n_internal_faces])
    print("batch of", len(meshes), "cases bit-identical to separate handles over 6 steps")
