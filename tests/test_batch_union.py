"""The batched cases run as one disjoint-union mesh.  The oracle on the union
(common time step) must reproduce every case's separate run bit for bit: the
union keeps each cell's face order, the stencils and the point sources."""
import numpy as np

from tests.helpers import bits_equal, ramp_alpha


def test_union_matches_separate_runs(oracle_lib):
    from geometricvofext_amd.mesh import hex_box, union_mesh
    meshes = [hex_box(7), hex_box(5, perturb=0.2, seed=3), hex_box(6)]
    centres = [(0.45, 0.5, 0.5), (0.5, 0.45, 0.55), (0.52, 0.5, 0.48)]
    u, off = union_mesh(meshes)
    u.check()
    ou = oracle_lib.OracleMesh(u)
    a0 = np.concatenate([ramp_alpha(oracle_lib.OracleMesh(m).C, [(c, 0.27)], 1.5 / 6)
                         for m, c in zip(meshes, centres)])
    U, phi = ou.deformation_field()
    su = oracle_lib.OracleSolver(ou, nAlphaBounds=3, clip=False)
    su.set_alpha(a0)
    su.set_flow(phi, U)
    sep = []
    for k, m in enumerate(meshes):
        om = oracle_lib.OracleMesh(m)
        s = oracle_lib.OracleSolver(om, nAlphaBounds=3, clip=False)
        s.set_alpha(a0[off[k]:off[k + 1]])
        Uk, phik = om.deformation_field()
        # the union's flux restricted to the case equals the case's own
        assert bits_equal(Uk, U[off[k]:off[k + 1]])
        s.set_flow(phik, Uk)
        sep.append(s)
    dt = 0.01
    for step in range(4):
        su.reconstruct()
        su.advect(dt)
        au = su.alpha()
        mu = su.mixed()
        for k, s in enumerate(sep):
            s.reconstruct()
            s.advect(dt)
            assert bits_equal(au[off[k]:off[k + 1]], s.alpha()), (step, k)
            mk = s.mixed()
            sel = (mu["cells"] >= off[k]) & (mu["cells"] < off[k + 1])
            assert np.array_equal(mu["cells"][sel] - off[k], mk["cells"])
            assert bits_equal(mu["d"][sel], mk["d"])
