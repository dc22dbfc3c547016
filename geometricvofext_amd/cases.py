"""Seeded initial shapes and flows of the synthetic workloads (harness side).

* ``random_spheres``: spheres with seeded centres and radii (tetrahedral
  rotation case: 8 spheres, radii 0.05-0.12, overlaps unioned by the
  initialiser), inside a ball about the cube centre.
* ``random_ellipsoids``: ellipsoids with seeded centres, semi-axes and
  orientations (learned-normal case: 3 ellipsoids).
* ``rotation_axis``: a seeded unit axis through the cube centre; one full
  revolution in time T takes |omega| = 2 pi / T.
"""
from __future__ import annotations

import numpy as np


def random_rotation(rng) -> np.ndarray:
    """Uniform random rotation matrix from a unit quaternion (Shoemake)."""
    u1, u2, u3 = rng.uniform(size=3)
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def random_spheres(seed: int, n: int = 8, rmin: float = 0.05, rmax: float = 0.12, reach: float = 0.45):
    """Spheres inside the ball of radius `reach` about the cube centre, so a
    rotation about any axis through the centre keeps them off the walls."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        r = rng.uniform(rmin, rmax)
        while True:
            c = rng.uniform(-1.0, 1.0, 3) * (reach - r)
            if np.linalg.norm(c) <= reach - r:
                break
        out.append((tuple(c + 0.5), float(r)))
    return out


def random_ellipsoids(seed: int, n: int = 3, amin: float = 0.08, amax: float = 0.15):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = rng.uniform(amin, amax, 3)
        c = rng.uniform(0.3, 0.7, 3)
        out.append((c, a, random_rotation(rng)))
    return out


def rotation_axis(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def rotation_omega(seed: int, period: float) -> np.ndarray:
    """Angular velocity of one revolution about the seeded axis in `period`."""
    return rotation_axis(seed) * (2.0 * np.pi / period)


def batch_cases(n: int = 256, seed: int = 2025, nmin: int = 16, nmax: int = 48, tmin: float = 0.5,
                tmax: float = 1.0):
    """The batched workload: n independent cases, each a uniform hex mesh of
    N^3 cells with N ~ U{nmin..nmax}; 1-4 ellipsoids (semi-axes U(0.05, 0.2),
    uniformly random orientation) centred inside the ball of radius
    0.45 - (largest semi-axis) about the cube centre; either solid-body
    rotation (one revolution about a seeded axis through the cube centre in
    the case's end time T) or the deformation field (period 2T, so the shape
    returns at T); T ~ U(tmin, tmax); Courant number U(0.3, 0.5)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        N = int(rng.integers(nmin, nmax + 1))
        shapes = []
        for _ in range(int(rng.integers(1, 5))):
            a = rng.uniform(0.05, 0.2, 3)
            reach = 0.45 - float(a.max())
            while True:
                c = rng.uniform(-1.0, 1.0, 3) * reach
                if np.linalg.norm(c) <= reach:
                    break
            shapes.append((c + 0.5, a, random_rotation(rng)))
        flow = "rotation" if rng.integers(2) == 0 else "deformation"
        T = float(rng.uniform(tmin, tmax))
        v = rng.normal(size=3)
        out.append(dict(N=N, shapes=shapes, flow=flow, end_time=T, co=float(rng.uniform(0.3, 0.5)),
                        axis=v / np.linalg.norm(v)))
    return out


def make_batch_case(spec, alpha_dict):
    """Handle, initial alpha, base flow and time control of one batched case
    (the harness of plicVofAdvectionFoam per case: RunTime with the case's
    Courant limit and end time; the flow factor's period 2T for the
    deformation, none for the rotation)."""
    import geometricvofext_amd as g
    from geometricvofext_amd import harness
    # (optional spec keys: "kind" "tet" for N^3 x 6 Kuhn tetrahedra, and
    # "perturb" / "seed" for displaced interior vertices)
    if spec.get("kind", "hex") == "tet":
        mesh = g.tet_box(spec["N"], perturb=spec.get("perturb", 0.0), seed=spec.get("seed", 0))
    else:
        mesh = g.hex_box(spec["N"], perturb=spec.get("perturb", 0.0), seed=spec.get("seed", 0))
    s = g.SolveVofEqu(mesh, dict(alpha_dict))
    s.init_alpha(spec["shapes"], max_depth=4)
    a0 = s.alpha()
    if spec["flow"] == "rotation":
        U0, phi0 = s.rotation_field(spec["axis"] * (2.0 * np.pi / spec["end_time"]), (0.5, 0.5, 0.5))
        period = 0.0
    else:
        U0, phi0 = s.deformation_field()
        period = 2.0 * spec["end_time"]
    s.upload_base_flow(U0, phi0)
    s.scale_flow(0.0)
    rt = harness.RunTime(end_time=spec["end_time"], max_co=spec["co"], max_alpha_co=spec["co"])
    return dict(solver=s, mesh=mesh, alpha0=a0, U0=U0, phi0=phi0, period=period, rt=rt)


def batch_time_step(cases, co, aco):
    """Advance every unfinished case's time control by one step from its
    Courant sums (setDeltaT.H, Time::operator++): returns (dt per case, 0 for
    a finished case, and the flow factor per case)."""
    from geometricvofext_amd import harness
    dt = np.zeros(len(cases))
    fac = np.zeros(len(cases))
    for k, c in enumerate(cases):
        rt = c["rt"]
        if not rt.run():
            fac[k] = c.get("factor", 0.0)
            continue
        rt.set_delta_t(0.5 * co[k] * rt.delta_t, 0.5 * aco[k] * rt.delta_t)
        rt.increment()
        dt[k] = rt.delta_t
        fac[k] = c["factor"] = harness.u_factor(rt.value, rt.delta_t, c["period"])
    return dt, fac
