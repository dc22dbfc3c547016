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
