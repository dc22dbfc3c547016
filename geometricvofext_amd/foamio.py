"""OpenFOAM field IO for the harness: binary ``volScalarField`` reader/writer and
the cell permutation of ``renumberMesh`` on a single-block hex mesh.

The reference's exact solutions (``tutorials/test/exactSolutions/<t>/
alpha.water.exact``) are binary fields with header ``arch
"LSB;label=32;scalar=64"`` stored in the cell order that ``renumberMesh
-overwrite`` produced (``tutorials/test/plicVofAdvectionFoam/Allrun:10``).
"""
from __future__ import annotations

import re
from collections import deque

import numpy as np


def read_internal_field(path: str) -> np.ndarray:
    """Internal field of an OpenFOAM volScalarField (binary or ascii, uniform or not)."""
    with open(path, "rb") as fh:
        raw = fh.read()
    m = re.search(rb"internalField\s+nonuniform\s+List<scalar>\s+(\d+)\s*\(", raw)
    if m:
        n = int(m.group(1))
        start = m.end()
        header = raw[:m.start()]
        if b"format      binary" in header or b"binary" in header.split(b"format", 1)[-1][:32]:
            return np.frombuffer(raw, dtype="<f8", count=n, offset=start).copy()
        body = raw[start:].split(b")", 1)[0]
        return np.array(body.split(), dtype=np.float64)
    m = re.search(rb"internalField\s+uniform\s+([-+0-9.eE]+)\s*;", raw)
    if m:
        raise ValueError("uniform internalField has no size; give it to the caller")
    raise ValueError(f"no internalField in {path}")


def write_internal_field(path: str, values: np.ndarray, name: str = "alpha.water") -> None:
    """Write a binary OpenFOAM volScalarField with zeroGradient walls."""
    values = np.ascontiguousarray(values, dtype="<f8")
    head = (
        "FoamFile\n{\n    version     2.0;\n    format      binary;\n"
        '    arch        "LSB;label=32;scalar=64";\n    class       volScalarField;\n'
        f"    object      {name};\n}}\n\ndimensions      [0 0 0 0 0 0 0];\n\n"
        f"internalField   nonuniform List<scalar> {values.size}\n("
    )
    tail = ")\n;\n\nboundaryField\n{\n    \".*\"\n    {\n        type            zeroGradient;\n    }\n}\n"
    with open(path, "wb") as fh:
        fh.write(head.encode())
        fh.write(values.tobytes())
        fh.write(tail.encode())


def renumber_order(N: int) -> np.ndarray:
    """``order[r]`` = natural (i + N j + N^2 k) index of renumbered cell r.

    Forward Cuthill-McKee from cell 0: a queue of cells; each visited cell
    enqueues its unvisited face neighbours (ascending index) stable-sorted by
    their number of face neighbours, the scheme of OpenFOAM's renumberMesh
    default.  ``natural[order] = renumbered``.
    """
    n = N ** 3
    idx = np.arange(n)
    i, j, k = idx % N, (idx // N) % N, idx // (N * N)
    deg = (i > 0).astype(np.int64) + (i < N - 1) + (j > 0) + (j < N - 1) + (k > 0) + (k < N - 1)
    visited = np.zeros(n, dtype=bool)
    order = np.empty(n, dtype=np.int64)
    q = deque([0])
    r = 0
    steps = (-N * N, -N, -1, 1, N, N * N)
    while q:
        c = q.popleft()
        if visited[c]:
            continue
        visited[c] = True
        order[r] = c
        r += 1
        ci, cj, ck = c % N, (c // N) % N, c // (N * N)
        ok = (ck > 0, cj > 0, ci > 0, ci < N - 1, cj < N - 1, ck < N - 1)
        nb = [c + s for s, o in zip(steps, ok) if o and not visited[c + s]]
        nb.sort(key=lambda x: deg[x])
        q.extend(nb)
    assert r == n
    return order


def to_natural(renumbered: np.ndarray, N: int, order: np.ndarray | None = None) -> np.ndarray:
    if order is None:
        order = renumber_order(N)
    out = np.empty_like(renumbered)
    out[order] = renumbered
    return out
