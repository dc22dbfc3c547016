"""Polyhedral mesh generators in OpenFOAM polyMesh layout (the blockMesh step of
the reference's test cases).

A mesh is the raw OpenFOAM face-addressed description:

* ``points``        (n_points, 3) float64
* ``face_offsets``  (n_faces + 1,) int32 CSR offsets into ``face_points``
* ``face_points``   int32 point labels, each face ordered so that its right-hand
  normal points out of its owner cell
* ``owner``         (n_faces,) int32
* ``neighbour``     (n_internal_faces,) int32, owner < neighbour
* patches: (name, start, size, type) with boundary faces stored after the
  internal faces, patch by patch.

Internal faces are in upper-triangular order (sorted by owner, then by
neighbour) as OpenFOAM's ``polyMesh`` constructor from cell shapes produces
them; the face stored for an internal face is the owner's face as the hex cell
model defines it.  Point, cell and boundary-face orders follow blockMesh for a
single block (``tutorials/test/plicVofAdvectionFoam/system/blockMeshDict:17-94``):
points and cells i-fastest, patches in the order top, left, back, right,
bottom, front.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

PATCH_WALL = 0
PATCH_EMPTY = 1


@dataclass
class PolyMesh:
    points: np.ndarray
    face_offsets: np.ndarray
    face_points: np.ndarray
    owner: np.ndarray
    neighbour: np.ndarray
    n_cells: int
    patch_names: list = field(default_factory=list)
    patch_start: np.ndarray = None
    patch_size: np.ndarray = None
    patch_type: np.ndarray = None
    name: str = ""

    @property
    def n_points(self) -> int:
        return int(self.points.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def n_internal_faces(self) -> int:
        return int(self.neighbour.shape[0])

    def face(self, f: int) -> np.ndarray:
        return self.face_points[self.face_offsets[f]:self.face_offsets[f + 1]]

    def check(self) -> None:
        """Structural invariants of a polyMesh (cheap, vectorised)."""
        nf, nif = self.n_faces, self.n_internal_faces
        assert self.face_offsets.shape == (nf + 1,)
        assert self.face_offsets[0] == 0 and self.face_offsets[-1] == self.face_points.size
        assert np.all(self.owner[:nif] < self.neighbour)
        assert self.owner.min() >= 0 and self.owner.max() < self.n_cells
        assert int(self.patch_start[0]) == nif if len(self.patch_start) else True
        assert int(self.patch_start[-1] + self.patch_size[-1]) == nf
        # upper-triangular order of internal faces
        key = self.owner[:nif].astype(np.int64) * self.n_cells + self.neighbour
        assert np.all(np.diff(key) > 0)


# OpenFOAM hex cell model (vertex v0..v7 = (i,j,k),(i+1,j,k),(i+1,j+1,k),(i,j+1,k), same at k+1)
_HEX_FACES = {
    "xmin": (0, 4, 7, 3),
    "xmax": (1, 2, 6, 5),
    "ymin": (0, 1, 5, 4),
    "ymax": (3, 7, 6, 2),
    "zmin": (0, 3, 2, 1),
    "zmax": (4, 5, 6, 7),
}


def _hex_vertices(N, pidx):
    """(N^3, 8) point labels of each hex cell, cells i-fastest."""
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    i = i.transpose(2, 1, 0).ravel()
    j = j.transpose(2, 1, 0).ravel()
    k = k.transpose(2, 1, 0).ravel()
    v = np.empty((N ** 3, 8), dtype=np.int64)
    v[:, 0] = pidx(i, j, k)
    v[:, 1] = pidx(i + 1, j, k)
    v[:, 2] = pidx(i + 1, j + 1, k)
    v[:, 3] = pidx(i, j + 1, k)
    v[:, 4] = pidx(i, j, k + 1)
    v[:, 5] = pidx(i + 1, j, k + 1)
    v[:, 6] = pidx(i + 1, j + 1, k + 1)
    v[:, 7] = pidx(i, j + 1, k + 1)
    return v, i, j, k


def hex_box(N: int, perturb: float = 0.0, seed: int = 0, length: float = 1.0) -> PolyMesh:
    """Uniform N^3 hex mesh of the cube [0, length]^3, all six patches walls.

    ``perturb`` > 0 displaces every interior vertex by U(-perturb*h, perturb*h)
    per axis from ``numpy.random.default_rng(seed)``, which makes hex faces
    non-planar (warped).
    """
    h = length / N
    n1 = N + 1
    g = np.arange(n1, dtype=np.float64) * h
    g[-1] = length
    pk, pj, pi = np.meshgrid(g, g, g, indexing="ij")
    points = np.stack([pi.ravel(), pj.ravel(), pk.ravel()], axis=1)
    if perturb > 0.0:
        rng = np.random.default_rng(seed)
        disp = rng.uniform(-perturb * h, perturb * h, size=points.shape)
        ii = np.arange(n1)
        interior1 = (ii > 0) & (ii < N)
        m = (interior1[None, None, :] & interior1[None, :, None] & interior1[:, None, None]).ravel()
        points[m] += disp[m]

    def pidx(i, j, k):
        return i + n1 * (j + n1 * k)

    v, ci, cj, ck = _hex_vertices(N, pidx)
    nc = N ** 3
    cell = np.arange(nc, dtype=np.int64)

    # internal faces, upper-triangular: per owner, neighbours c+1 < c+N < c+N^2
    nb = np.stack([cell + 1, cell + N, cell + N * N], axis=1)
    valid = np.stack([ci < N - 1, cj < N - 1, ck < N - 1], axis=1)
    loc = [_HEX_FACES["xmax"], _HEX_FACES["ymax"], _HEX_FACES["zmax"]]
    fverts = np.stack([v[:, list(l)] for l in loc], axis=1)  # (nc, 3, 4)
    own_i = np.broadcast_to(cell[:, None], (nc, 3))[valid]
    nei_i = nb[valid]
    fv_i = fverts[valid]

    # boundary patches: order as blockMeshDict (top, left, back, right, bottom, front)
    def patch(face_key, sel, order_keys):
        cells = cell[sel]
        keys = [a[sel] for a in order_keys]
        o = np.lexsort(keys[::-1]) if keys else np.arange(cells.size)
        cells = cells[o]
        return cells, v[cells][:, list(_HEX_FACES[face_key])]

    patches = [
        ("top", "zmax", ck == N - 1, [ci, cj]),      # i outer, j inner
        ("left", "ymin", cj == 0, [ci, ck]),        # i outer, k inner
        ("back", "xmin", ci == 0, [ck, cj]),        # k outer, j inner
        ("right", "ymax", cj == N - 1, [ci, ck]),
        ("bottom", "zmin", ck == 0, [ci, cj]),
        ("front", "xmax", ci == N - 1, [ck, cj]),
    ]
    b_own, b_fv, names, starts, sizes = [], [], [], [], []
    start = own_i.size
    for name, key, sel, okeys in patches:
        c, fv = patch(key, sel, okeys)
        b_own.append(c)
        b_fv.append(fv)
        names.append(name)
        starts.append(start)
        sizes.append(c.size)
        start += c.size
    owner = np.concatenate([own_i] + b_own).astype(np.int32)
    fv_all = np.concatenate([fv_i] + b_fv).astype(np.int32)
    nf = owner.size
    return PolyMesh(
        points=np.ascontiguousarray(points),
        face_offsets=(np.arange(nf + 1, dtype=np.int64) * 4).astype(np.int32),
        face_points=np.ascontiguousarray(fv_all.ravel()),
        owner=owner,
        neighbour=nei_i.astype(np.int32),
        n_cells=nc,
        patch_names=names,
        patch_start=np.array(starts, dtype=np.int32),
        patch_size=np.array(sizes, dtype=np.int32),
        patch_type=np.full(len(names), PATCH_WALL, dtype=np.int32),
        name=f"hex{N}" + (f"_p{perturb}_s{seed}" if perturb > 0 else ""),
    )


# Kuhn (Freudenthal) split of the unit cube into 6 tetrahedra sharing the
# main diagonal v0-v6; every cube uses the same diagonal, so the split is
# conforming across cubes.  Local vertex numbering as the hex model above.
_KUHN_TETS = np.array([
    [0, 1, 2, 6],   # x, y, z
    [0, 1, 5, 6],   # x, z, y
    [0, 3, 2, 6],   # y, x, z
    [0, 3, 7, 6],   # y, z, x
    [0, 4, 5, 6],   # z, x, y
    [0, 4, 7, 6],   # z, y, x
])


def tet_box(N: int, perturb: float = 0.0, seed: int = 0, length: float = 1.0) -> PolyMesh:
    """N^3 cubes each split into 6 Kuhn tetrahedra (6 N^3 cells), six wall patches.

    Cells are cube-major (cube index i-fastest), then the six tets of
    ``_KUHN_TETS``.  A triangle face starts at its lowest local tet vertex and
    is oriented outward from its owner.  Interior vertices are displaced by
    U(-perturb*h, perturb*h) per axis when ``perturb`` > 0.
    """
    base = hex_box(N, perturb=perturb, seed=seed, length=length)
    points = base.points
    n1 = N + 1

    def pidx(i, j, k):
        return i + n1 * (j + n1 * k)

    v, ci, cj, ck = _hex_vertices(N, pidx)
    tv = v[:, _KUHN_TETS]                      # (N^3, 6, 4)
    tv = tv.reshape(-1, 4)                     # cell = 6*cube + t
    nc = tv.shape[0]
    # fix orientation so that (b-a).((c-a)x(d-a)) > 0 for every tet
    P = points
    a, b, c, d = (P[tv[:, q]] for q in range(4))
    vol6 = np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a))
    flip = vol6 < 0
    tv[flip, 2], tv[flip, 3] = tv[flip, 3].copy(), tv[flip, 2].copy()
    # faces opposite local vertex q, outward for a positively oriented tet
    # (a,b,c,d): opposite a -> (b,c,d); opposite b -> (a,d,c); opposite c -> (a,b,d); opposite d -> (a,c,b)
    loc = np.array([[1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]])
    fv = tv[:, loc]                            # (nc, 4, 3)
    fv = fv.reshape(-1, 3)
    fcell = np.repeat(np.arange(nc, dtype=np.int64), 4)
    key = np.sort(fv, axis=1)
    kk = (key[:, 0] * (points.shape[0]) + key[:, 1]) * points.shape[0] + key[:, 2]
    order = np.argsort(kk, kind="stable")
    kk_s = kk[order]
    same = np.zeros(kk_s.size, dtype=bool)
    same[1:] = kk_s[1:] == kk_s[:-1]
    # pairs: positions p where same[p] -> (order[p-1], order[p])
    p = np.nonzero(same)[0]
    fa, fb = order[p - 1], order[p]
    ca, cb = fcell[fa], fcell[fb]
    own_f = np.where(ca < cb, fa, fb)
    owners = fcell[own_f]
    neis = np.where(ca < cb, cb, ca)
    o = np.lexsort((neis, owners))
    own_f, owners, neis = own_f[o], owners[o], neis[o]
    internal_fv = fv[own_f]
    # boundary faces: unpaired
    paired = np.zeros(kk_s.size, dtype=bool)
    paired[p] = True
    paired[p - 1] = True
    bfaces = np.sort(order[~paired])
    bcell = fcell[bfaces]
    bfv = fv[bfaces]
    # classify boundary faces by cube side using the unperturbed lattice coords
    n1 = N + 1
    pi_ = bfv % n1
    pj_ = (bfv // n1) % n1
    pk_ = bfv // (n1 * n1)
    sides = [
        ("top", np.all(pk_ == N, axis=1)),
        ("left", np.all(pj_ == 0, axis=1)),
        ("back", np.all(pi_ == 0, axis=1)),
        ("right", np.all(pj_ == N, axis=1)),
        ("bottom", np.all(pk_ == 0, axis=1)),
        ("front", np.all(pi_ == N, axis=1)),
    ]
    taken = np.zeros(bfaces.size, dtype=bool)
    b_own, b_fv, names, starts, sizes = [], [], [], [], []
    start = owners.size
    for name, sel in sides:
        sel = sel & ~taken
        taken |= sel
        idx = np.nonzero(sel)[0]
        idx = idx[np.argsort(bcell[idx], kind="stable")]
        b_own.append(bcell[idx])
        b_fv.append(bfv[idx])
        names.append(name)
        starts.append(start)
        sizes.append(idx.size)
        start += idx.size
    assert taken.all()
    owner = np.concatenate([owners] + b_own).astype(np.int32)
    fv_all = np.concatenate([internal_fv] + b_fv).astype(np.int32)
    nf = owner.size
    return PolyMesh(
        points=points,
        face_offsets=(np.arange(nf + 1, dtype=np.int64) * 3).astype(np.int32),
        face_points=np.ascontiguousarray(fv_all.ravel()),
        owner=owner,
        neighbour=neis.astype(np.int32),
        n_cells=nc,
        patch_names=names,
        patch_start=np.array(starts, dtype=np.int32),
        patch_size=np.array(sizes, dtype=np.int32),
        patch_type=np.full(len(names), PATCH_WALL, dtype=np.int32),
        name=f"tet{N}" + (f"_p{perturb}_s{seed}" if perturb > 0 else ""),
    )


def agglomerated_box(N: int, group: np.ndarray, perturb: float = 0.0, seed: int = 0,
                     name: str = "agglomerated") -> PolyMesh:
    """Polyhedral mesh of the unit cube from an N^3 hex lattice whose cells are
    merged into boxes: ``group`` (N^3, lattice cells i-fastest) names the cell
    each lattice cell belongs to, and each group must be an axis-aligned box of
    lattice cells.  The lattice faces between two cells (or between a cell
    and the domain boundary) merge into one polygon, the loop of lattice
    edges around them, so a face of a large cell next to smaller cells is
    split as the small cells' faces are and keeps the edge midpoints they
    share: the layout of a 2:1 hexRef8 refinement.  Cells take labels in the
    order of their first lattice cell; internal faces are upper-triangular;
    patches as ``hex_box``; unused lattice points (centres of merged faces,
    cell interiors) are dropped.  ``perturb`` displaces interior lattice
    points as in ``hex_box`` (faces become warped)."""
    base = hex_box(N, perturb=perturb, seed=seed)
    n1 = N + 1
    group = np.asarray(group, dtype=np.int64).reshape(-1)
    assert group.size == N ** 3
    # cell labels by first lattice cell
    _, first = np.unique(group, return_index=True)
    order = np.argsort(first)
    label_of = np.empty(group.max() + 1, dtype=np.int64)
    label_of[np.unique(group)[order]] = np.arange(order.size)
    cell = label_of[group]
    nc = int(order.size)

    def pidx(i, j, k):
        return i + n1 * (j + n1 * k)

    # lattice faces: (cell below/outside, cell above/outside, 4 points with
    # the right-hand normal along +axis)
    faces = []   # (owner cell, neighbour cell or -1 - patch, point loop) per lattice face
    cid = lambda i, j, k: cell[i + N * (j + N * k)]
    rng_ = range(N)
    # the +axis quad of lattice face at (i,j,k) on the low side of cell (i,j,k) along the axis
    def quad(axis, i, j, k):
        if axis == 0:
            return [pidx(i, j, k), pidx(i, j + 1, k), pidx(i, j + 1, k + 1), pidx(i, j, k + 1)]
        if axis == 1:
            return [pidx(i, j, k), pidx(i, j, k + 1), pidx(i + 1, j, k + 1), pidx(i + 1, j, k)]
        return [pidx(i, j, k), pidx(i + 1, j, k), pidx(i + 1, j + 1, k), pidx(i, j + 1, k)]
    patch_of = {(2, 1): 0, (1, 0): 1, (0, 0): 2, (1, 1): 3, (2, 0): 4, (0, 1): 5}   # (axis, high side)
    merged = {}
    for axis in range(3):
        for k in range(N + (axis == 2)):
            for j in range(N + (axis == 1)):
                for i in range(N + (axis == 0)):
                    lo = [i, j, k]
                    lo[axis] -= 1
                    a = cid(*lo) if lo[axis] >= 0 else None
                    b = cid(i, j, k) if [i, j, k][axis] < N else None
                    q = quad(axis, i, j, k)        # normal +axis: from a to b
                    if a is not None and b is not None:
                        if a == b:
                            continue
                        key = (min(a, b), max(a, b))
                        loop = q if a < b else q[::-1]
                    elif b is None:
                        key = (a, -1 - patch_of[(axis, 1)])
                        loop = q
                    else:
                        key = (b, -1 - patch_of[(axis, 0)])
                        loop = q[::-1]
                    merged.setdefault(key, []).append(loop)
    # one polygon per key: the directed edges left after opposite pairs cancel
    polys = {}
    for key, loops in merged.items():
        edges = {}
        for L in loops:
            for t in range(4):
                u, v = L[t], L[(t + 1) % 4]
                if (v, u) in edges:
                    del edges[(v, u)]
                else:
                    edges[(u, v)] = True
        nxt = {u: v for (u, v) in edges}
        assert len(nxt) == len(edges), "merged faces must form one simple loop"
        start = min(nxt)
        poly = [start]
        while nxt[poly[-1]] != start:
            poly.append(nxt[poly[-1]])
        assert len(poly) == len(edges), "merged faces must form one loop"
        polys[key] = poly
    internal = sorted(k for k in polys if k[1] >= 0)
    bnd = sorted((k for k in polys if k[1] < 0), key=lambda k: (-1 - k[1], k[0]))
    fl = [polys[k] for k in internal] + [polys[k] for k in bnd]
    owner = np.array([k[0] for k in internal] + [k[0] for k in bnd], dtype=np.int64)
    neighbour = np.array([k[1] for k in internal], dtype=np.int64)
    used = np.unique(np.concatenate([np.asarray(f) for f in fl]))
    relabel = np.full(base.points.shape[0], -1, dtype=np.int64)
    relabel[used] = np.arange(used.size)
    fpts = relabel[np.concatenate([np.asarray(f) for f in fl])]
    offs = np.zeros(len(fl) + 1, dtype=np.int64)
    np.cumsum([len(f) for f in fl], out=offs[1:])
    names = ["top", "left", "back", "right", "bottom", "front"]
    sizes = [sum(1 for k in bnd if -1 - k[1] == p) for p in range(6)]
    starts = list(len(internal) + np.concatenate([[0], np.cumsum(sizes)[:-1]]))
    m = PolyMesh(
        points=np.ascontiguousarray(base.points[used]),
        face_offsets=offs.astype(np.int32),
        face_points=fpts.astype(np.int32),
        owner=owner.astype(np.int32),
        neighbour=neighbour.astype(np.int32),
        n_cells=nc,
        patch_names=names,
        patch_start=np.array(starts, dtype=np.int32),
        patch_size=np.array(sizes, dtype=np.int32),
        patch_type=np.full(6, PATCH_WALL, dtype=np.int32),
        name=name + (f"_p{perturb}_s{seed}" if perturb > 0 else ""),
    )
    m.check()
    return m


def refined_hex_box(N: int, lo=(0.25, 0.25, 0.25), hi=(0.75, 0.75, 0.75), perturb: float = 0.0,
                    seed: int = 0) -> PolyMesh:
    """N^3 hex mesh (N even in each refined direction's lattice) whose cells
    with centres inside the box [lo, hi) are refined 2x2x2, as one hexRef8
    level: cells next to the refined region keep the edge midpoints and have
    their shared faces split in four (up to 24 faces and 20 points)."""
    Nf = 2 * N
    i, j, k = np.meshgrid(np.arange(Nf), np.arange(Nf), np.arange(Nf), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))
    ci, cj, ck = i // 2, j // 2, k // 2
    cc = np.stack([(ci + 0.5) / N, (cj + 0.5) / N, (ck + 0.5) / N], axis=1)
    refine = np.all((cc >= np.asarray(lo)) & (cc < np.asarray(hi)), axis=1)
    coarse_id = ci + N * (cj + N * ck)
    fine_id = N ** 3 + i + Nf * (j + Nf * k)
    group = np.where(refine, fine_id, coarse_id)
    return agglomerated_box(Nf, group, perturb=perturb, seed=seed, name=f"refined{N}")


def brick_box(N: int, seed: int = 0, perturb: float = 0.0) -> PolyMesh:
    """Polyhedral mesh of mixed boxes: the (2N)^3 lattice in 2x2x2 blocks,
    each block (seeded) one cell, two 2x2x1 bricks along a random axis, or
    eight lattice cells.  Cells have 6 to 16 faces."""
    Nf = 2 * N
    rng = np.random.default_rng(seed)
    kind = rng.integers(0, 3, N ** 3)
    axis = rng.integers(0, 3, N ** 3)
    i, j, k = np.meshgrid(np.arange(Nf), np.arange(Nf), np.arange(Nf), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))
    b = (i // 2) + N * ((j // 2) + N * (k // 2))
    half = np.choose(axis[b], [i % 2, j % 2, k % 2])
    fine = i + Nf * (j + Nf * k)
    group = np.where(kind[b] == 0, 8 * b, np.where(kind[b] == 1, 8 * b + 1 + half, 8 * N ** 3 + fine))
    return agglomerated_box(Nf, group, perturb=perturb, seed=seed, name=f"bricks{N}_s{seed}")


def union_mesh(meshes, name: str = "batch") -> tuple[PolyMesh, np.ndarray]:
    """Disjoint union of independent meshes as one polyMesh, for running many
    cases in the same launches.  Internal faces come case by case (each case's
    upper-triangular order, shifted), then every case's boundary patches.
    Within every cell the relative order of its faces, and every stencil and
    point-source order, stay those of the case on its own, so each case's
    results are bit-identical to a separate run.  Returns the mesh and the
    cell offsets of the cases (n_cases + 1)."""
    pt_off = np.cumsum([0] + [m.n_points for m in meshes])
    c_off = np.cumsum([0] + [m.n_cells for m in meshes])
    nif_tot = sum(m.n_internal_faces for m in meshes)
    pts = np.concatenate([m.points for m in meshes])
    fl_int, fl_bnd, own_int, own_bnd, nei = [], [], [], [], []
    names, starts, sizes, types = [], [], [], []
    start = nif_tot
    for k, m in enumerate(meshes):
        nif = m.n_internal_faces
        lens = np.diff(m.face_offsets)
        fp = m.face_points.astype(np.int64) + pt_off[k]
        split = m.face_offsets[nif]
        fl_int.append((lens[:nif], fp[:split]))
        own_int.append(m.owner[:nif].astype(np.int64) + c_off[k])
        nei.append(m.neighbour.astype(np.int64) + c_off[k])
        fl_bnd.append((lens[nif:], fp[split:]))
        own_bnd.append(m.owner[nif:].astype(np.int64) + c_off[k])
        for p, pname in enumerate(m.patch_names):
            names.append(f"case{k}_{pname}")
            starts.append(start + int(m.patch_start[p]) - nif)
            sizes.append(int(m.patch_size[p]))
            types.append(int(m.patch_type[p]))
        start += m.n_faces - nif
    lens = np.concatenate([l for l, _ in fl_int] + [l for l, _ in fl_bnd])
    fpts = np.concatenate([f for _, f in fl_int] + [f for _, f in fl_bnd]).astype(np.int32)
    owner = np.concatenate(own_int + own_bnd).astype(np.int32)
    offs = np.zeros(lens.size + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    u = PolyMesh(
        points=pts,
        face_offsets=offs.astype(np.int32),
        face_points=fpts,
        owner=owner,
        neighbour=np.concatenate(nei).astype(np.int32),
        n_cells=int(c_off[-1]),
        patch_names=names,
        patch_start=np.array(starts, dtype=np.int32),
        patch_size=np.array(sizes, dtype=np.int32),
        patch_type=np.array(types, dtype=np.int32),
        name=name,
    )
    return u, c_off.astype(np.int32)
