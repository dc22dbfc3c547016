"""CPU restatement of the learned-normal pieces (numpy, FP64).

Test infrastructure only: imported by tests/ as the checker of the HIP
kernels in geometricvofext_amd/csrc/learned_kernels.hip; never imported by
the product package.  The learned normal has no counterpart in the
reference, so these functions restate this library's own design:

* ``cube_fraction``: liquid fraction of the unit cube below m.xi = a
  (Scardovelli & Zaleski, J. Comput. Phys. 164, 2000), regrouped so that no
  case divides a cancelling difference by a small m.  Pinned against the
  oracle's exact polyhedral cut (calcSubCell, cutCell.C:343-542) in
  tests/test_learned_oracle.py.
* ``canon_of`` / ``canonical_input``: the 48-fold symmetry canonicalisation
  of a 3x3x3 stencil.  Pinned by invariance under all 48 cube symmetries.
* ``mlp_forward`` / ``mlp_backward``: the 27-128-128-3 GELU network with the
  BF16 rounding points of the MFMA kernels (``rounded=True``) or in plain
  FP64 (``rounded=False``).  The FP64 backward is pinned by central finite
  differences.
* ``cube_plane_offset`` / ``residual``: the plane of a normal that holds a
  cube's own alpha (the inverse of the cube fraction: closed forms of
  Scardovelli & Zaleski on the outer pieces, bracketed Newton on the cubic
  ones), and its mean |fraction - alpha| over the 26 lattice neighbours.
  The offset is pinned against the oracle's findSignedDistance on unit
  cubes and by the round trip through ``plane_cube_fraction``.
"""
from __future__ import annotations

import math

import numpy as np

NIN, KIN, NH, NOUT = 27, 32, 128, 3
P_W1 = 0
P_B1 = P_W1 + NH * NIN
P_W2 = P_B1 + NH
P_B2 = P_W2 + NH * NH
P_W3 = P_B2 + NH
P_B3 = P_W3 + NOUT * NH
N_PARAMS = P_B3 + NOUT


# ---------------------------------------------------------------------------
# BF16
# ---------------------------------------------------------------------------
def f32_to_bf16(x) -> np.ndarray:
    """Round-to-nearest-even float32 -> bf16 bit patterns (uint16)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_to_f(h) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def rbf(x) -> np.ndarray:
    """Value of x after float32 then bf16 rounding."""
    return bf16_to_f(f32_to_bf16(np.asarray(x, dtype=np.float64).astype(np.float32)))


# ---------------------------------------------------------------------------
# cube fraction
# ---------------------------------------------------------------------------
def _cf_sorted(m1, m2, m3, a):
    m12 = m1 + m2
    if a <= 0.0:
        return 0.0
    if a < m1:
        return a * a * a / (6.0 * m1 * m2 * m3)
    base = 3.0 * a * a - 3.0 * a * m1 + m1 * m1
    if a < m2:
        return base / (6.0 * m2 * m3)
    if a < m12 and a < m3:
        b = a - m2
        return (base - b * b * b / m1) / (6.0 * m2 * m3)
    if m3 < m12:
        b, c = a - m2, a - m3
        return (base - (b * b * b + c * c * c) / m1) / (6.0 * m2 * m3)
    return (2.0 * a - m12) / (2.0 * m3)


def cube_fraction(m0, m1, m2, a):
    s = m0 + m1 + m2
    if not s > 0.0:
        return 1.0 if a > 0.0 else 0.0
    m0, m1, m2, a = m0 / s, m1 / s, m2 / s, a / s
    if a <= 0.0:
        return 0.0
    if a >= 1.0:
        return 1.0
    m = sorted((m0, m1, m2))
    if a <= 0.5:
        return _cf_sorted(m[0], m[1], m[2], a)
    return 1.0 - _cf_sorted(m[0], m[1], m[2], 1.0 - a)


def plane_cube_fraction(n, D, lo, h):
    """Liquid fraction of [lo, lo+h]^3 on the side n.x + D < 0."""
    rhs = -D - (n[0] * lo[0] + n[1] * lo[1] + n[2] * lo[2])
    m = [n[0] * h, n[1] * h, n[2] * h]
    for k in range(3):
        if m[k] < 0.0:
            rhs -= m[k]
            m[k] = -m[k]
    return cube_fraction(m[0], m[1], m[2], rhs)


# ---------------------------------------------------------------------------
# stencil symmetry
# ---------------------------------------------------------------------------
def lat_index(dx, dy, dz):
    return 9 * (dx + 1) + 3 * (dy + 1) + (dz + 1)


def youngs_normal(a):
    """-grad alpha with 1-2-1 weights (unit spacing, no normalisation)."""
    g = [0.0, 0.0, 0.0]
    for u in (-1, 0, 1):
        for v in (-1, 0, 1):
            w = (2.0 if u == 0 else 1.0) * (2.0 if v == 0 else 1.0)
            g[0] += w * (a[lat_index(1, u, v)] - a[lat_index(-1, u, v)])
            g[1] += w * (a[lat_index(u, 1, v)] - a[lat_index(u, -1, v)])
            g[2] += w * (a[lat_index(u, v, 1)] - a[lat_index(u, v, -1)])
    return [-g[0], -g[1], -g[2]]


def canon_of(a):
    """(perm, sgn): canonical axis c <- mesh axis perm[c], sign sgn[c]."""
    n = youngs_normal(a)
    sg = [-1.0 if v < 0.0 else 1.0 for v in n]
    mag = [abs(v) for v in n]
    perm = sorted(range(3), key=lambda k: -mag[k])  # stable: ties keep index order
    return perm, [sg[p] for p in perm]


def canon_source(perm, sgn, Lc):
    e = (Lc // 9 - 1, (Lc // 3) % 3 - 1, Lc % 3 - 1)
    s = [0, 0, 0]
    for k in range(3):
        s[perm[k]] = -e[k] if sgn[k] < 0.0 else e[k]
    return lat_index(*s)


def to_canon(perm, sgn, v):
    return [sgn[k] * v[perm[k]] for k in range(3)]


def from_canon(perm, sgn, vc):
    v = [0.0, 0.0, 0.0]
    for k in range(3):
        v[perm[k]] = sgn[k] * vc[k]
    return v


def canonical_input(a):
    """BF16 network input (32 uint16) of a mesh-frame stencil a[27]."""
    perm, sgn = canon_of(a)
    vals = np.array([2.0 * a[canon_source(perm, sgn, L)] - 1.0 for L in range(NIN)])
    x = np.zeros(KIN, np.uint16)
    x[:NIN] = f32_to_bf16(vals.astype(np.float32))
    return x


# ---------------------------------------------------------------------------
# network
# ---------------------------------------------------------------------------
def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[P_W1:P_B1].reshape(NH, NIN)
    b1 = p[P_B1:P_W2]
    W2 = p[P_W2:P_B2].reshape(NH, NH)
    b2 = p[P_B2:P_W3]
    W3 = p[P_W3:P_B3].reshape(NOUT, NH)
    b3 = p[P_B3:N_PARAMS]
    return W1, b1, W2, b2, W3, b3


def _gelu(x):
    erf = np.vectorize(math.erf)
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def _gelu_grad(x):
    erf = np.vectorize(math.erf)
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def mlp_forward(params, x, rounded=True):
    """Outputs (n,3) for inputs x (n,27) float, or BF16 bit patterns (n,32)."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    if rounded:  # the device keeps biases and master weights in FP32
        b1, b2, b3 = (np.asarray(b, np.float32).astype(np.float64) for b in (b1, b2, b3))
    X = bf16_to_f(x[:, :NIN]) if x.dtype == np.uint16 else np.asarray(x, np.float64)
    if rounded:
        W1, W2, W3 = rbf(W1), rbf(W2), rbf(W3)
    Z1 = X @ W1.T + b1
    H1 = rbf(_gelu(Z1)) if rounded else _gelu(Z1)
    Z2 = H1 @ W2.T + b2
    H2 = rbf(_gelu(Z2)) if rounded else _gelu(Z2)
    O = H2 @ W3.T + b3
    return O, (X, Z1, H1, Z2, H2)


def mlp_loss_grad(params, x, y, rounded=True):
    """(loss = sum |o - y|^2 / n, gradient in packed order).  With rounded=True
    the weight gradients use the BF16 values of dZ2 and dZ1 and the bias
    gradients their unrounded values, as the MFMA kernel does."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    n = x.shape[0]
    O, (X, Z1, H1, Z2, H2) = mlp_forward(params, x, rounded)
    Y = np.asarray(y, np.float64)[:, :NOUT]
    diff = O - Y
    loss = float(np.sum(diff * diff)) / n
    dO = 2.0 * diff / n
    W2u, W3u = (rbf(W2), rbf(W3)) if rounded else (W2, W3)
    gW3 = dO.T @ H2
    gb3 = dO.sum(0)
    dZ2 = (dO @ W3u) * _gelu_grad(Z2)
    dZ2r = rbf(dZ2) if rounded else dZ2
    gW2 = dZ2r.T @ H1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2r @ W2u) * _gelu_grad(Z1)
    dZ1r = rbf(dZ1) if rounded else dZ1
    gW1 = dZ1r.T @ X
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


# ---------------------------------------------------------------------------
# residual of a plane over the lattice neighbours
# ---------------------------------------------------------------------------
def _area_sorted(m1, m2, m3, a):
    """d _cf_sorted / da (the cut area), piece by piece."""
    m12 = m1 + m2
    if a <= 0.0:
        return 0.0
    if a < m1:
        return a * a / (2.0 * m1 * m2 * m3)
    base = 6.0 * a - 3.0 * m1
    if a < m2:
        return base / (6.0 * m2 * m3)
    if a < m12 and a < m3:
        b = a - m2
        return (base - 3.0 * b * b / m1) / (6.0 * m2 * m3)
    if m3 < m12:
        b, c = a - m2, a - m3
        return (base - 3.0 * (b * b + c * c) / m1) / (6.0 * m2 * m3)
    return 1.0 / m3


def _level_sorted(m1, m2, m3, V):
    """Inverse of _cf_sorted on [0, 1/2]: closed forms on the first two pieces
    and the linear one (Scardovelli & Zaleski 2000), Newton's method inside
    the bracket of a cubic piece, bisecting when a step leaves it."""
    if V <= 0.0:
        return 0.0
    m12 = m1 + m2
    V1 = _cf_sorted(m1, m2, m3, m1)
    if V < V1:
        return math.pow(6.0 * m1 * m2 * m3 * V, 1.0 / 3.0)
    V2 = _cf_sorted(m1, m2, m3, m2)
    if V < V2:
        return 0.5 * m1 + math.sqrt(max(2.0 * m2 * m3 * V - m1 * m1 / 12.0, 0.0))
    if m12 <= m3:
        V3 = _cf_sorted(m1, m2, m3, m12)
        if V >= V3:
            return m3 * V + 0.5 * m12
        lo, hi, Vlo, Vhi = m2, m12, V2, V3
    else:
        V3 = _cf_sorted(m1, m2, m3, m3)
        if V < V3:
            lo, hi, Vlo, Vhi = m2, m3, V2, V3
        else:
            lo, hi, Vlo, Vhi = m3, 0.5, V3, _cf_sorted(m1, m2, m3, 0.5)
    a = lo + (hi - lo) * ((V - Vlo) / (Vhi - Vlo)) if Vhi > Vlo else lo
    for _ in range(12):
        f = _cf_sorted(m1, m2, m3, a) - V
        if f == 0.0:
            break
        if f > 0.0:
            hi = a
        else:
            lo = a
        fp = _area_sorted(m1, m2, m3, a)
        an = a - f / fp if fp > 0.0 else 0.5 * (lo + hi)
        if abs(an - a) <= 2e-16:
            a = an
            break
        if not (lo < an < hi):
            an = 0.5 * (lo + hi)
        a = an
    return a


def cube_plane_offset(n, a, lo, h):
    """Plane constant D of the cube [lo, lo+h]^3 with liquid fraction a on the
    side n.x + D < 0: the inverse of plane_cube_fraction (reflect the axes
    with n < 0, normalise and sort the coefficients, invert the sorted cube
    fraction, undo the reflection)."""
    m = [n[0] * h, n[1] * h, n[2] * h]
    neg = [mk if mk < 0.0 else 0.0 for mk in m]
    m = [abs(mk) for mk in m]
    s = m[0] + m[1] + m[2]
    if not s > 0.0 or a <= 0.0:
        lev = 0.0
    elif a >= 1.0:
        lev = 1.0
    else:
        sm = sorted(mk / s for mk in m)
        lev = _level_sorted(*sm, a) if a <= 0.5 else 1.0 - _level_sorted(*sm, 1.0 - a)
    rhs = lev * s
    for k in range(3):
        if neg[k] < 0.0:
            rhs += neg[k]
    return -rhs - (n[0] * lo[0] + n[1] * lo[1] + n[2] * lo[2])


def residual(n, d, lo, h, alpha27):
    """Mean |fraction - alpha| over the 26 lattice neighbour cubes lo + e h of
    the cell cube [lo, lo+h]^3 for the plane n.x + d = 0."""
    acc = 0.0
    for L in range(NIN):
        if L == 13:
            continue
        e = (L // 9 - 1, (L // 3) % 3 - 1, L % 3 - 1)
        clo = [lo[k] + e[k] * h for k in range(3)]
        acc += abs(plane_cube_fraction(n, d, clo, h) - alpha27[L])
    return acc / 26.0
