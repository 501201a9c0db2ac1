"""Flow past a circular cylinder: a structured O-grid around the cylinder
inside a square, eight H-grid blocks filling a rectangular channel, the
quadrilaterals extruded one cell in z.  The shape of the reference's
cylinder_re100 case (D = 1, channel x in [-10, 20] D, y in [-10, 10] D,
inflow at -x, pressure outlet at +x, symmetry far-field and front/back), with
hexahedra instead of its triangular prisms and sized by the caller.

Patches: 0 inlet, 1 outlet, 2 bottom, 3 top, 4 front, 5 back, 6 cylinder.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

FV, FG, ZG, IO, SYM = range(5)
PATCHES = ("inlet", "outlet", "bottom", "top", "front", "back", "cylinder")


def _geometric(n: int, length: float, first: float) -> np.ndarray:
    """n+1 points on [0, length] whose first interval is `first` and whose
    intervals grow geometrically (uniform when first >= length / n)."""
    if first * n >= length:
        return np.linspace(0.0, length, n + 1)
    lo, hi = 1.0, 2.0
    while first * (hi**n - 1) / (hi - 1) < length:
        hi *= 2
    for _ in range(200):  # bisection on the growth ratio
        q = 0.5 * (lo + hi)
        if first * (q**n - 1) / (q - 1) < length:
            lo = q
        else:
            hi = q
    q = 0.5 * (lo + hi)
    x = np.concatenate([[0.0], np.cumsum(first * q ** np.arange(n))])
    return x * (length / x[-1])


def _block(xs: np.ndarray, ys: np.ndarray):
    """Vertices and CCW quads of the tensor grid xs x ys."""
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel()], 1)
    nx, ny = len(xs) - 1, len(ys) - 1
    idx = np.arange((nx + 1) * (ny + 1)).reshape(nx + 1, ny + 1)
    q = np.stack([idx[:-1, :-1], idx[1:, :-1], idx[1:, 1:], idx[:-1, 1:]], -1).reshape(-1, 4)
    return pts, q


def _smooth_corners(verts, quads, B, radius, rho, sweeps):
    """Damped Laplacian sweeps v += (w/2)(mean of edge neighbours - v) with a
    Gaussian weight w of width rho around the square's corners, where five
    cells meet.  A fixed number of sweeps moves the corner region part of the
    way towards the harmonic positions without flattening the O-grid's
    radial clustering; the cylinder wall and the far field do not move."""
    nv = len(verts)
    a = quads.ravel()
    b = np.roll(quads, -1, 1).ravel()
    key = np.unique(np.concatenate([a, b]).astype(np.int64) * nv + np.concatenate([b, a]))
    ends, other = key // nv, key % nv
    deg = np.bincount(ends, minlength=nv).astype(float)
    corners = np.array([[B, B], [-B, B], [-B, -B], [B, -B]])
    dist = np.min(np.linalg.norm(verts[:, None, :] - corners[None], axis=2), 1)
    w = np.exp(-(dist / rho) ** 2)
    w[np.hypot(verts[:, 0], verts[:, 1]) < radius * (1 + 1e-9) + 1e-12] = 0.0
    w[w < 1e-6] = 0.0
    active = np.flatnonzero(w > 0)
    sel = w[ends] > 0
    ends, other = ends[sel], other[sel]
    pos = np.full(nv, -1)
    pos[active] = np.arange(len(active))
    avg = sp.csr_matrix((1.0 / deg[ends], (pos[ends], other)), shape=(len(active), nv))
    v = verts.copy()
    wa = 0.5 * w[active, None]
    for _ in range(sweeps):
        v[active] += wa * (avg @ v - v[active])
    return v


def cylinder_quads(n_side=128, n_radial=200, n_far=150, n_up=120, n_down=480, first=0.005, square=2.0,
                   x_min=-10.0, x_max=20.0, y_half=10.0, radius=0.5, smooth_radius=0.8, smooth_sweeps=None):
    """2-D vertices and CCW quads.  n_side cells per square side (4 n_side
    around the cylinder), n_radial O-grid layers starting at `first`, n_far
    cells from the square to y = +-y_half, n_up / n_down cells upstream /
    downstream of the square.  The neighbourhood (smooth_radius) of the
    square's corners, where five cells meet, is relaxed towards the mean of
    the neighbours, which lowers the non-orthogonality there."""
    B = square
    blocks = []
    # O-grid: circle point + f (square point - circle point), periodic in theta.
    t = np.arange(n_side) / n_side
    sq, th = [], []
    for k, (p0, p1) in enumerate((((B, -B), (B, B)), ((B, B), (-B, B)), ((-B, B), (-B, -B)), ((-B, -B), (B, -B)))):
        p0, p1 = np.array(p0), np.array(p1)
        sq.append(p0[None] + t[:, None] * (p1 - p0)[None])
        th.append(-np.pi / 4 + k * np.pi / 2 + t * np.pi / 2)
    sq, th = np.concatenate(sq), np.concatenate(th)
    circ = radius * np.stack([np.cos(th), np.sin(th)], 1)
    f = _geometric(n_radial, 1.0, first / (B - radius))
    pts = circ[None] + f[:, None, None] * (sq - circ)[None]  # (n_radial+1, 4 n_side, 2)
    nt = 4 * n_side
    idx = np.arange((n_radial + 1) * nt).reshape(n_radial + 1, nt)
    jn = (np.arange(nt) + 1) % nt
    q = np.stack([idx[:-1, :], idx[:-1, jn], idx[1:, jn], idx[1:, :]], -1).reshape(-1, 4)
    blocks.append((pts.reshape(-1, 2), q))
    # H-grid: three columns (upstream, square, downstream) x three rows.
    h = 2 * B / n_side
    ys_mid = np.linspace(-B, B, n_side + 1)
    far = _geometric(n_far, y_half - B, h)
    ys_top = B + far
    ys_bot = -(B + far)[::-1]
    xs_mid = np.linspace(-B, B, n_side + 1)
    xs_up = -(B + _geometric(n_up, -x_min - B, h))[::-1]
    xs_down = B + _geometric(n_down, x_max - B, h)
    for xs in (xs_up, xs_mid, xs_down):
        for ys in (ys_bot, ys_mid, ys_top):
            if xs is xs_mid and ys is ys_mid:
                continue
            blocks.append(_block(xs, ys))
    # Merge coincident vertices of adjacent blocks.
    allp, allq, off = [], [], 0
    for p, qq in blocks:
        allp.append(p)
        allq.append(qq + off)
        off += len(p)
    allp = np.concatenate(allp)
    allq = np.concatenate(allq)
    key = np.round(allp * 1e8).astype(np.int64)
    _, first_idx, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    verts = allp[first_idx]
    quads = inv.ravel()[allq]
    if smooth_sweeps is None:  # enough sweeps to reach over smooth_radius / h cells
        smooth_sweeps = int(max(300, 2 * (smooth_radius / h) ** 2))
    verts = _smooth_corners(verts, quads, B, radius, smooth_radius, smooth_sweeps)
    # CCW orientation.
    a = verts[quads]
    area = 0.5 * np.sum(a[:, :, 0] * np.roll(a[:, :, 1], -1, 1) - np.roll(a[:, :, 0], -1, 1) * a[:, :, 1], 1)
    quads[area < 0] = quads[area < 0][:, ::-1]
    return verts, quads


def cylinder_arrays(thickness=0.1, **kw) -> dict:
    """Mesh arrays (the layout MeshData.from_arrays takes) of the extruded
    O+H grid: interior faces first (owner < neighbour, ordered by owner),
    then boundary faces by patch, every face oriented out of its owner."""
    x_min, x_max, y_half = kw.get("x_min", -10.0), kw.get("x_max", 20.0), kw.get("y_half", 10.0)
    radius = kw.get("radius", 0.5)
    v2, quads = cylinder_quads(**kw)
    nv, nc = len(v2), len(quads)
    verts = np.concatenate([np.c_[v2, np.zeros(nv)], np.c_[v2, np.full(nv, thickness)]])
    # Side faces from the quad edges (a, b) in CCW order: points a, b, b', a'.
    a = quads.ravel()
    b = np.roll(quads, -1, 1).ravel()
    cell = np.repeat(np.arange(nc), 4)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    order = np.lexsort((hi, lo))
    lo_s, hi_s = lo[order], hi[order]
    same = (lo_s[1:] == lo_s[:-1]) & (hi_s[1:] == hi_s[:-1])
    first = np.flatnonzero(same)
    e1, e2 = order[first], order[first + 1]  # the two sides of each interior edge
    c1, c2 = cell[e1], cell[e2]
    own = np.minimum(c1, c2)
    nei = np.maximum(c1, c2)
    eo = np.where(c1 < c2, e1, e2)  # the owner's CCW edge gives the outward orientation
    interior = np.stack([a[eo], b[eo], b[eo] + nv, a[eo] + nv], 1)
    io = np.lexsort((nei, own))
    own, nei, interior = own[io], nei[io], interior[io]
    paired = np.zeros(len(a), bool)
    paired[e1] = paired[e2] = True
    be = np.flatnonzero(~paired)
    mid = 0.5 * (v2[a[be]] + v2[b[be]])
    tol = 1e-9 * max(x_max - x_min, y_half)
    patch = np.full(len(be), -1)
    patch[np.abs(mid[:, 0] - x_min) < tol] = 0
    patch[np.abs(mid[:, 0] - x_max) < tol] = 1
    patch[np.abs(mid[:, 1] + y_half) < tol] = 2
    patch[np.abs(mid[:, 1] - y_half) < tol] = 3
    patch[np.hypot(mid[:, 0], mid[:, 1]) < radius * 1.01] = 6
    if np.any(patch < 0):
        raise ValueError("unclassified boundary edge")
    side = np.stack([a[be], b[be], b[be] + nv, a[be] + nv], 1)
    faces, owners, patches = [interior], [own], [np.full(len(own), -1)]
    for pid in (0, 1, 2, 3):
        sel = patch == pid
        faces.append(side[sel]); owners.append(cell[be][sel]); patches.append(np.full(sel.sum(), pid))
    cells = np.arange(nc)
    faces.append(quads[:, ::-1]); owners.append(cells); patches.append(np.full(nc, 4))       # front, normal -z
    faces.append(quads + nv); owners.append(cells); patches.append(np.full(nc, 5))           # back, normal +z
    sel = patch == 6
    faces.append(side[sel]); owners.append(cell[be][sel]); patches.append(np.full(sel.sum(), 6))
    F = np.concatenate(faces)
    nf = len(F)
    neighbour = np.full(nf, -1, np.int64)
    neighbour[:len(nei)] = nei
    return dict(vertices=verts.ravel(), fpo=np.arange(0, 4 * nf + 1, 4, dtype=np.int32),
                fpi=F.ravel().astype(np.int32), owner=np.concatenate(owners).astype(np.int32),
                neighbour=neighbour.astype(np.int32), patch=np.concatenate(patches).astype(np.int32),
                npatch=len(PATCHES), ncells=nc)


def cylinder_bcs(u_inf=1.0):
    """Inlet (U_inf, 0, 0); outlet zeroGradient U and p = 0; symmetry far
    field and front/back; no-slip cylinder (cases/cylinder_re100/README.md)."""
    ut = np.array([FV, ZG, SYM, SYM, SYM, SYM, FV], np.int32)
    uv = np.zeros((7, 3))
    uv[0] = (u_inf, 0.0, 0.0)
    pt = np.array([ZG, FV, SYM, SYM, SYM, SYM, ZG], np.int32)
    return (ut, uv), (pt, np.zeros(7))


def potential_flow(centres: np.ndarray, u_inf=1.0, radius=0.5) -> np.ndarray:
    """Uniform flow past a cylinder: u = U (1 - R^2 (x^2 - y^2) / r^4),
    v = -2 U R^2 x y / r^4 (divergence-free, zero normal velocity on r = R)."""
    x, y = centres[:, 0], centres[:, 1]
    r4 = (x * x + y * y) ** 2
    U = np.zeros((len(x), 3))
    U[:, 0] = u_inf * (1 - radius**2 * (x * x - y * y) / r4)
    U[:, 1] = -2 * u_inf * radius**2 * x * y / r4
    return U
