"""Quadrilateral spectral-element meshes and global boundary numbering.

This is the host-side set-up that semtex's ``Mesh`` + ``enumerate`` produce
for a solver: nodal coordinates per element, the element-boundary to global
map ``btog``, the essential-BC masks ``bmask``/``emask``, ``nglobal``,
``nsolve`` and the band width ``nband`` (src/numbersys.h:4-80,
utility/enumerate.cpp:110-199), plus the boundary segments of
``BoundarySys`` (src/bsys.cpp:42-157).

Numbering rules kept from the reference:
  * element-boundary nodes are visited in the CCW edge-major order of
    ``emap`` (femlib/mapping.c:46-132);
  * essential (Dirichlet) nodes are numbered last, so unknowns are
    ``0..nsolve-1`` (utility/enumerate.cpp:635-692);
  * ``nband`` = 1 + max over elements of (max - min) global number among the
    element's non-essential boundary nodes (utility/enumerate.cpp:923-965).

The ordering of the unknowns is free (the solution does not depend on it,
only the band width does).  ``number_nodes`` tries reverse Cuthill-McKee (the
reference's ``enumerate -O1`` heuristic) and layered orderings that fold a
periodic direction, and keeps the one with the smallest band.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

ESSENTIAL, NATURAL, AXIS = 0, 1, 2


def gll(np_: int):
    """GLL points, weights and derivative matrix D[i,j] = h_j'(z_i).

    Same rule as femlib/polylib.c:312 ``zwglj`` / :513 ``Dglj`` with
    alpha = beta = 0, computed by Newton iteration on the Legendre
    recurrence (von Winckel's published ``lglnodes`` method).
    """
    N = np_ - 1
    if np_ == 1:
        return np.zeros(1), np.full(1, 2.0), np.zeros((1, 1))
    x = -np.cos(np.pi * np.arange(np_) / N)
    for _ in range(100):
        P = np.zeros((np_, N + 1))
        P[:, 0] = 1.0
        P[:, 1] = x
        for k in range(2, N + 1):
            P[:, k] = ((2 * k - 1) * x * P[:, k - 1] - (k - 1) * P[:, k - 2]) / k
        xold = x
        x = xold - (x * P[:, N] - P[:, N - 1]) / ((N + 1) * P[:, N])
        x[0], x[-1] = -1.0, 1.0
        if np.max(np.abs(x - xold)) < 1e-16:
            break
    x = 0.5 * (x - x[::-1])
    PN = np.polynomial.legendre.legval(x, np.eye(N + 1)[N])
    w = 2.0 / (N * (N + 1) * PN ** 2)
    D = np.zeros((np_, np_))
    for i in range(np_):
        for j in range(np_):
            if i != j:
                D[i, j] = PN[i] / (PN[j] * (x[i] - x[j]))
    D[0, 0] = -0.25 * N * (N + 1)
    D[N, N] = 0.25 * N * (N + 1)
    return x, w, D


def edgemaps(np_: int):
    """emap (CCW edges, then interior row-major) and its inverse pmap."""
    nm = np_ - 1
    em = [0]
    n = 0
    for step in (1, np_, -1, -np_):
        for _ in range(nm):
            n += step
            em.append(n)
    em = em[: 4 * nm]
    for i in range(1, nm):
        for j in range(1, nm):
            em.append(i * np_ + j)
    em = np.array(em, dtype=np.int32)
    pm = np.empty_like(em)
    pm[em] = np.arange(np_ * np_, dtype=np.int32)
    return em, pm


@dataclass
class Mesh2D:
    """One 2D spectral-element mesh with its numbering and boundary segments."""

    np: int
    nel: int
    x: np.ndarray            # [nel, np*np] float64
    y: np.ndarray            # [nel, np*np] float64
    btog: np.ndarray         # [nel, next] int32
    bmask: np.ndarray        # [nel, next] int32
    emask: np.ndarray        # [nel] int32
    nglobal: int
    nsolve: int
    nband: int
    bnd_elmt: np.ndarray     # [nbound] int32
    bnd_side: np.ndarray     # [nbound] int32
    bnd_kind: np.ndarray     # [nbound] int32
    cyl: int = 0
    info: dict = field(default_factory=dict)

    @property
    def next(self) -> int:
        return 4 * (self.np - 1)

    @property
    def nint(self) -> int:
        return (self.np - 2) ** 2

    @property
    def npnp(self) -> int:
        return self.np * self.np

    @property
    def nbound(self) -> int:
        return int(self.bnd_elmt.shape[0])

    @property
    def plane_size(self) -> int:
        """Geometry::planeSize: nel*np*np rounded up to even (src/geometry.cpp:108)."""
        n = self.nel * self.npnp
        return n + (n & 1)

    def side_xy(self, b: int):
        """Coordinates of boundary segment b's np nodes in CCW traverse order."""
        e, s, np_ = int(self.bnd_elmt[b]), int(self.bnd_side[b]), self.np
        idx = side_node_index(np_, s)
        return self.x[e, idx], self.y[e, idx]


def side_node_index(np_: int, side: int) -> np.ndarray:
    """Row-major node indices of one element side in CCW order (Element::sideGeom)."""
    k = np.arange(np_)
    if side == 0:
        return k
    if side == 1:
        return k * np_ + (np_ - 1)
    if side == 2:
        return (np_ - 1) * np_ + (np_ - 1 - k)
    return (np_ - 1 - k) * np_


def _bandwidth(btog, bmask, nsolve):
    free = np.where(bmask == 0, btog, -1)
    hi = free.max(axis=1)
    lo = np.where(bmask == 0, btog, np.iinfo(np.int32).max).min(axis=1)
    ok = hi >= 0
    if not ok.any():
        return 1
    return int((hi[ok] - lo[ok]).max()) + 1


def number_nodes(keys_elem, essential_node, coords, periods, ordering="auto"):
    """Number unique boundary nodes: unknowns first (ordered for a small
    band), essential nodes last.

    keys_elem: [nel, next] unique-node id per element-boundary node.
    essential_node: [nuniq] bool.  coords: [nuniq, 2] representative x, y.
    periods: (Lx or None, Ly or None).
    Returns btog [nel, next] int32, nsolve, chosen ordering name.
    """
    nuniq = essential_node.shape[0]
    free_ids = np.flatnonzero(~essential_node)
    ess_ids = np.flatnonzero(essential_node)
    candidates = {}

    def layered(major, fold_minor):
        # sort by the major coordinate, then by the (folded if periodic)
        # minor coordinate so that periodic neighbours stay close.
        minor = 1 - major
        c = coords[free_ids]
        a = np.round(c[:, major], 9)
        b = c[:, minor].copy()
        P = periods[minor]
        if fold_minor and P:
            lo = coords[:, minor].min()
            t = (b - lo) / P
            f = np.minimum(t, 1.0 - t)
            b = np.round(f, 9) * 2 + (t >= 0.5) * 1e-12
        return free_ids[np.lexsort((b, a))]

    candidates["layer_y"] = layered(1, True)
    candidates["layer_x"] = layered(0, True)
    try:
        import scipy.sparse as sp
        from scipy.sparse.csgraph import reverse_cuthill_mckee

        rows, cols = [], []
        pos = -np.ones(nuniq, dtype=np.int64)
        pos[free_ids] = np.arange(free_ids.size)
        for ke in keys_elem:
            f = pos[ke]
            f = f[f >= 0]
            rows.append(np.repeat(f, f.size))
            cols.append(np.tile(f, f.size))
        if free_ids.size:
            A = sp.csr_matrix((np.ones(sum(r.size for r in rows)),
                               (np.concatenate(rows), np.concatenate(cols))),
                              shape=(free_ids.size, free_ids.size))
            perm = reverse_cuthill_mckee(A, symmetric_mode=True)
            candidates["rcm"] = free_ids[perm]
    except ImportError:  # pragma: no cover - scipy is present in the image
        pass

    best = None
    for name, order in candidates.items():
        if ordering != "auto" and name != ordering:
            continue
        gid = -np.ones(nuniq, dtype=np.int64)
        gid[order] = np.arange(order.size)
        gid[ess_ids] = free_ids.size + np.arange(ess_ids.size)
        btog = gid[keys_elem].astype(np.int32)
        bmask = essential_node[keys_elem].astype(np.int32)
        bw = _bandwidth(btog, bmask, free_ids.size)
        if best is None or bw < best[0]:
            best = (bw, btog, name)
    return best[1], int(free_ids.size), best[2]


def rect_mesh(nx: int, ny: int, np_: int, x0=0.0, x1=1.0, y0=0.0, y1=1.0,
              bc=("D", "D", "D", "D"), periodic_x=False, periodic_y=False,
              cyl=False, ordering="auto", xs=None, ys=None) -> Mesh2D:
    """nx x ny grid of straight-sided quadrilateral elements.

    bc gives the boundary kind on (bottom y=y0, right x=x1, top y=y1,
    left x=x0): "D" essential, "N" natural, "A" axis (natural with no
    flux: kind AXIS, a boundary segment with a place in the line data).
    A periodic direction ignores its two entries.  xs/ys optionally give the
    element break points (length nx+1 / ny+1) for graded meshes.
    Elements are numbered with x fastest; every element's vertices are
    listed CCW from its lower-left corner (the session-file convention).
    """
    z, _, _ = gll(np_)
    xs = np.linspace(x0, x1, nx + 1) if xs is None else np.asarray(xs, float)
    ys = np.linspace(y0, y1, ny + 1) if ys is None else np.asarray(ys, float)
    x0, x1, y0, y1 = xs[0], xs[-1], ys[0], ys[-1]
    nel = nx * ny
    npnp = np_ * np_
    X = np.empty((nel, npnp))
    Y = np.empty((nel, npnp))
    for ey in range(ny):
        for ex in range(nx):
            e = ey * nx + ex
            xa, xb = xs[ex], xs[ex + 1]
            ya, yb = ys[ey], ys[ey + 1]
            gx = xa + 0.5 * (1.0 + z) * (xb - xa)
            gy = ya + 0.5 * (1.0 + z) * (yb - ya)
            X[e] = np.tile(gx, np_)
            Y[e] = np.repeat(gy, np_)
    return _finish(np_, X, Y, nx, ny, (x0, x1, y0, y1), bc, periodic_x,
                   periodic_y, cyl, ordering)


def _finish(np_, X, Y, nx, ny, box, bc, periodic_x, periodic_y, cyl, ordering):
    x0, x1, y0, y1 = box
    nel = X.shape[0]
    emap, _ = edgemaps(np_)
    next_ = 4 * (np_ - 1)
    Lx = (x1 - x0) if periodic_x else None
    Ly = (y1 - y0) if periodic_y else None
    scale = max(x1 - x0, y1 - y0)
    tol = 1e-9 * scale

    bx = X[:, emap[:next_]].copy()
    by = Y[:, emap[:next_]].copy()
    kx, ky = bx.copy(), by.copy()
    if Lx:
        kx = np.where(np.abs(kx - x1) < tol, x0, kx)
    if Ly:
        ky = np.where(np.abs(ky - y1) < tol, y0, ky)
    q = np.stack([np.round(kx / tol).astype(np.int64),
                  np.round(ky / tol).astype(np.int64)], axis=-1).reshape(-1, 2)
    uniq, inv = np.unique(q, axis=0, return_inverse=True)
    keys = inv.reshape(nel, next_)
    nuniq = uniq.shape[0]
    coords = np.zeros((nuniq, 2))
    coords[keys.ravel(), 0] = kx.ravel()
    coords[keys.ravel(), 1] = ky.ravel()

    # boundary segments and essential nodes
    bnd_e, bnd_s, bnd_k = [], [], []
    essential = np.zeros(nuniq, dtype=bool)
    nm = np_ - 1
    for e in range(nel):
        ex, ey = e % nx, e // nx
        for side, (on, kind) in enumerate((
                (ey == 0 and not periodic_y, bc[0]),
                (ex == nx - 1 and not periodic_x, bc[1]),
                (ey == ny - 1 and not periodic_y, bc[2]),
                (ex == 0 and not periodic_x, bc[3]))):
            if not on:
                continue
            k = {"D": ESSENTIAL, "N": NATURAL, "A": AXIS}[kind]
            bnd_e.append(e)
            bnd_s.append(side)
            bnd_k.append(k)
            if k == ESSENTIAL:
                loc = [side * nm + i for i in range(nm)] + [0 if side == 3 else side * nm + nm]
                essential[keys[e, loc]] = True

    btog, nsolve, how = number_nodes(keys, essential, coords, (Lx, Ly), ordering)
    bmask = essential[keys].astype(np.int32)
    emask = bmask.any(axis=1).astype(np.int32)
    nband = _bandwidth(btog, bmask, nsolve)
    return Mesh2D(np=np_, nel=nel, x=np.ascontiguousarray(X),
                  y=np.ascontiguousarray(Y), btog=np.ascontiguousarray(btog),
                  bmask=np.ascontiguousarray(bmask), emask=emask,
                  nglobal=int(nuniq), nsolve=nsolve, nband=nband,
                  bnd_elmt=np.array(bnd_e, np.int32),
                  bnd_side=np.array(bnd_s, np.int32),
                  bnd_kind=np.array(bnd_k, np.int32), cyl=int(bool(cyl)),
                  info={"ordering": how, "nx": nx, "ny": ny,
                        "box": (x0, x1, y0, y1)})
