"""Shared test helpers: independent dense reference assemblies and problem
set-ups (forcing / boundary values) for the elliptic sessions."""
import numpy as np

from semtex_amd.mesh import edgemaps, gll


def dense_element(np_, X, Y, hcon):
    """Elemental Helmholtz matrix from nodal gradient matrices (a different
    route from the reference's Q1..Q3 split)."""
    _, w, D = gll(np_)
    I = np.eye(np_)
    Dr, Ds = np.kron(I, D), np.kron(D, I)
    xr, xs, yr, ys = Dr @ X, Ds @ X, Dr @ Y, Ds @ Y
    J = xr * ys - xs * yr
    rx, ry, sx, sy = ys / J, -xs / J, -yr / J, xr / J
    Gx = rx[:, None] * Dr + sx[:, None] * Ds
    Gy = ry[:, None] * Dr + sy[:, None] * Ds
    W = np.outer(w, w).ravel() * J
    return Gx.T @ (W[:, None] * Gx) + Gy.T @ (W[:, None] * Gy) + hcon * np.diag(W), W


def dense_global_solve(mesh, hcon, f, line=None):
    """Solve the full (uncondensed) assembled Helmholtz system densely.

    Unknown numbering: global boundary ids from btog, then element
    interiors.  f is [nel, np*np] physical forcing; returns [nel, np*np].
    """
    np_, nel = mesh.np, mesh.nel
    em, _ = edgemaps(np_)
    nb, ni = mesh.next, mesh.nint
    ntot = mesh.nglobal + nel * ni
    A = np.zeros((ntot, ntot))
    b = np.zeros(ntot)
    gids = []
    for e in range(nel):
        g = np.empty(np_ * np_, dtype=np.int64)
        g[em[:nb]] = mesh.btog[e]
        g[em[nb:]] = mesh.nglobal + e * ni + np.arange(ni)
        gids.append(g)
        He, W = dense_element(np_, mesh.x[e], mesh.y[e], hcon)
        A[np.ix_(g, g)] += He
        np.add.at(b, g, -W * f[e])
    # boundary data
    uess = np.zeros(ntot)
    is_ess = np.zeros(ntot, bool)
    is_ess[mesh.nsolve:mesh.nglobal] = True
    if line is not None:
        from semtex_amd.mesh import side_node_index
        z, w, D = gll(np_)
        nm = np_ - 1
        for bnd in range(mesh.nbound):
            e, s = int(mesh.bnd_elmt[bnd]), int(mesh.bnd_side[bnd])
            loc = [s * nm + i for i in range(nm)] + [0 if s == 3 else s * nm + nm]
            g = mesh.btog[e, loc]
            vals = line[bnd * np_:(bnd + 1) * np_]
            if mesh.bnd_kind[bnd] == 0:
                uess[g] = vals
            else:
                idx = side_node_index(np_, s)
                xs, ys = mesh.x[e, idx], mesh.y[e, idx]
                # straight side: |dx/dr| constant = length/2
                L = np.hypot(xs[-1] - xs[0], ys[-1] - ys[0])
                np.add.at(b, g, vals * w * L / 2)
    free = ~is_ess
    rhs = b[free] - A[np.ix_(free, is_ess)] @ uess[is_ess]
    u = uess.copy()
    u[free] = np.linalg.solve(A[np.ix_(free, free)], rhs)
    out = np.empty((nel, np_ * np_))
    for e in range(nel):
        out[e] = u[gids[e]]
    return out


def boundary_lines(mesh, func):
    """Line data (nbound*np) holding func(x, y) at each segment's nodes."""
    line = np.zeros(mesh.nbound * mesh.np)
    for b in range(mesh.nbound):
        xs, ys = mesh.side_xy(b)
        line[b * mesh.np:(b + 1) * mesh.np] = func(xs, ys)
    return line
