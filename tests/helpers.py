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
            elif mesh.bnd_kind[bnd] == 1:
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


def kovas5_case():
    """mesh/kovas5: Kovasznay flow in the y-z plane (regress/kovas5.ok).
    Returns the velocity and pressure meshes, the DNS parameters, the
    exact-solution function ex(X, Y, Z) -> [u, v, w], the z planes and the
    physical-space velocity boundary lines [3][nz][nbound*np]."""
    from semtex_amd.mesh import rect_mesh
    Re = 40.0
    lam = Re / 2 - np.sqrt(0.25 * Re * Re + 4 * np.pi ** 2)
    nz, Lz, dt, np_ = 4, 1.0, 0.005, 8
    xs, ys = [-0.5, 0.0, 0.5], [-0.5, 0.0, 1.0]
    mu = rect_mesh(2, 2, np_, bc=("D", "D", "D", "D"), xs=xs, ys=ys)
    mp = rect_mesh(2, 2, np_, bc=("N", "N", "N", "N"), xs=xs, ys=ys)

    def ex(X, Y, Z):
        tp = 2 * np.pi
        return [0 * X, 1 - np.exp(lam * Y) * np.cos(tp * Z),
                lam / tp * np.exp(lam * Y) * np.sin(tp * Z)]

    z = np.arange(nz) * Lz / nz
    lines = np.zeros((3, nz, mu.nbound * np_))
    for b in range(mu.nbound):
        xb, yb = mu.side_xy(b)
        for k in range(nz):
            e = ex(xb, yb, np.full_like(xb, z[k]))
            for i in range(3):
                lines[i, k, b * np_:(b + 1) * np_] = e[i]
    return dict(mesh_u=mu, mesh_p=mp, nz=nz, beta=2 * np.pi / Lz, kinvis=1.0 / Re, dt=dt,
                steps=200, ex=ex, z=z, lines=lines)
