"""A second, independently ordered FP64 restatement of one mode's DIRECT
plane solve (Field::solve, src/field.cpp:598-709), for measuring how far two
valid FP64 implementations of the same solve can differ.

TEST INFRASTRUCTURE ONLY.  The elemental matrices come from the oracle's
HelmholtzRow restatement (or_helmholtz_row, src/element.cpp:1411-1481), so
the discrete problem is the same; everything after that takes a different
rounding path from oracle/sem_oracle.c: the interior blocks are inverted by
LAPACK (numpy.linalg.inv, partial pivoting) instead of the oracle's own
Gauss-Jordan, the Schur complements and condensed right-hand sides use BLAS
products, and the assembled condensed system is factorised once as a dense
matrix by LAPACK LU (scipy.linalg.lu_factor) instead of the band Cholesky.
The singular mode keeps the reference's rule: its last unknown is dropped
(src/matrix.cpp:154-162), exactly as in the oracle's numbering.

Dense storage limits it to meshes of a few thousand global nodes (the
12x24-element N_P=8 channel: 3,828, a 117 MB matrix).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sl

from . import oracle as O


class DenseAltSystem:
    """One mode's system (lambda2, betak2) on a Cartesian mesh."""

    def __init__(self, mesh, lambda2, betak2, nsolve):
        m = mesh
        self.m = m
        np_ = m.np
        self.npnp = np_ * np_
        self.em, _ = O.edgemaps(np_)
        nb = m.next
        self.nsolve = nsolve
        self.geo, self.Bm, self.Hinv = [], [], []
        A = np.zeros((nsolve, nsolve))
        for e in range(m.nel):
            _, geo = O.mapping(np_, m.x[e], m.y[e], m.cyl)
            H = O.helmholtz_full(np_, geo, m.y[e], lambda2, betak2, m.cyl)
            H = H[np.ix_(self.em, self.em)]
            hbb, hbi, hii = H[:nb, :nb], H[:nb, nb:], H[nb:, nb:]
            Hinv = np.linalg.inv(hii)
            Bm = hbi @ Hinv
            Sc = hbb - Bm @ hbi.T
            g = m.btog[e]
            ok = g < nsolve
            A[np.ix_(g[ok], g[ok])] += Sc[np.ix_(ok, ok)]
            self.geo.append(geo)
            self.Bm.append(Bm)
            self.Hinv.append(Hinv)
        self.lu = sl.lu_factor(A) if nsolve else None
        # natural-BC quadrature (Natural::sum) from the oracle's side geometry
        self.nat = []
        from semtex_amd.mesh import NATURAL, side_node_index
        for b in range(m.nbound):
            if m.bnd_kind[b] != NATURAL:
                continue
            e, s = int(m.bnd_elmt[b]), int(m.bnd_side[b])
            area, nx, ny = np.zeros(np_), np.zeros(np_), np.zeros(np_)
            O.lib().or_side_geom(np_, np.ascontiguousarray(self.geo[e].ravel()),
                                 np.ascontiguousarray(m.x[e]), np.ascontiguousarray(m.y[e]),
                                 s, area, nx, ny)
            # the side's nodes in CCW order -> global numbers via the element
            idx = side_node_index(np_, s)
            pos = np.empty(self.npnp, dtype=np.int64)
            pos[self.em] = np.arange(self.npnp)
            glob = m.btog[e][pos[idx]]
            self.nat.append((b, glob, area))

    def solve_plane(self, f, line=None):
        """Natural / zero-essential boundary data only (the DNS pressure)."""
        m, np_, nb, npnp = self.m, self.m.np, self.m.next, self.npnp
        f = np.asarray(f, float).reshape(-1)
        R = np.zeros(m.nglobal)
        Fi = []
        for e in range(m.nel):
            fe = -(self.geo[e][O.Q8] * f[e * npnp:(e + 1) * npnp])[self.em]
            np.add.at(R, m.btog[e], fe[:nb] - self.Bm[e] @ fe[nb:])
            Fi.append(fe[nb:])
        if line is not None:
            for b, glob, area in self.nat:
                np.add.at(R, glob, line[b * np_:(b + 1) * np_] * area)
        ub = np.zeros(m.nglobal)
        if self.nsolve:
            ub[:self.nsolve] = sl.lu_solve(self.lu, R[:self.nsolve])
        u = np.zeros(m.nel * npnp)
        for e in range(m.nel):
            u_b = ub[m.btog[e]]
            u_i = self.Hinv[e].T @ Fi[e] - self.Bm[e].T @ u_b
            ue = np.empty(npnp)
            ue[self.em[:nb]] = u_b
            ue[self.em[nb:]] = u_i
            u[e * npnp:(e + 1) * npnp] = ue
        return u
