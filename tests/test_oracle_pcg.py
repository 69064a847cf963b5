"""Pins the oracle's JACPCG restatement (src/field.cpp:711-805,
src/matrix.cpp:269-310, src/element.cpp:1484-1533):

* the Jacobi preconditioner is the inverse diagonal of the global matrix
  assembled independently from dense element matrices (affine elements);
  on curved elements (Q3 != 0) it carries the reference's cross term
  2 Q3 dvr[j*j] dvs[i*i] (linear indices, src/element.cpp:1519), checked
  against the dense diagonal plus that term's difference from the true one;
* with a tight tolerance the PCG solution equals the direct (condensed,
  banded Cholesky) oracle solve, with Dirichlet and Neumann data and
  Fourier constants, for two numberings;
* started from the converged solution it takes no iteration;
* at the reference's default tolerance (TOL_REL = 1e-8) it converges in
  a few iterations to within the tolerance's reach of the direct solve.
"""
import numpy as np
import pytest

from oracle.oracle import OracleMesh, OracleMsys
from semtex_amd.mesh import edgemaps, rect_mesh
from tests.helpers import dense_element


def global_diag(mesh, hcon):
    em, _ = edgemaps(mesh.np)
    nb, ni = mesh.next, mesh.nint
    d = np.zeros(mesh.nglobal + mesh.nel * ni)
    for e in range(mesh.nel):
        g = np.empty(mesh.npnp, dtype=np.int64)
        g[em[:nb]] = mesh.btog[e]
        g[em[nb:]] = mesh.nglobal + e * ni + np.arange(ni)
        He, _ = dense_element(mesh.np, mesh.x[e], mesh.y[e], hcon)
        np.add.at(d, g, np.diag(He))
    return d


@pytest.mark.parametrize("lam2,bk2", [(1.0, 0.0), (2.5, 4.0)])
def test_preconditioner_is_inverse_assembled_diagonal(lam2, bk2):
    m = rect_mesh(3, 2, 7, 0.0, 2.0, -1.0, 0.5, bc=("D", "N", "D", "N"), periodic_x=True)
    S = OracleMsys(OracleMesh(m), lam2, bk2)
    pc = S.pcg_diag()
    want = 1.0 / global_diag(m, lam2 + bk2)
    assert np.max(np.abs(pc - want) / np.abs(want)) < 1e-12


def _problem(ordering):
    m = rect_mesh(4, 3, 7, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "N", "N"),
                  periodic_x=True, ordering=ordering)
    rng = np.random.default_rng(5)
    f = rng.standard_normal(m.nel * m.npnp)
    line = np.zeros(m.nbound * m.np)
    for b in range(m.nbound):
        xs, ys = m.side_xy(b)
        line[b * m.np:(b + 1) * m.np] = np.cos(xs) + 0.3 * ys if m.bnd_kind[b] == 0 else np.sin(xs)
    return m, f, line


@pytest.mark.parametrize("ordering", ["layer_y", "rcm"])
@pytest.mark.parametrize("lam2,bk2", [(3.0, 0.0), (1.0, 9.0)])
def test_tight_pcg_equals_direct(ordering, lam2, bk2):
    m, f, line = _problem(ordering)
    S = OracleMsys(OracleMesh(m), lam2, bk2)
    ud = S.solve_plane(f, line)
    up, it = S.solve_plane_pcg(f, line, tol_rel=1e-15, step_max=5000)
    assert 0 < it < 5000
    assert np.linalg.norm(up - ud) / np.linalg.norm(ud) < 1e-11


def test_converged_guess_takes_no_iteration():
    m, f, line = _problem("layer_y")
    S = OracleMsys(OracleMesh(m), 2.0, 1.0)
    ud = S.solve_plane(f, line)
    up, it = S.solve_plane_pcg(f, line, u0=ud, tol_rel=1e-8)
    assert it == 0
    assert np.linalg.norm(up - ud) / np.linalg.norm(ud) < 1e-14


def test_default_tolerance_converges_to_its_reach():
    m, f, line = _problem("layer_y")
    S = OracleMsys(OracleMesh(m), 1.0 / (1e-3 * 5e-4), 4.0)   # DNS start-up constant
    ud = S.solve_plane(f, line)
    up, it = S.solve_plane_pcg(f, line, tol_rel=1e-8)
    assert 0 < it < 50
    assert np.linalg.norm(up - ud) / np.linalg.norm(ud) < 1e-6


@pytest.mark.parametrize("lam2,bk2", [(1.0, 0.0), (2.5, 4.0)])
def test_preconditioner_cross_term_on_curved_elements(lam2, bk2):
    """HelmholtzDiag's cross term on non-affine (curved) elements, where Q3
    is nonzero.  The reference forms it as 2 Q3[ij] dvr[j*j] dvs[i*i]
    (src/element.cpp:1519,1530), reading the derivative matrix at the
    linear indices j*j and i*i rather than its diagonal (j*np+j, i*np+i).
    Expected: the diagonal of the independently assembled dense element
    matrices (true cross term 2 Q3 D[j,j] D[i,i]) plus, per element node,
    2 Q3 (D.flat[j*j] D.flat[i*i] - D[j,j] D[i,i]), with Q3 = W (rx sx +
    ry sy) computed here from the nodal geometry.  A dropped cross term,
    the true diagonal in its place, a missing factor 2 or a transposed
    derivative matrix each fail this."""
    from semtex_amd.mesh import gll
    m = rect_mesh(3, 2, 7, 0.0, 2.0, -1.0, 0.5, bc=("D", "D", "D", "D"))
    x0, y0 = m.x.copy(), m.y.copy()
    m.x = x0 + 0.06 * np.sin(np.pi * y0) * np.sin(0.5 * np.pi * x0)
    m.y = y0 + 0.05 * np.sin(np.pi * x0)
    S = OracleMsys(OracleMesh(m), lam2, bk2)
    pc = S.pcg_diag()
    np_ = m.np
    _, w, D = gll(np_)
    I = np.eye(np_)
    Dr, Ds = np.kron(I, D), np.kron(D, I)
    em, _ = edgemaps(np_)
    nb, ni = m.next, m.nint
    want = global_diag(m, lam2 + bk2)
    jj = np.array([D.flat[j * j] for j in range(np_)])
    corr = np.zeros(m.npnp)
    for i in range(np_):
        for j in range(np_):
            corr[i * np_ + j] = jj[j] * jj[i] - D[j, j] * D[i, i]
    q3_seen = 0.0
    for e in range(m.nel):
        X, Y = m.x[e], m.y[e]
        xr, xs, yr, ys = Dr @ X, Ds @ X, Dr @ Y, Ds @ Y
        J = xr * ys - xs * yr
        rx, ry, sx, sy = ys / J, -xs / J, -yr / J, xr / J
        Q3 = np.outer(w, w).ravel() * J * (rx * sx + ry * sy)
        q3_seen = max(q3_seen, float(np.max(np.abs(Q3))))
        g = np.empty(m.npnp, dtype=np.int64)
        g[em[:nb]] = m.btog[e]
        g[em[nb:]] = m.nglobal + e * ni + np.arange(ni)
        np.add.at(want, g, 2.0 * Q3 * corr)
    assert q3_seen > 1e-3                       # the elements are genuinely curved
    assert np.max(np.abs(pc - 1.0 / want) * np.abs(want)) < 1e-12
    # and the reference's indexing is not the true diagonal here
    assert np.max(np.abs(pc * global_diag(m, lam2 + bk2) - 1.0)) > 1e-6
