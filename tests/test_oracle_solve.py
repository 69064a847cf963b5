"""Pin the oracle's banded Cholesky and whole Field::solve DIRECT path.

* pbtrf/pbtrs against dense numpy factorisation and solves;
* the condensed solve against a dense solve of the full (uncondensed)
  assembled system built from an independent element assembly;
* the reference's own analytic elliptic sessions (elliptic/testsuite:
  mesh/helmholtz1, poisson1, laplace4, laplace5, laplace7) and the
  config-1 manufactured solution, against their closed-form solutions.
"""
import numpy as np
import pytest

from semtex_amd.mesh import rect_mesh
from tests.helpers import boundary_lines, dense_global_solve


def random_band_spd(n, kd, rng):
    A = np.zeros((n, n))
    for d in range(1, kd + 1):
        v = rng.uniform(-1, 1, n - d)
        A += np.diag(v, d) + np.diag(v, -d)
    A += np.diag(np.abs(A).sum(axis=1) + rng.uniform(0.5, 1.5, n))
    ab = np.zeros((n, kd + 1))
    for j in range(n):
        for i in range(max(0, j - kd), j + 1):
            ab[j, kd + i - j] = A[i, j]
    return A, ab


@pytest.mark.parametrize("n,kd", [(1, 0), (7, 3), (200, 15), (513, 64)])
def test_band_cholesky_vs_dense(oracle, n, kd):
    rng = np.random.default_rng(n + kd)
    A, ab = random_band_spd(n, kd, rng)
    rc, U = oracle.pbtrf(ab, n, kd)
    assert rc == 0
    Ud = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - kd), j + 1):
            Ud[i, j] = U[j, kd + i - j]
    assert np.allclose(Ud, np.linalg.cholesky(A).T, rtol=0, atol=1e-12 * np.abs(A).max())
    b = rng.standard_normal(n)
    x = oracle.pbtrs(U, n, kd, b)
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xr) < 1e-13 * np.linalg.norm(xr) * 10


def test_band_cholesky_reports_indefinite(oracle):
    A, ab = random_band_spd(20, 2, np.random.default_rng(0))
    ab[7, 2] = -1.0        # diagonal of column 7
    rc, _ = oracle.pbtrf(ab, 20, 2)
    assert rc == 8


def solve2d(oracle, mesh, lam2, f, line=None):
    om = oracle.OracleMesh(mesh)
    S = oracle.OracleMsys(om, lam2, 0.0)
    return S.solve_plane(f.ravel(), line).reshape(mesh.nel, -1)


@pytest.mark.parametrize("ordering", ["rcm", "layer_y"])
def test_condensed_solve_equals_dense_uncondensed(oracle, ordering):
    m = rect_mesh(3, 2, 6, 0.0, 1.5, 0.0, 1.0, bc=("D", "N", "D", "N"),
                  ordering=ordering)
    rng = np.random.default_rng(5)
    f = rng.standard_normal((m.nel, m.npnp))
    line = rng.standard_normal(m.nbound * m.np)
    # continuous essential data: take it from a smooth function instead
    ess = boundary_lines(m, lambda x, y: np.cos(x) + y)
    for b in range(m.nbound):
        if m.bnd_kind[b] == 0:
            line[b * m.np:(b + 1) * m.np] = ess[b * m.np:(b + 1) * m.np]
    u = solve2d(oracle, m, 2.0, f, line)
    ur = dense_global_solve(m, 2.0, f, line)
    assert np.max(np.abs(u - ur)) < 1e-12 * np.abs(ur).max()


def test_condensed_solve_periodic_vs_dense(oracle):
    m = rect_mesh(4, 3, 5, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "D", "N"),
                  periodic_x=True)
    f = np.random.default_rng(9).standard_normal((m.nel, m.npnp))
    u = solve2d(oracle, m, 1.0 + 4.0, f)
    ur = dense_global_solve(m, 5.0, f)
    assert np.max(np.abs(u - ur)) < 1e-12 * np.abs(ur).max()


def unit_square(np_, **kw):
    return rect_mesh(2, 2, np_, 0.0, 1.0, 0.0, 1.0, **kw)


def test_session_helmholtz1(oracle):
    """mesh/helmholtz1: N_P=11, LAMBDA2=16, forcing sin(pi x) sin(pi y),
    exact -sin(pi x) sin(pi y)/(2 pi^2 + 16)."""
    m = unit_square(11)
    f = np.sin(np.pi * m.x) * np.sin(np.pi * m.y)
    u = solve2d(oracle, m, 16.0, f)
    ex = -f / (2 * np.pi ** 2 + 16.0)
    assert np.max(np.abs(u - ex)) < 1e-10


def test_session_poisson1(oracle):
    m = unit_square(11)
    f = np.sin(np.pi * m.x) * np.sin(np.pi * m.y)
    u = solve2d(oracle, m, 0.0, f)
    assert np.max(np.abs(u + f / (2 * np.pi ** 2))) < 1e-10


def test_session_laplace5_nonhomogeneous_dirichlet(oracle):
    m = unit_square(11)
    g = lambda x, y: np.sin(x) * np.exp(-y)  # noqa: E731
    u = solve2d(oracle, m, 0.0, np.zeros((m.nel, m.npnp)), boundary_lines(m, g))
    assert np.max(np.abs(u - g(m.x, m.y))) < 1e-11


def test_session_laplace7_mixed_dirichlet_neumann(oracle):
    """Essential on the top edge, natural (du/dn) on the other three."""
    m = unit_square(11, bc=("N", "N", "D", "N"))
    line = np.zeros(m.nbound * m.np)
    for b in range(m.nbound):
        xs, ys = m.side_xy(b)
        side = m.bnd_side[b]
        v = {0: np.sin(xs) * np.exp(-ys),      # bottom: -du/dy
             1: np.cos(xs) * np.exp(-ys),      # right:  du/dx
             2: np.sin(xs) * np.exp(-ys),      # top: Dirichlet value
             3: -np.cos(xs) * np.exp(-ys)}[side]
        line[b * m.np:(b + 1) * m.np] = v
    u = solve2d(oracle, m, 0.0, np.zeros((m.nel, m.npnp)), line)
    assert np.max(np.abs(u - np.sin(m.x) * np.exp(-m.y))) < 1e-10


def test_session_laplace4_neumann_top(oracle):
    m = unit_square(12, bc=("D", "D", "N", "D"))
    line = np.zeros(m.nbound * m.np)
    for b in range(m.nbound):
        if m.bnd_kind[b] == 1:
            xs, _ = m.side_xy(b)
            line[b * m.np:(b + 1) * m.np] = np.sin(np.pi * xs)
    u = solve2d(oracle, m, 0.0, np.zeros((m.nel, m.npnp)), line)
    ex = np.sin(np.pi * m.x) * np.sinh(np.pi * m.y) / (np.pi * np.cosh(np.pi))
    assert np.max(np.abs(u - ex)) < 1e-10


def test_square_4x4_np8_manufactured(oracle):
    """4x4 elements, N_P=8, lambda2=1, u = sin(pi x) sin(pi y)."""
    m = rect_mesh(4, 4, 8)
    ex = np.sin(np.pi * m.x) * np.sin(np.pi * m.y)
    f = -(2 * np.pi ** 2 + 1.0) * ex
    u = solve2d(oracle, m, 1.0, f)
    assert np.max(np.abs(u - ex)) < 2e-8
    ur = dense_global_solve(m, 1.0, f)
    assert np.max(np.abs(u - ur)) < 1e-13


def test_fourier_modes_3d_laplace(oracle):
    """2.5D: u = sin(pi x) sin(pi y) (cos 2z + sin 3z) solved mode by mode,
    exact per mode: -(2 pi^2 + k^2)^-1 (beta = 1)."""
    m = rect_mesh(2, 2, 10)
    nz = 8
    base = np.sin(np.pi * m.x) * np.sin(np.pi * m.y)
    planes = np.zeros((nz, m.nel * m.npnp))
    planes[4] = base.ravel() * 1.0   # mode 2, Re
    planes[7] = base.ravel() * 2.0   # mode 3, Im
    out = oracle.solve_modes(m, 0.0, 1.0, planes)
    assert np.max(np.abs(out[4] + base.ravel() / (2 * np.pi ** 2 + 4))) < 1e-10
    assert np.max(np.abs(out[7] + 2 * base.ravel() / (2 * np.pi ** 2 + 9))) < 1e-10
    assert np.max(np.abs(out[[0, 2, 3, 5, 6]])) < 1e-14


@pytest.mark.parametrize("np_,tol", [(8, 2e-9), (10, 5e-12)])
def test_cylindrical_axisymmetric_manufactured(np_, tol):
    """Cylindrical mode 0 (Q1..Q4 carry r, the forcing weight Q8 does not:
    src/element.cpp:800-806, the caller supplies r f): u = (1 - r^2) cos(pi x)
    on [0,2] x [0,1], axis on r = 0 (natural), u = 0 on r = 1, periodic x;
    converges spectrally to the exact solution."""
    from oracle.oracle import OracleMesh, OracleMsys
    from semtex_amd.mesh import rect_mesh
    m = rect_mesh(4, 3, np_, 0.0, 2.0, 0.0, 1.0, bc=("A", "N", "D", "N"), periodic_x=True, cyl=1)
    lam2 = 2.0
    u = (1 - m.y ** 2) * np.cos(np.pi * m.x)
    f = -np.pi ** 2 * u - 4 * np.cos(np.pi * m.x) - lam2 * u
    uo = OracleMsys(OracleMesh(m), lam2, 0.0).solve_plane((m.y * f).ravel())
    assert np.abs(uo - u.ravel()).max() < tol
