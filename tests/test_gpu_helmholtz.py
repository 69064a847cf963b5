"""GPU parity of the FP64 Helmholtz solve (sx_field_solve) against the CPU
oracle, on the same seeded inputs.  Bar: relative L2 <= 1e-12 (FP64 path).

Cases: the 2D unit-square configuration (4x4, N_P=8) and the reference's
analytic elliptic sessions (homogeneous / nonhomogeneous Dirichlet,
Neumann); small multi-mode Fourier problems with boundary data; a singular
(pure-Neumann, lambda2=0) mode 0; the full 16x16 N_P=10 N_Z=64 batch
against the oracle on selected modes, plus the residual of every mode
through the library's own elemental operator.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def gpu_solve(torch, mesh, lambda2, beta, nz, nmodes, F, lines=None):
    from semtex_amd.solver import Field, GpuMesh, ModalMatrixSys
    gm = GpuMesh(mesh)
    M = ModalMatrixSys(lambda2, beta, 0, nmodes, gm)
    u = Field(gm, nz)
    if lines is not None:
        u.lines.copy_(torch.from_numpy(lines))
    f = torch.from_numpy(np.ascontiguousarray(F)).cuda()
    u.solve(f, M)
    torch.cuda.synchronize()
    return u.data.cpu().numpy(), gm, M


def oracle_solve(oracle, mesh, lambda2, beta, nz, F, lines=None):
    n = mesh.nel * mesh.npnp
    planes = np.ascontiguousarray(F[:, :n])
    return oracle.solve_modes(mesh, lambda2, beta, planes, lines)


def test_square_4x4_np8_parity(torch_cuda, oracle):
    from semtex_amd.problems import helmholtz_2d_square
    P = helmholtz_2d_square()
    m = P["mesh"]
    print("case: 4x4 np=8 lambda2=1", flush=True)
    F = np.zeros((1, m.plane_size))
    F[0, :m.nel * m.npnp] = P["forcing"][0]
    ug, _, _ = gpu_solve(torch_cuda, m, 1.0, 0.0, 1, 1, F)
    uo = oracle_solve(oracle, m, 1.0, 0.0, 1, F)
    n = m.nel * m.npnp
    assert rel_l2(ug[0, :n], uo[0]) < TOL
    assert np.max(np.abs(ug[0, :n] - P["exact"][0])) < 2e-8


@pytest.mark.parametrize("case", ["helmholtz1", "laplace5", "laplace7", "laplace4"])
def test_elliptic_sessions_parity(torch_cuda, oracle, case):
    from semtex_amd.mesh import rect_mesh
    from tests.helpers import boundary_lines
    print("case:", case, flush=True)
    if case == "helmholtz1":
        m = rect_mesh(2, 2, 11)
        f = np.sin(np.pi * m.x) * np.sin(np.pi * m.y)
        lam2, line = 16.0, None
        exact = -f / (2 * np.pi ** 2 + 16)
    elif case == "laplace5":
        m = rect_mesh(2, 2, 11)
        f = np.zeros((m.nel, m.npnp))
        g = lambda x, y: np.sin(x) * np.exp(-y)  # noqa: E731
        lam2, line, exact = 0.0, boundary_lines(m, g), g(m.x, m.y)
    elif case == "laplace7":
        m = rect_mesh(2, 2, 11, bc=("N", "N", "D", "N"))
        f = np.zeros((m.nel, m.npnp))
        line = np.zeros(m.nbound * m.np)
        for b in range(m.nbound):
            xs, ys = m.side_xy(b)
            line[b * m.np:(b + 1) * m.np] = {
                0: np.sin(xs) * np.exp(-ys), 1: np.cos(xs) * np.exp(-ys),
                2: np.sin(xs) * np.exp(-ys), 3: -np.cos(xs) * np.exp(-ys)}[int(m.bnd_side[b])]
        lam2, exact = 0.0, np.sin(m.x) * np.exp(-m.y)
    else:
        m = rect_mesh(2, 2, 12, bc=("D", "D", "N", "D"))
        f = np.zeros((m.nel, m.npnp))
        line = np.zeros(m.nbound * m.np)
        for b in range(m.nbound):
            if m.bnd_kind[b] == 1:
                xs, _ = m.side_xy(b)
                line[b * m.np:(b + 1) * m.np] = np.sin(np.pi * xs)
        lam2 = 0.0
        exact = np.sin(np.pi * m.x) * np.sinh(np.pi * m.y) / (np.pi * np.cosh(np.pi))
    n = m.nel * m.npnp
    F = np.zeros((1, m.plane_size))
    F[0, :n] = f.ravel()
    L = None if line is None else line.reshape(1, -1)
    ug, _, _ = gpu_solve(torch_cuda, m, lam2, 0.0, 1, 1, F, L)
    uo = oracle_solve(oracle, m, lam2, 0.0, 1, F, L)
    assert rel_l2(ug[0, :n], uo[0]) < TOL
    assert np.max(np.abs(ug[0, :n] - exact.ravel())) < 1e-9


@pytest.mark.parametrize("ordering", ["layer_y", "rcm"])
def test_fourier_modes_with_boundary_data(torch_cuda, oracle, ordering):
    """6x4 periodic mesh, N_P=7, 8 planes (4 modes), nonzero Dirichlet and
    Neumann data on every plane."""
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.problems import fourier_forcing
    print("case: fourier 6x4 np=7 nz=8", ordering, flush=True)
    m = rect_mesh(6, 4, 7, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "N", "N"),
                  periodic_x=True, ordering=ordering)
    nz = 8
    F = fourier_forcing(m, nz, seed=7)
    rng = np.random.default_rng(11)
    lines = np.zeros((nz, m.nbound * m.np))
    for k in range(nz):
        if k == 1:
            continue
        for b in range(m.nbound):
            xs, ys = m.side_xy(b)
            if m.bnd_kind[b] == 0:
                lines[k, b * m.np:(b + 1) * m.np] = np.cos((k + 1) * xs) * (1 + 0.1 * k)
            else:
                lines[k, b * m.np:(b + 1) * m.np] = rng.standard_normal(m.np)
    ug, _, _ = gpu_solve(torch_cuda, m, 2.0, 1.0, nz, nz // 2, F, lines)
    uo = oracle_solve(oracle, m, 2.0, 1.0, nz, F, lines)
    n = m.nel * m.npnp
    for k in range(nz):
        if k == 1:
            continue
        assert rel_l2(ug[k, :n], uo[k]) < TOL, k


def test_singular_neumann_mode0(torch_cuda, oracle):
    """lambda2 = 0, natural BCs everywhere: mode 0 is singular and the last
    unknown is dropped (src/matrix.cpp:154-162); mode 1 is regular."""
    from semtex_amd.mesh import rect_mesh
    print("case: singular neumann", flush=True)
    m = rect_mesh(3, 3, 6, bc=("N", "N", "N", "N"))
    nz = 4
    rng = np.random.default_rng(3)
    F = np.zeros((nz, m.plane_size))
    n = m.nel * m.npnp
    for k in (0, 2, 3):
        F[k, :n] = rng.standard_normal(n)
    ug, _, M = gpu_solve(torch_cuda, m, 0.0, 1.0, nz, 2, F)
    assert M.nsolve(0) == m.nsolve - 1 and M.nsolve(1) == m.nsolve
    uo = oracle_solve(oracle, m, 0.0, 1.0, nz, F)
    for k in (0, 2, 3):
        assert rel_l2(ug[k, :n], uo[k]) < 1e-11, k


def test_indefinite_system_reports_factor_error(torch_cuda):
    from semtex_amd import _lib
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.solver import GpuMesh, ModalMatrixSys
    m = rect_mesh(2, 2, 5)
    gm = GpuMesh(m)
    with pytest.raises(_lib.SxError):
        ModalMatrixSys(-400.0, 0.0, 0, 1, gm)


def test_full_fourier_batch_parity_and_residual(torch_cuda, oracle):
    """16x16 elements, N_P=10, N_Z=64: oracle parity on modes 0, 1, 5, 31;
    every mode's residual |A u - b| through the elemental operator."""
    import torch
    from semtex_amd.problems import helmholtz_fourier
    from semtex_amd.solver import Field, helmholtz_op
    print("case: 16x16 np=10 nz=64 full batch", flush=True)
    P = helmholtz_fourier()
    m, nz = P["mesh"], P["nz"]
    F = P["forcing"]
    ug, gm, M = gpu_solve(torch, m, P["lambda2"], P["beta"], nz, P["nmodes"], F)
    n = m.nel * m.npnp
    sel = [0, 2, 3, 10, 11, 62, 63]
    planes = np.zeros((nz, n))
    planes[sel] = F[sel, :n]
    from oracle.oracle import OracleMesh, OracleMsys
    om = OracleMesh(m)
    for k in sel:
        mode = k >> 1
        S = OracleMsys(om, P["lambda2"], (P["beta"] * mode) ** 2)
        uo = S.solve_plane(F[k, :n])
        assert rel_l2(ug[k, :n], uo) < TOL, k
    # residual of every plane: weak form  H u = -M f  on the unknowns,
    # assembled with the library's elemental operator (direct stiffness sum)
    u_t = torch.from_numpy(ug).cuda()
    hu = torch.zeros_like(u_t)
    helmholtz_op(M, u_t, hu)
    hu = hu.cpu().numpy()[:, :n].reshape(nz, m.nel, m.npnp)
    from semtex_amd.mesh import edgemaps
    em, _ = edgemaps(m.np)
    q8 = np.array([oracle.mapping(m.np, m.x[e], m.y[e])[1][oracle.Q8] for e in range(m.nel)])
    for k in range(nz):
        if k == 1:
            continue
        r = hu[k] + q8 * F[k, :n].reshape(m.nel, m.npnp)
        rb = np.zeros(m.nglobal)
        np.add.at(rb, m.btog, r[:, em[:m.next]])
        resid = np.concatenate([rb[:m.nsolve], r[:, em[m.next:]].ravel()])
        scale = np.linalg.norm((q8 * F[k, :n].reshape(m.nel, m.npnp)).ravel())
        assert np.linalg.norm(resid) < 1e-11 * scale, k


def test_mixed_precision_refinement_reaches_fp64(torch_cuda, oracle):
    """SX_MIXED: FP32 elemental condensation/recovery + FP64 band solve +
    FP64 residual refinement agrees with the FP64 path to <= 1e-10 relative
    (the mixed path's accuracy bar) on the full 16x16 N_P=10 N_Z=64 batch, and with
    boundary data on a small periodic problem."""
    import torch
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.problems import fourier_forcing, helmholtz_fourier
    from semtex_amd.solver import SX_MIXED, Field, GpuMesh, ModalMatrixSys
    print("case: mixed precision 16x16 np=10 nz=64", flush=True)
    P = helmholtz_fourier()
    m, nz = P["mesh"], P["nz"]
    gm = GpuMesh(m)
    M64 = ModalMatrixSys(P["lambda2"], P["beta"], 0, P["nmodes"], gm)
    M32 = ModalMatrixSys(P["lambda2"], P["beta"], 0, P["nmodes"], gm, precision=SX_MIXED)
    F = torch.from_numpy(P["forcing"]).cuda()
    u64, u32 = Field(gm, nz), Field(gm, nz)
    u64.solve(F.clone(), M64)
    u32.solve(F.clone(), M32, mixed_tol=1e-13, max_iter=10)
    torch.cuda.synchronize()
    n = m.nel * m.npnp
    a, b = u32.data.cpu().numpy()[:, :n], u64.data.cpu().numpy()[:, :n]
    planes = [k for k in range(nz) if k != 1]
    rel = np.linalg.norm(a[planes] - b[planes]) / np.linalg.norm(b[planes])
    print("mixed iterations", u32.last_iters, "rel", rel, flush=True)
    assert rel < 1e-10
    assert 1 <= u32.last_iters <= 10
    # boundary data (essential + natural) on a small periodic problem
    m2 = rect_mesh(6, 4, 7, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "N", "N"), periodic_x=True)
    gm2 = GpuMesh(m2)
    nz2 = 8
    F2 = fourier_forcing(m2, nz2, seed=3)
    lines = np.zeros((nz2, m2.nbound * m2.np))
    for k in range(nz2):
        if k == 1:
            continue
        for bb in range(m2.nbound):
            xs, ys = m2.side_xy(bb)
            lines[k, bb * m2.np:(bb + 1) * m2.np] = np.cos((k + 1) * xs) + 0.5 * (m2.bnd_kind[bb] == 1)
    ua, ub = Field(gm2, nz2), Field(gm2, nz2)
    ua.lines.copy_(torch.from_numpy(lines))
    ub.lines.copy_(torch.from_numpy(lines))
    Ma = ModalMatrixSys(2.0, 1.0, 0, nz2 // 2, gm2)
    Mb = ModalMatrixSys(2.0, 1.0, 0, nz2 // 2, gm2, precision=SX_MIXED)
    ua.solve(torch.from_numpy(F2).cuda(), Ma)
    ub.solve(torch.from_numpy(F2).cuda(), Mb, mixed_tol=1e-13)
    torch.cuda.synchronize()
    n2 = m2.nel * m2.npnp
    a2, b2 = ub.data.cpu().numpy()[:, :n2], ua.data.cpu().numpy()[:, :n2]
    pl = [k for k in range(nz2) if k != 1]
    assert np.linalg.norm(a2[pl] - b2[pl]) / np.linalg.norm(b2[pl]) < 1e-10


def test_cylindrical_mode0_parity(torch_cuda, oracle):
    """Axisymmetric (cylindrical, mode 0) problem with the axis on y = 0:
    GPU vs oracle to 1e-12, and the manufactured solution to 5e-12."""
    from semtex_amd.mesh import rect_mesh
    print("case: cylindrical mode 0", flush=True)
    m = rect_mesh(4, 3, 10, 0.0, 2.0, 0.0, 1.0, bc=("A", "N", "D", "N"), periodic_x=True, cyl=1)
    lam2 = 2.0
    u = (1 - m.y ** 2) * np.cos(np.pi * m.x)
    f = -np.pi ** 2 * u - 4 * np.cos(np.pi * m.x) - lam2 * u
    n = m.nel * m.npnp
    F = np.zeros((1, m.plane_size))
    F[0, :n] = (m.y * f).ravel()
    ug, _, _ = gpu_solve(torch_cuda, m, lam2, 0.0, 1, 1, F)
    uo = oracle_solve(oracle, m, lam2, 0.0, 1, F)
    assert rel_l2(ug[0, :n], uo[0]) < TOL
    assert np.abs(ug[0, :n] - u.ravel()).max() < 5e-12
