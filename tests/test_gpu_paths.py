"""GPU parity (1e-12, the FP64 bar) of the solve paths the configurations
do not reach, against the oracle on the same inputs:

* wide bands: 28x8 elements, N_P=10 (nband 495, D = 8 panel blocks), both
  the twisted and the single-piece layout;
* N_P = 14 (element matrices too large for the LDS-tiled condensation and
  recovery: the per-element kernels);
* the offset-major band sweeps (SX_BAND_SWEEP=offset);
* a system of modes 2..3 (base mode 2: every plane is solved, no Nyquist
  skip, betak2 = (beta (2 + m))^2).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _solve(torch, m, lam2, beta, nz, F, base_mode=0, lines=None):
    from semtex_amd.solver import Field, GpuMesh, ModalMatrixSys
    gm = GpuMesh(m)
    nmodes = nz // 2 if nz > 1 else 1
    M = ModalMatrixSys(lam2, beta, base_mode, nmodes, gm)
    u = Field(gm, nz)
    if lines is not None:
        u.lines.copy_(torch.from_numpy(lines))
    u.solve(torch.from_numpy(np.ascontiguousarray(F)).cuda(), M)
    torch.cuda.synchronize()
    return u.data.cpu().numpy(), M.layout()


def _check(oracle, m, lam2, beta, F, ug, planes, base_mode=0, lines=None):
    from oracle.oracle import OracleMesh, OracleMsys
    om = OracleMesh(m)
    n = m.nel * m.npnp
    for k in planes:
        mode = base_mode + (k >> 1)
        S = OracleMsys(om, lam2, (beta * mode) ** 2)
        uo = S.solve_plane(F[k, :n], None if lines is None else lines[k])
        rel = np.linalg.norm(ug[k, :n] - uo) / np.linalg.norm(uo)
        assert rel < TOL, (k, rel)


@pytest.mark.parametrize("twist", ["1", "0"])
def test_wide_band_eight_panel_blocks(torch_cuda, oracle, monkeypatch, twist):
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.problems import fourier_forcing
    print("case: 28x8 np=10 nz=4 (D=8), twist", twist, flush=True)
    monkeypatch.setenv("SX_BAND_TWIST", twist)
    m = rect_mesh(28, 8, 10, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "D", "N"), periodic_x=True)
    F = fourier_forcing(m, 4, seed=31)
    ug, lay = _solve(torch_cuda, m, 1.5, 1.0, 4, F)
    assert lay["panels"] == 1 and lay["D"] == 8 and lay["twisted"] == int(twist), lay
    _check(oracle, m, 1.5, 1.0, F, ug, [0, 2, 3])


def test_np14_per_element_kernels(torch_cuda, oracle):
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.problems import fourier_forcing
    print("case: 3x2 np=14 nz=4", flush=True)
    m = rect_mesh(3, 2, 14, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "D", "N"), periodic_x=True)
    F = fourier_forcing(m, 4, seed=32)
    ug, _ = _solve(torch_cuda, m, 2.0, 1.0, 4, F)
    _check(oracle, m, 2.0, 1.0, F, ug, [0, 2, 3])


def test_offset_major_sweeps(torch_cuda, oracle, monkeypatch):
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.problems import fourier_forcing
    print("case: 6x4 np=7 nz=8 offset-major sweeps", flush=True)
    monkeypatch.setenv("SX_BAND_SWEEP", "offset")
    m = rect_mesh(6, 4, 7, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "D", "N"), periodic_x=True)
    F = fourier_forcing(m, 8, seed=33)
    ug, lay = _solve(torch_cuda, m, 2.0, 1.0, 8, F)
    assert lay["panels"] == 0, lay
    _check(oracle, m, 2.0, 1.0, F, ug, [0] + list(range(2, 8)))


def test_base_mode_two_solves_every_plane(torch_cuda, oracle):
    from semtex_amd.mesh import rect_mesh
    print("case: 6x4 np=7 base mode 2, 2 modes", flush=True)
    m = rect_mesh(6, 4, 7, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "N", "N"), periodic_x=True)
    rng = np.random.default_rng(34)
    n = m.nel * m.npnp
    F = np.zeros((4, m.plane_size))
    F[:, :n] = rng.standard_normal((4, n))
    lines = np.zeros((4, m.nbound * m.np))
    for k in range(4):
        for b in range(m.nbound):
            xs, _ = m.side_xy(b)
            lines[k, b * m.np:(b + 1) * m.np] = (np.sin((k + 1) * xs) if m.bnd_kind[b] == 0
                                                 else rng.standard_normal(m.np))
    ug, _ = _solve(torch_cuda, m, 0.5, 1.0, 4, F, base_mode=2, lines=lines)
    _check(oracle, m, 0.5, 1.0, F, ug, [0, 1, 2, 3], base_mode=2, lines=lines)


@pytest.mark.parametrize("mfma", ["1", "0"])
def test_tiled_condensation_mfma_and_vector(torch_cuda, oracle, monkeypatch, mfma):
    """The class-tiled condensation/recovery on FP64 MFMA tiles (default where
    the 16-padded operands fit the LDS) and on the vector FMA tiles
    (SX_TILE_MFMA=0), N_P = 7 and 11, with Dirichlet and Neumann data."""
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.problems import fourier_forcing
    monkeypatch.setenv("SX_TILE_MFMA", mfma)
    for np_ in (7, 11):
        print("case: tiled kernels mfma", mfma, "np", np_, flush=True)
        m = rect_mesh(5, 3, np_, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "N", "N"),
                      periodic_x=True)
        F = fourier_forcing(m, 4, seed=40 + np_)
        rng = np.random.default_rng(np_)
        lines = np.zeros((4, m.nbound * m.np))
        for k in range(4):
            for b in range(m.nbound):
                xs, _ = m.side_xy(b)
                lines[k, b * m.np:(b + 1) * m.np] = (np.cos((k + 1) * xs) if m.bnd_kind[b] == 0
                                                     else rng.standard_normal(m.np))
        ug, _ = _solve(torch_cuda, m, 1.5, 1.0, 4, F, lines=lines)
        _check(oracle, m, 1.5, 1.0, F, ug, [0, 2, 3], lines=lines)
