"""The twisted (two-piece) band solve against the single-piece sweeps and
the oracle.  SX_BAND_TWIST=1/0 forces the layout when a system is factored;
the layout query confirms which one ran.

* Helmholtz, 12x10 elements, N_P=8, 8 planes, nonzero Dirichlet and
  Neumann data: both layouts match the oracle to 1e-12 (FP64 bar);
* DNS channel, 8x10 elements, N_P=8, 8 planes (the grouped velocity solve
  with 6 right-hand sides per mode and the pressure solve): three steps
  with the twisted layout equal the single-piece ones to 1e-11, the DNS
  parity bar (the two factorisations eliminate in different orders; the
  pure-Neumann pressure solve differs by about 3e-12).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _lines(m, nz, seed):
    rng = np.random.default_rng(seed)
    lines = np.zeros((nz, m.nbound * m.np))
    for k in range(nz):
        if k == 1:
            continue
        for b in range(m.nbound):
            xs, _ = m.side_xy(b)
            lines[k, b * m.np:(b + 1) * m.np] = (np.cos((k + 1) * xs) if m.bnd_kind[b] == 0
                                                 else rng.standard_normal(m.np))
    return lines


@pytest.mark.parametrize("twist", ["1", "0"])
def test_twisted_helmholtz_matches_oracle(torch_cuda, oracle, monkeypatch, twist):
    import torch
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.problems import fourier_forcing
    from semtex_amd.solver import Field, GpuMesh, ModalMatrixSys
    print("case: helmholtz 12x10 np=8 nz=8 twist", twist, flush=True)
    monkeypatch.setenv("SX_BAND_TWIST", twist)
    m = rect_mesh(12, 10, 8, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "N", "N"), periodic_x=True)
    nz = 8
    F = fourier_forcing(m, nz, seed=5)
    lines = _lines(m, nz, 9)
    gm = GpuMesh(m)
    M = ModalMatrixSys(3.0, 1.0, 0, nz // 2, gm)
    lay = M.layout()
    assert lay["panels"] == 1 and lay["twisted"] == int(twist), lay
    u = Field(gm, nz)
    u.lines.copy_(torch.from_numpy(lines))
    u.solve(torch.from_numpy(F).cuda(), M)
    torch.cuda.synchronize()
    ug = u.data.cpu().numpy()
    n = m.nel * m.npnp
    uo = oracle.solve_modes(m, 3.0, 1.0, np.ascontiguousarray(F[:, :n]), lines)
    for k in range(nz):
        if k == 1:
            continue
        rel = np.linalg.norm(ug[k, :n] - uo[k]) / np.linalg.norm(uo[k])
        assert rel < 1e-12, (k, rel)


def _dns_run(torch, monkeypatch, twist, steps=3):
    from semtex_amd.mesh import edgemaps, rect_mesh
    from semtex_amd.solver import DNS, SX_SKEW, GpuMesh, ModalMatrixSys
    monkeypatch.setenv("SX_BAND_TWIST", twist)
    bcu, bcp = ("D", "N", "D", "N"), ("N", "N", "N", "N")
    mu = rect_mesh(8, 10, 8, 0.0, 2 * np.pi, -1.0, 1.0, bc=bcu, periodic_x=True)
    mp = rect_mesh(8, 10, 8, 0.0, 2 * np.pi, -1.0, 1.0, bc=bcp, periodic_x=True)
    nz, beta, nu, dt = 8, 1.0, 1.0 / 100, 2e-3
    gu, gp = GpuMesh(mu), GpuMesh(mp)
    # the same meshes' systems take the forced layout
    lay = ModalMatrixSys(1.0 / (dt * nu), beta, 0, nz // 2, gu).layout()
    assert lay["twisted"] == int(twist), lay
    D = DNS(gu, gp, nz, beta, nu, dt, const_force=(2 * nu, 0.0, 0.0), advection=SX_SKEW)
    n, S = mu.nel * mu.npnp, D.stride
    rng = np.random.default_rng(20261202)
    wall = 1.0 - mu.y.ravel() ** 2
    u0 = np.zeros((3, nz, S))
    em, _ = edgemaps(mu.np)
    for i in range(3):
        u0[i, :, :n] = (wall if i == 0 else 0.0) + 0.1 * rng.uniform(-1, 1, (nz, n)) * wall
        for k in range(nz):
            f = u0[i, k, :n].reshape(mu.nel, -1)
            g = np.zeros(mu.nglobal)
            g[mu.btog] = f[:, em[:mu.next]]
            f[:, em[:mu.next]] = g[mu.btog]
            u0[i, k, :n] = f.ravel()
    D.set_state(torch.from_numpy(u0).cuda())
    D.step(steps)
    torch.cuda.synchronize()
    return [D.field(i).cpu().numpy()[:, :n] for i in range(4)]


def test_twisted_dns_equals_single_piece(torch_cuda, monkeypatch):
    print("case: dns channel 8x10 np=8 nz=8, twisted vs single piece", flush=True)
    a = _dns_run(torch_cuda, monkeypatch, "1")
    b = _dns_run(torch_cuda, monkeypatch, "0")
    planes = [0] + list(range(2, 8))
    for i in range(4):
        rel = np.linalg.norm(a[i][planes] - b[i][planes]) / np.linalg.norm(b[i][planes])
        assert rel < 1e-11, (i, rel)
