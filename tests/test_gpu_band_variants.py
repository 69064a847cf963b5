"""Every panel-sweep variant of the band solve against the oracle at the
FP64 bar (1e-12), on systems whose layout query confirms the variant ran:

* SX_BAND_UNIT=0: inverted diagonal blocks in the chain (kernel 1);
* unit-diagonal panels W = T_J U (kernel 2), 1 and 2 right-hand sides per
  workgroup (SX_PANEL_NR);
* split sweeps, one chain on two workgroups exchanging records through L2
  (kernel 3, SX_BAND_SPLIT=1), 1 and 2 right-hand sides per workgroup;
each with the twisted and the single-piece layout, for D = 5 (16x16
elements, N_P=10) and D = 3 (12x10 elements, N_P=8), with nonzero Dirichlet
and Neumann data.  Repeated split solves (each launch advances the chains'
epochs) give bitwise-equal results, and DNS steps whose grouped velocity
solve runs on split sweeps with 3 right-hand sides per workgroup equal the
default path to 1e-11 (the DNS parity bar).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VARIANTS = [  # (SX_BAND_UNIT, SX_BAND_SPLIT, SX_PANEL_NR, expected kernel)
    ("0", "0", "1", 1),
    ("1", "0", "1", 2),
    ("1", "0", "2", 2),
    ("1", "1", "1", 3),
    ("1", "1", "2", 3),
]


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
            lines[k, b * m.np:(b + 1) * m.np] = (np.sin((k + 2) * xs) if m.bnd_kind[b] == 0
                                                 else rng.standard_normal(m.np))
    return lines


@pytest.fixture(scope="module")
def problems(oracle):
    """Meshes, forcing, boundary data and oracle solutions, computed once."""
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.problems import fourier_forcing
    out = {}
    for key, (nx, ny, np_, D) in {"D5": (16, 16, 10, 5), "D3": (12, 10, 8, 3)}.items():
        m = rect_mesh(nx, ny, np_, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "N", "N"),
                      periodic_x=True)
        nz = 4
        F = fourier_forcing(m, nz, seed=50 + D)
        lines = _lines(m, nz, 60 + D)
        n = m.nel * m.npnp
        uo = oracle.solve_modes(m, 2.5, 1.0, np.ascontiguousarray(F[:, :n]), lines)
        out[key] = (m, nz, F, lines, uo, D)
    return out


def _solve(torch, m, nz, F, lines, reps=1):
    from semtex_amd.solver import Field, GpuMesh, ModalMatrixSys
    gm = GpuMesh(m)
    M = ModalMatrixSys(2.5, 1.0, 0, nz // 2, gm)
    u = Field(gm, nz)
    u.lines.copy_(torch.from_numpy(lines))
    sols = []
    for _ in range(reps):
        u.solve(torch.from_numpy(F).cuda(), M)
        torch.cuda.synchronize()
        sols.append(u.data.cpu().numpy().copy())
    return sols, M.layout()


@pytest.mark.parametrize("key", ["D5", "D3"])
@pytest.mark.parametrize("twist", ["1", "0"])
@pytest.mark.parametrize("unit,split,nr,kernel", VARIANTS)
def test_band_variant_matches_oracle(torch_cuda, problems, monkeypatch, key, twist, unit, split,
                                     nr, kernel):
    m, nz, F, lines, uo, D = problems[key]
    print("case:", key, "twist", twist, "unit", unit, "split", split, "nr", nr, flush=True)
    monkeypatch.setenv("SX_BAND_TWIST", twist)
    monkeypatch.setenv("SX_BAND_UNIT", unit)
    monkeypatch.setenv("SX_BAND_SPLIT", split)
    monkeypatch.setenv("SX_PANEL_NR", nr)
    (ug,), lay = _solve(torch_cuda, m, nz, F, lines)
    assert lay["D"] == D and lay["twisted"] == int(twist), lay
    assert lay["kernel"] == kernel and lay["nr"] == int(nr), lay
    n = m.nel * m.npnp
    for k in (0, 2, 3):
        rel = np.linalg.norm(ug[k, :n] - uo[k]) / np.linalg.norm(uo[k])
        assert rel < 1e-12, (k, rel)


def test_split_repeated_solves_are_bitwise_equal(torch_cuda, problems, monkeypatch):
    m, nz, F, lines, _, _ = problems["D5"]
    print("case: D5 split, 4 solves in a row", flush=True)
    monkeypatch.setenv("SX_BAND_SPLIT", "1")
    sols, lay = _solve(torch_cuda, m, nz, F, lines, reps=4)
    assert lay["kernel"] == 3, lay
    for s in sols[1:]:
        assert np.array_equal(s, sols[0])


def _dns(torch, monkeypatch, env):
    from semtex_amd.problems import dns_channel
    from semtex_amd.solver import DNS, SX_SKEW, GpuMesh
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    P = dns_channel(nz=8)
    gu, gp = GpuMesh(P["mesh_u"]), GpuMesh(P["mesh_p"])
    D = DNS(gu, gp, P["nz"], P["beta"], P["kinvis"], P["dt"], const_force=P["const_force"],
            advection=SX_SKEW)
    n = P["mesh_u"].nel * P["mesh_u"].npnp
    u0 = torch.zeros((3, P["nz"], D.stride), dtype=torch.float64)
    u0[:, :, :n] = torch.from_numpy(P["u0"])
    D.set_state(u0.cuda())
    D.step(3)
    torch.cuda.synchronize()
    return [D.field(i).cpu().numpy()[:, :n] for i in range(4)]


def test_dns_split_three_rhs_per_workgroup(torch_cuda, monkeypatch):
    print("case: dns channel 12x24 np=8 nz=8, split sweeps NR=3 vs default", flush=True)
    a = _dns(torch_cuda, monkeypatch, {"SX_BAND_SPLIT": "1", "SX_PANEL_NR": "3"})
    monkeypatch.delenv("SX_BAND_SPLIT")
    monkeypatch.delenv("SX_PANEL_NR")
    b = _dns(torch_cuda, monkeypatch, {})
    planes = [0] + list(range(2, 8))
    for i in range(4):
        rel = np.linalg.norm(a[i][planes] - b[i][planes]) / np.linalg.norm(b[i][planes])
        assert rel < 1e-11, (i, rel)


def test_split_wait_failure_reports_status_then_recovers(torch_cuda, problems, monkeypatch):
    """The split sweeps' bounded waits, forced from the host to give up at
    once (SX_SPLIT_SPIN=-1): the kernels still exit, the solution is NaN,
    sx_msys_status reports SX_ERR_SYNC once (then clears), and the next
    solves on the same system with the normal bound are exact: every
    workgroup of the failed launch used one epoch and the epoch advanced
    only after all of them finished, so no record of the failed launch
    matches a later one."""
    from semtex_amd import _lib
    from semtex_amd.solver import Field, GpuMesh, ModalMatrixSys
    m, nz, F, lines, uo, _ = problems["D5"]
    print("case: D5 split, forced wait failure, status, then normal solves", flush=True)
    monkeypatch.setenv("SX_BAND_SPLIT", "1")
    gm = GpuMesh(m)
    M = ModalMatrixSys(2.5, 1.0, 0, nz // 2, gm)
    u = Field(gm, nz)
    u.lines.copy_(torch_cuda.from_numpy(lines))
    n = m.nel * m.npnp
    M.status()                                        # clean before anything ran
    monkeypatch.setenv("SX_SPLIT_SPIN", "-1")
    u.solve(torch_cuda.from_numpy(F).cuda(), M)
    torch_cuda.cuda.synchronize()
    assert M.layout()["kernel"] == 3, M.layout()
    assert np.isnan(u.data.cpu().numpy()[0, :n]).any()
    with pytest.raises(_lib.SxError) as ei:
        M.status()
    assert ei.value.code == _lib.SX_ERR_SYNC
    M.status()                                        # reported once, then cleared
    monkeypatch.delenv("SX_SPLIT_SPIN")
    for rep in range(3):
        u.solve(torch_cuda.from_numpy(F).cuda(), M)
        torch_cuda.cuda.synchronize()
        M.status()
        good = u.data.cpu().numpy()
        for k in (0, 2, 3):
            rel = np.linalg.norm(good[k, :n] - uo[k]) / np.linalg.norm(uo[k])
            assert rel < 1e-12, (rep, k, rel)


def test_split_status_through_host_entry_point(torch_cuda, problems, monkeypatch):
    """sx_field_solve_host checks the status itself: a forced wait failure
    returns SX_ERR_SYNC instead of SX_OK."""
    import ctypes as C

    from semtex_amd import _lib
    from semtex_amd.solver import GpuMesh, ModalMatrixSys
    m, nz, F, lines, uo, _ = problems["D5"]
    print("case: D5 split, sx_field_solve_host with a forced wait failure", flush=True)
    monkeypatch.setenv("SX_BAND_SPLIT", "1")
    gm = GpuMesh(m)
    M = ModalMatrixSys(2.5, 1.0, 0, nz // 2, gm)
    L = _lib.load()
    f = np.ascontiguousarray(F)
    u = np.zeros_like(f)
    ln = np.ascontiguousarray(lines)
    p = lambda a: C.c_void_p(a.ctypes.data)  # noqa: E731
    monkeypatch.setenv("SX_SPLIT_SPIN", "-1")
    rc = L.sx_field_solve_host(M.handle, nz, f.shape[1], p(f), p(u), p(ln))
    assert rc == _lib.SX_ERR_SYNC, rc
    monkeypatch.delenv("SX_SPLIT_SPIN")
    rc = L.sx_field_solve_host(M.handle, nz, f.shape[1], p(f), p(u), p(ln))
    assert rc == _lib.SX_OK, L.sx_last_error()
    n = m.nel * m.npnp
    for k in (0, 2, 3):
        rel = np.linalg.norm(u[k, :n] - uo[k]) / np.linalg.norm(uo[k])
        assert rel < 1e-12, (k, rel)
