"""Cylindrical 3D coupling on the GPU (sx_couple, sx_msys_create_named,
per-mode-class numbering) against the oracle:

* sx_couple on field planes and on boundary lines equals the oracle's
  couple (AuxField::couple, Field::coupleBCs FORWARD) bit for bit, and its
  inverse restores the input;
* the solve of u, v~, w~ and p over their mode-class numberings with
  Field::modeConstant's Fourier constant equals the oracle to 1e-12 and
  reproduces the manufactured solutions of the oracle's tests;
* the coupled viscous substep of the radial and azimuthal velocities --
  couple the forcing and the boundary lines, solve v~ and w~, uncouple
  (dns/integrate.cpp:185-195) -- equals the same sequence in the oracle.
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


def _cuda(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _planes(mesh, nz, rng):
    F = np.zeros((nz, mesh.plane_size))
    n = mesh.nel * mesh.npnp
    for k in range(nz):
        if k != 1:
            F[k, :n] = rng.standard_normal(n)
    return F


def test_couple_matches_oracle_on_fields_and_lines(torch_cuda):
    import torch
    from oracle import oracle as O
    from semtex_amd.solver import SX_FORWARD, SX_INVERSE, couple
    print("case: sx_couple, nz=16, fields (n < stride) and lines", flush=True)
    rng = np.random.default_rng(11)
    nz, n, S = 16, 300, 320
    V, W = rng.standard_normal((nz, S)), rng.standard_normal((nz, S))
    v, w = _cuda(torch, V), _cuda(torch, W)
    couple(v, w, SX_FORWARD, n=n)
    Vo, Wo = O.couple(V[:, :n], W[:, :n], +1)
    assert np.array_equal(v.cpu().numpy()[:, :n], Vo)
    assert np.array_equal(w.cpu().numpy()[:, :n], Wo)
    assert np.array_equal(v.cpu().numpy()[:, n:], V[:, n:])      # past n untouched
    couple(v, w, SX_INVERSE, n=n)
    assert np.max(np.abs(v.cpu().numpy() - V)) < 1e-15
    assert np.max(np.abs(w.cpu().numpy() - W)) < 1e-15
    # boundary lines [nz][nline], base mode 2 (every mode coupled)
    L1, L2 = rng.standard_normal((8, 40)), rng.standard_normal((8, 40))
    a, b = _cuda(torch, L1), _cuda(torch, L2)
    couple(a, b, SX_FORWARD, base_mode=2)
    Lo1, Lo2 = O.couple(L1, L2, +1, base_mode=2)
    assert np.array_equal(a.cpu().numpy(), Lo1) and np.array_equal(b.cpu().numpy(), Lo2)


@pytest.mark.parametrize("name", ["u", "v", "w", "p"])
def test_named_modal_solve_matches_oracle_and_manufactured(torch_cuda, name):
    import torch
    from oracle import oracle as O
    from semtex_amd.problems import cylindrical_meshes
    from semtex_amd.solver import CylModalSys, GpuMesh
    from test_oracle_cylindrical import manufactured
    print(f"case: cylindrical modal solve '{name}', 4x3 np=11 nz=8", flush=True)
    nz, lam2 = 8, 2.0
    meshes = cylindrical_meshes(name, nx=4, ny=3, np_=11)
    gms = {id(m): GpuMesh(m) for m in meshes}
    M = CylModalSys(name, lam2, 1.0, nz // 2, [gms[id(m)] for m in meshes])
    m0 = meshes[0]
    n = m0.nel * m0.npnp
    U, F = manufactured(meshes, name, nz, lam2)
    Fp = np.zeros((nz, m0.plane_size))
    Fp[:, :n] = F
    u = torch.zeros((nz, m0.plane_size), dtype=torch.float64, device="cuda")
    M.solve(_cuda(torch, Fp), u)
    torch.cuda.synchronize()
    got = u.cpu().numpy()[:, :n]
    ref = O.solve_field_cyl(meshes, name, lam2, 1.0, F)
    for k in range(nz):
        if k == 1:
            continue
        rel = np.linalg.norm(got[k] - ref[k]) / np.linalg.norm(ref[k])
        assert rel < TOL, (name, k, rel)
        assert np.max(np.abs(got[k] - U[k])) < 1e-12, (name, k)


def test_coupled_viscous_substep_matches_oracle(torch_cuda):
    """Random forcing and boundary data on the wall: couple F_v, F_w and
    the wall lines, solve v~ (k+1) and w~ (k-1) on their numberings,
    uncouple; nz = 12 (modes 0..5, base mode 0)."""
    import torch
    from oracle import oracle as O
    from semtex_amd.problems import cylindrical_meshes
    from semtex_amd.solver import SX_FORWARD, SX_INVERSE, CylModalSys, GpuMesh, couple
    print("case: coupled v~/w~ viscous substep, 4x3 np=9 nz=12", flush=True)
    nz, lam2 = 12, 30.0
    rng = np.random.default_rng(5)
    mv = cylindrical_meshes("v", nx=4, ny=3, np_=9)
    mw = cylindrical_meshes("w", nx=4, ny=3, np_=9)
    gm = {}
    for m in mv + mw:
        gm.setdefault(id(m), GpuMesh(m))
    Mv = CylModalSys("v", lam2, 1.0, nz // 2, [gm[id(m)] for m in mv])
    Mw = CylModalSys("w", lam2, 1.0, nz // 2, [gm[id(m)] for m in mw])
    m0 = mv[0]
    n = m0.nel * m0.npnp
    Fv, Fw = _planes(m0, nz, rng), _planes(m0, nz, rng)
    nl = m0.nbound * m0.np
    # values on every boundary segment; only the essential ones are used
    Lv, Lw = rng.standard_normal((nz, nl)), rng.standard_normal((nz, nl))
    Lv[1] = Lw[1] = 0.0
    fv, fw, lv, lw = (_cuda(torch, a) for a in (Fv, Fw, Lv, Lw))
    couple(fv, fw, SX_FORWARD, n=n)
    couple(lv, lw, SX_FORWARD)
    uv = torch.zeros_like(fv)
    uw = torch.zeros_like(fw)
    Mv.solve(fv, uv, lv)
    Mw.solve(fw, uw, lw)
    couple(uv, uw, SX_INVERSE, n=n)
    torch.cuda.synchronize()
    # the oracle, step by step
    Fvo, Fwo = O.couple(Fv[:, :n], Fw[:, :n], +1)
    Lvo, Lwo = O.couple(Lv, Lw, +1)
    Vt = O.solve_field_cyl(mv, "v", lam2, 1.0, Fvo, Lvo)
    Wt = O.solve_field_cyl(mw, "w", lam2, 1.0, Fwo, Lwo)
    Vo, Wo = O.couple(Vt, Wt, -1)
    planes = [k for k in range(nz) if k != 1]
    for got, ref in ((uv, Vo), (uw, Wo)):
        g = got.cpu().numpy()[planes, :n]
        rel = np.linalg.norm(g - ref[planes]) / np.linalg.norm(ref[planes])
        assert rel < TOL, rel
