"""GPU DNS step (sx_dns_*) against the reference's own regression goldens
and, step by step, against the DNS oracle.

* regress/taylor2.ok and taylor4.ok: 20 steps of 2D / 3D Taylor flow with
  the default alternating skew-symmetric advection on the GPU reproduce
  norm_inf(exact - u) to the 4 printed digits.
* regress/kovas5.ok: z-varying Dirichlet velocity on every side and
  computed pressure conditions on every side, 200 steps, JACPCG start-up:
  the same 4 printed digits for u, v, w.
* A walled, forced 3D channel (no-slip walls, computed pressure BCs,
  CONST_X, full skew-symmetric advection, random perturbation), N_TIME 1,
  2 and 3: every Fourier plane of u, v, w (and p, Nyquist excluded) after
  each of 8 steps agrees with the oracle to 1e-11 relative L2; the same
  with z-varying Dirichlet data on the walls.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "regress_taylor.json")))


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def phys_of(F):
    from oracle.dns_oracle import inv
    return inv(F)


@pytest.mark.parametrize("startup", ["direct", "pcg"])
@pytest.mark.parametrize("case", ["taylor2", "taylor4"])
def test_gpu_dns_reproduces_regress_goldens(torch_cuda, oracle, case, startup):
    """Both start-ups; the JACPCG one at the sessions' TOL_REL = 1e-12."""
    import torch
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.solver import DNS, SX_ALTSKEW, SX_JACPCG, GpuMesh
    print("case:", case, startup, flush=True)
    nz = 1 if case == "taylor2" else 8
    m = rect_mesh(2, 2, 11, 0.0, 2.0, 0.0, 2.0, periodic_x=True, periodic_y=True)
    gm = GpuMesh(m)
    Lz, nu, dt = 2.0, 0.01, 0.02
    D = DNS(gm, gm, nz, 2 * np.pi / Lz, nu, dt, advection=SX_ALTSKEW)
    if startup == "pcg":
        D.set_startup(SX_JACPCG, tol_rel=1e-12, step_max=500)
    n, S = m.nel * m.npnp, D.stride
    z = np.arange(nz) * Lz / nz
    X = np.broadcast_to(m.x.ravel(), (nz, n))
    Y = np.broadcast_to(m.y.ravel(), (nz, n))
    Z = np.broadcast_to(z[:, None], (nz, n))
    pi = np.pi

    def exact(t):
        e = np.exp(-2 * pi * pi * nu * t)
        if case == "taylor2":
            return [-np.cos(pi * X) * np.sin(pi * Y) * e, np.sin(pi * X) * np.cos(pi * Y) * e]
        return [-np.cos(pi * X) * np.sin(pi * Z) * e, 0 * X, np.sin(pi * X) * np.cos(pi * Z) * e]

    u0 = np.zeros((D.ncom, nz, S))
    for i, f in enumerate(exact(0.0)):
        u0[i, :, :n] = f
    D.set_state(torch.from_numpy(u0).cuda())
    D.step(20)
    torch.cuda.synchronize()
    E = exact(0.4)
    for i, comp in enumerate("uvw"[:D.ncom]):
        got = phys_of(D.field(i).cpu().numpy()[:, :n])
        err = float(np.max(np.abs(got - E[i])))
        want = GOLD[case][comp]
        if want == "noise-level":
            assert err < 1e-12, (comp, err)
        else:
            assert f"{err:.4e}" == f"{want:.4e}", (comp, err, want)


@pytest.mark.parametrize("startup,nz,nord", [("direct", 8, 2), ("pcg", 8, 2), ("direct", 32, 2),
                                             ("direct", 8, 1), ("direct", 8, 3), ("pcg", 8, 3)])
def test_gpu_dns_channel_step_parity(torch_cuda, oracle, startup, nz, nord):
    import torch
    from oracle.dns_oracle import DnsOracle
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.solver import DNS, SX_JACPCG, SX_SKEW, GpuMesh
    print(f"case: channel 3x4 np=6 nz={nz} skew N_TIME={nord}, 8 steps", startup, flush=True)
    bcu, bcp = ("D", "N", "D", "N"), ("N", "N", "N", "N")
    mu = rect_mesh(3, 4, 6, 0.0, 2 * np.pi, -1.0, 1.0, bc=bcu, periodic_x=True)
    mp = rect_mesh(3, 4, 6, 0.0, 2 * np.pi, -1.0, 1.0, bc=bcp, periodic_x=True)
    beta, nu, dt = 1.0, 1.0 / 50, 5e-3
    cf = (2 * nu, 0.0, 0.0)
    gu, gp = GpuMesh(mu), GpuMesh(mp)
    D = DNS(gu, gp, nz, beta, nu, dt, const_force=cf, advection=SX_SKEW, nord=nord)
    O = DnsOracle(mu, mp, nz, beta, nu, dt, const_force=cf, advection="skew", startup=startup,
                  tol_rel=1e-12, nord=nord)
    if startup == "pcg":
        D.set_startup(SX_JACPCG, tol_rel=1e-12, step_max=500)
    n, S = mu.nel * mu.npnp, D.stride
    rng = np.random.default_rng(20261202)
    y = mu.y.ravel()
    wall = 1.0 - y ** 2
    u0 = np.zeros((3, nz, S))
    for i in range(3):
        pert = 0.3 * rng.uniform(-1, 1, (nz, n)) * wall      # vanishes at the walls
        u0[i, :, :n] = (wall if i == 0 else 0.0) + pert
    # make the perturbation single-valued on shared element-boundary nodes
    from semtex_amd.mesh import edgemaps
    em, _ = edgemaps(mu.np)
    for i in range(3):
        for k in range(nz):
            f = u0[i, k, :n].reshape(mu.nel, -1)
            g = np.zeros(mu.nglobal)
            g[mu.btog] = f[:, em[:mu.next]]
            f[:, em[:mu.next]] = g[mu.btog]
            u0[i, k, :n] = f.ravel()
    O.set_initial(u0[:, :, :n])
    D.set_state(torch.from_numpy(u0).cuda())
    for s in range(8):
        O.step()
        D.step(1)
        torch.cuda.synchronize()
        for i in range(4):
            g = D.field(i).cpu().numpy()[:, :n]
            o = O.u[i] if i < 3 else O.p
            planes = [k for k in range(nz) if k != 1]
            rel = np.linalg.norm(g[planes] - o[planes]) / np.linalg.norm(o[planes])
            assert rel < 1e-11, (s, i, rel)


def _channel_run(torch, monkeypatch, tile, nz=32, steps=3):
    from semtex_amd.mesh import edgemaps, rect_mesh
    from semtex_amd.solver import DNS, SX_SKEW, GpuMesh
    monkeypatch.setenv("SX_FFT_THREADS", tile)
    bcu, bcp = ("D", "N", "D", "N"), ("N", "N", "N", "N")
    mu = rect_mesh(4, 6, 6, 0.0, 2 * np.pi, -1.0, 1.0, bc=bcu, periodic_x=True)
    mp = rect_mesh(4, 6, 6, 0.0, 2 * np.pi, -1.0, 1.0, bc=bcp, periodic_x=True)
    beta, nu, dt = 1.0, 1.0 / 100, 2e-3
    D = DNS(GpuMesh(mu), GpuMesh(mp), nz, beta, nu, dt, const_force=(2 * nu, 0.0, 0.0),
            advection=SX_SKEW)
    n, S = mu.nel * mu.npnp, D.stride
    rng = np.random.default_rng(7)
    wall = 1.0 - mu.y.ravel() ** 2
    em, _ = edgemaps(mu.np)
    u0 = np.zeros((3, nz, S))
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


def test_dns_transform_tiles_agree(torch_cuda, monkeypatch):
    """The DNS step's fused z transforms on 128- and 256-thread tiles (the
    tile only regroups columns; each column's arithmetic is the same)."""
    print("case: channel 4x6 np=6 nz=32 skew, 3 steps, transform tiles 128 vs 256", flush=True)
    a = _channel_run(torch_cuda, monkeypatch, "128")
    b = _channel_run(torch_cuda, monkeypatch, "256")
    planes = [0] + list(range(2, 32))
    for i in range(4):
        rel = np.linalg.norm(a[i][planes] - b[i][planes]) / np.linalg.norm(b[i][planes])
        assert rel < 1e-13, (i, rel)


def test_gpu_dns_reproduces_kovas5_golden(torch_cuda, oracle):
    import torch
    from helpers import kovas5_case
    from semtex_amd.solver import DNS, SX_ALTSKEW, SX_JACPCG, GpuMesh
    print("case: kovas5, 200 steps, JACPCG start-up", flush=True)
    K = kovas5_case()
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "regress_kovas.json")))["kovas5"]
    mu, nz = K["mesh_u"], K["nz"]
    D = DNS(GpuMesh(mu), GpuMesh(K["mesh_p"]), nz, K["beta"], K["kinvis"], K["dt"],
            advection=SX_ALTSKEW)
    D.set_startup(SX_JACPCG, tol_rel=1e-8, step_max=500)
    D.set_velocity_bc(torch.from_numpy(K["lines"]).cuda())
    n, S = mu.nel * mu.npnp, D.stride
    X = np.broadcast_to(mu.x.ravel(), (nz, n))
    Y = np.broadcast_to(mu.y.ravel(), (nz, n))
    Z = np.broadcast_to(K["z"][:, None], (nz, n))
    E = K["ex"](X, Y, Z)
    u0 = np.zeros((3, nz, S))
    for i in range(3):
        u0[i, :, :n] = E[i]
    D.set_state(torch.from_numpy(u0).cuda())
    D.step(K["steps"])
    torch.cuda.synchronize()
    for i, c in enumerate("uvw"):
        got = phys_of(D.field(i).cpu().numpy()[:, :n])
        err = float(np.max(np.abs(got - E[i])))
        assert f"{err:.4e}" == f"{gold[c]:.4e}", (c, err, gold[c])


def test_gpu_dns_channel_with_wall_data_step_parity(torch_cuda, oracle):
    """Walls with z- and x-varying Dirichlet velocity (a tangential wall
    motion u = 0.1 sin(x) cos(z), w = 0.05 cos(x); v = 0) against the oracle
    every step."""
    import torch
    from oracle.dns_oracle import DnsOracle
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.solver import DNS, SX_SKEW, GpuMesh
    print("case: channel 3x4 np=6 nz=8 with wall data, 6 steps", flush=True)
    bcu, bcp = ("D", "N", "D", "N"), ("N", "N", "N", "N")
    mu = rect_mesh(3, 4, 6, 0.0, 2 * np.pi, -1.0, 1.0, bc=bcu, periodic_x=True)
    mp = rect_mesh(3, 4, 6, 0.0, 2 * np.pi, -1.0, 1.0, bc=bcp, periodic_x=True)
    nz, beta, nu, dt = 8, 1.0, 1.0 / 50, 5e-3
    cf = (2 * nu, 0.0, 0.0)
    D = DNS(GpuMesh(mu), GpuMesh(mp), nz, beta, nu, dt, const_force=cf, advection=SX_SKEW)
    O = DnsOracle(mu, mp, nz, beta, nu, dt, const_force=cf, advection="skew")
    z = np.arange(nz) * 2 * np.pi / (beta * nz)
    lines = np.zeros((3, nz, mu.nbound * mu.np))
    for b in range(mu.nbound):
        xb, _ = mu.side_xy(b)
        for k in range(nz):
            sl = slice(b * mu.np, (b + 1) * mu.np)
            lines[0, k, sl] = 0.1 * np.sin(xb) * np.cos(z[k])
            lines[2, k, sl] = 0.05 * np.cos(xb)
    O.set_velocity_bc(lines)
    D.set_velocity_bc(torch.from_numpy(lines).cuda())
    n, S = mu.nel * mu.npnp, D.stride
    y = mu.y.ravel()
    wall = 1.0 - y ** 2
    rng = np.random.default_rng(31)
    from semtex_amd.mesh import edgemaps
    em, _ = edgemaps(mu.np)
    u0 = np.zeros((3, nz, S))
    for i in range(3):
        u0[i, :, :n] = (wall if i == 0 else 0.0) + 0.2 * rng.uniform(-1, 1, (nz, n)) * wall
        for k in range(nz):
            f = u0[i, k, :n].reshape(mu.nel, -1)
            g = np.zeros(mu.nglobal)
            g[mu.btog] = f[:, em[:mu.next]]
            f[:, em[:mu.next]] = g[mu.btog]
    O.set_initial(u0[:, :, :n])
    D.set_state(torch.from_numpy(u0).cuda())
    planes = [k for k in range(nz) if k != 1]
    for s in range(6):
        O.step()
        D.step(1)
        torch.cuda.synchronize()
        for i in range(4):
            g = D.field(i).cpu().numpy()[:, :n]
            o = O.u[i] if i < 3 else O.p
            rel = np.linalg.norm(g[planes] - o[planes]) / np.linalg.norm(o[planes])
            assert rel < 1e-11, (s, i, rel)


@pytest.mark.parametrize("nz", [16, 32])
def test_gpu_dns_dealiased_step_parity(torch_cuda, oracle, nz):
    """3/2-rule dealiasing (sx_dns_set_dealias): the walled channel against
    the oracle's dealiased step (DnsOracle.dealias), every field after each
    of 5 steps, at the DNS bar."""
    import torch
    from oracle.dns_oracle import DnsOracle
    from semtex_amd.mesh import edgemaps, rect_mesh
    from semtex_amd.solver import DNS, SX_SKEW, GpuMesh
    print(f"case: channel 3x4 np=6 nz={nz} skew dealiased, 5 steps", flush=True)
    bcu, bcp = ("D", "N", "D", "N"), ("N", "N", "N", "N")
    mu = rect_mesh(3, 4, 6, 0.0, 2 * np.pi, -1.0, 1.0, bc=bcu, periodic_x=True)
    mp = rect_mesh(3, 4, 6, 0.0, 2 * np.pi, -1.0, 1.0, bc=bcp, periodic_x=True)
    beta, nu, dt = 1.0, 1.0 / 50, 5e-3
    cf = (2 * nu, 0.0, 0.0)
    D = DNS(GpuMesh(mu), GpuMesh(mp), nz, beta, nu, dt, const_force=cf, advection=SX_SKEW)
    D.set_dealias(True)
    O = DnsOracle(mu, mp, nz, beta, nu, dt, const_force=cf, advection="skew")
    O.dealias = True
    n, S = mu.nel * mu.npnp, D.stride
    rng = np.random.default_rng(77)
    wall = 1.0 - mu.y.ravel() ** 2
    em, _ = edgemaps(mu.np)
    u0 = np.zeros((3, nz, S))
    for i in range(3):
        u0[i, :, :n] = (wall if i == 0 else 0.0) + 0.3 * rng.uniform(-1, 1, (nz, n)) * wall
        for k in range(nz):
            f = u0[i, k, :n].reshape(mu.nel, -1)
            g = np.zeros(mu.nglobal)
            g[mu.btog] = f[:, em[:mu.next]]
            f[:, em[:mu.next]] = g[mu.btog]
    O.set_initial(u0[:, :, :n])
    D.set_state(torch.from_numpy(u0).cuda())
    planes = [k for k in range(nz) if k != 1]
    for s in range(5):
        O.step()
        D.step(1)
        torch.cuda.synchronize()
        for i in range(4):
            g = D.field(i).cpu().numpy()[:, :n]
            o = O.u[i] if i < 3 else O.p
            rel = np.linalg.norm(g[planes] - o[planes]) / np.linalg.norm(o[planes])
            assert rel < 1e-11, (s, i, rel)


def test_gpu_dns_dump_and_restart_through_field_file(torch_cuda, tmp_path):
    """Domain::dump / Domain::restart through a field file: the state after
    3 steps, written in physical space (sx_field_write) and read back into
    a fresh DNS (sx_field_read + sx_dns_set_state), equals the original
    velocity to round-off; the header carries the step, time and fields."""
    import ctypes as C

    import torch
    from semtex_amd import _lib
    from semtex_amd.problems import dns_channel
    from semtex_amd.solver import DNS, SX_SKEW, GpuMesh, _FieldHeader
    print("case: channel 4x6 np=6 nz=16, dump after 3 steps and restart", flush=True)
    P = dns_channel(nx=4, ny=6, np_=6, nz=16)
    gu, gp = GpuMesh(P["mesh_u"]), GpuMesh(P["mesh_p"])
    kw = dict(const_force=P["const_force"], advection=SX_SKEW)
    A = DNS(gu, gp, P["nz"], P["beta"], P["kinvis"], P["dt"], **kw)
    n = P["mesh_u"].nel * P["mesh_u"].npnp
    u0 = torch.zeros((3, P["nz"], A.stride), dtype=torch.float64)
    u0[:, :, :n] = torch.from_numpy(P["u0"])
    A.set_state(u0.cuda())
    A.step(3)
    path = str(tmp_path / "chan.fld")
    A.dump(path, session="chan")
    h = _FieldHeader()
    _lib.check(_lib.load().sx_field_read_header(path.encode(), 0, C.byref(h)), "header")
    assert h.step == 3 and h.fields == b"uvwp" and h.nz == P["nz"] and h.nel == P["mesh_u"].nel
    assert abs(h.time - 3 * P["dt"]) < 1e-15
    B = DNS(gu, gp, P["nz"], P["beta"], P["kinvis"], P["dt"], **kw)
    B.restart(path)
    planes = [k for k in range(P["nz"]) if k != 1]
    for i in range(3):
        a = A.field(i).cpu().numpy()[planes, :n]
        b = B.field(i).cpu().numpy()[planes, :n]
        assert np.linalg.norm(a - b) < 1e-14 * np.linalg.norm(a), i
