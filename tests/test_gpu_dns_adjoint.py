"""Gradient of J = 1/2 ||u(T) - u_target||_M^2 through 20 DNS steps with
respect to the initial perturbation (the channel shrunk to 6x12 elements,
N_P=6, N_Z=32, Re=2000, dt=1e-3), with the taped step of
semtex_amd.adjoint checkpointed every 5 steps:

* the transposes of the z transform, of the solve (forcing and natural
  boundary data) and of the element operators (gradient, z derivative,
  smoothing kernels) pass the dot-product identity <A x, y> = <x, A^T y>,
  and the operator kernels equal the operators written as torch arithmetic;
* the taped step reproduces the library's DNS step (sx_dns_*) to 1e-11 per
  step;
* the directional derivative along a seeded direction matches central
  finite differences (h = 1e-6) of the library's DNS and of the CPU DNS
  oracle to 1e-6 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 1e-6
STEPS = 20


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def setup(torch_cuda):
    from semtex_amd.adjoint import DnsTape
    from semtex_amd.problems import dns_channel
    P = dns_channel(nx=6, ny=12, np_=6, nz=32)
    T = DnsTape(P["mesh_u"], P["mesh_p"], P["nz"], P["beta"], P["kinvis"], P["dt"],
                const_force=P["const_force"])
    return P, T


def _single_valued(m, a):
    from semtex_amd.mesh import edgemaps
    em, _ = edgemaps(m.np)
    ext = em[:m.next]
    for idx in np.ndindex(a.shape[:-1]):
        f = a[idx].reshape(m.nel, -1)
        g = np.zeros(m.nglobal)
        g[m.btog] = f[:, ext]
        f[:, ext] = g[m.btog]
    return a


def test_element_operator_kernels_and_transposes(torch_cuda, setup):
    """sx_grad_planes (x, y), sx_dz_planes and sx_smooth_planes equal the
    same operators written as torch arithmetic (Element::grad, the z
    derivative, Field::smooth) to round-off, and their transpose kernels
    pass the dot-product identity <A x, y> = <x, A^T y>."""
    torch = torch_cuda
    _, T = setup
    print("case: element operator kernels 6x12 np=6 nz=32", flush=True)
    rng = np.random.default_rng(3)
    shape = (T.nz, T.n)
    ops = [("grad_x", lambda f: T.grad_xy(f, 0), lambda f: T.grad_xy_torch(f, 0)),
           ("grad_y", lambda f: T.grad_xy(f, 1), lambda f: T.grad_xy_torch(f, 1)),
           ("grad_z", T.grad_z, T.grad_z_torch),
           ("smooth", T.smooth, T.smooth_torch)]
    for name, kern, ref in ops:
        x = torch.from_numpy(rng.standard_normal(shape)).cuda().requires_grad_(True)
        y = torch.from_numpy(rng.standard_normal(shape)).cuda()
        ax = kern(x)
        r = ref(x.detach())
        err = float((ax.detach() - r).abs().max() / r.abs().max())
        assert err < 1e-13, (name, err)
        lhs = float((ax * y).sum())
        ax.backward(y)
        rhs = float((x.detach() * x.grad).sum())
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0), (name, lhs, rhs)


def test_transform_transposes(torch_cuda, setup):
    torch = torch_cuda
    _, T = setup
    print("case: transform transposes nz=32", flush=True)
    rng = np.random.default_rng(1)
    for sign in (1, -1):
        x = torch.from_numpy(rng.standard_normal((32, 101))).cuda().requires_grad_(True)
        y = torch.from_numpy(rng.standard_normal((32, 101))).cuda()
        Ax = T.Transform.apply(x, sign)
        (Ax * y).sum().backward()
        lhs, rhs = float((Ax * y).sum().detach()), float((x * x.grad).sum().detach())
        assert abs(lhs - rhs) <= 1e-13 * max(abs(lhs), 1.0), (sign, lhs, rhs)


def test_solve_transpose_forcing_and_boundary_data(torch_cuda, setup):
    torch = torch_cuda
    P, T = setup
    print("case: pressure solve transpose", flush=True)
    rng = np.random.default_rng(2)
    nz, n = P["nz"], T.n
    f = torch.from_numpy(rng.standard_normal((nz, n))).cuda().requires_grad_(True)
    lines = torch.from_numpy(rng.standard_normal((nz, T.nline))).cuda().requires_grad_(True)
    keep = torch.from_numpy(rng.standard_normal((nz, n))).cuda().requires_grad_(True)
    y = torch.from_numpy(rng.standard_normal((nz, n))).cuda()
    u = T.Solve.apply(f, lines, keep, T.Mp, T.gp, T.pgeo)
    (u * y).sum().backward()
    lhs = float((u * y).sum().detach())
    rhs = float(((f * f.grad).sum() + (lines * lines.grad).sum() + (keep * keep.grad).sum()).detach())
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (lhs, rhs)
    # the unsolved Nyquist plane passes keep through
    assert torch.equal(u[1], keep[1])


def test_taped_step_matches_library_dns(torch_cuda, setup):
    torch = torch_cuda
    from semtex_amd.solver import DNS, SX_SKEW
    P, T = setup
    print("case: taped step vs sx_dns, 6x12 np=6 nz=32, 4 steps", flush=True)
    D = DNS(T.gu, T.gp, P["nz"], P["beta"], P["kinvis"], P["dt"],
            const_force=P["const_force"], advection=SX_SKEW)
    n, S = T.n, D.stride
    u0 = torch.zeros((3, P["nz"], S), dtype=torch.float64)
    u0[:, :, :n] = torch.from_numpy(P["u0"])
    D.set_state(u0.cuda())
    with torch.no_grad():
        st = T.initial(torch.from_numpy(P["u0"]).cuda())
        for s in range(4):
            st = T.step(st, s + 1)
            D.step(1)
            torch.cuda.synchronize()
            planes = [0] + list(range(2, P["nz"]))
            for i in range(4):
                g = (st[0][i] if i < 3 else st[1])[planes]
                ref = D.field(i)[planes, :n]
                rel = float(torch.linalg.norm(g - ref) / torch.linalg.norm(ref))
                assert rel < 1e-11, (s, i, rel)


def _J_np(T, U, Ut):
    """1/2 sum Q4 (u - u_t)^2 over components and planes (Nyquist excluded)."""
    d = U - Ut
    d[:, 1] = 0.0
    q4 = T.q4.cpu().numpy()
    return 0.5 * float((q4 * d.reshape(3, d.shape[1], T.nel, -1) ** 2).sum())


def test_dns_gradient_matches_finite_differences(torch_cuda, setup):
    torch = torch_cuda
    from oracle.dns_oracle import DnsOracle
    from semtex_amd.solver import DNS, SX_SKEW
    P, T = setup
    nz, n = P["nz"], T.n
    print(f"case: DNS adjoint, 6x12 np=6 nz=32, {STEPS} steps, checkpoint 5", flush=True)
    rng = np.random.default_rng(20261203)
    Ut = 0.1 * rng.standard_normal((3, nz, n))
    Ut[:, 1] = 0.0
    delta = _single_valued(P["mesh_u"], rng.uniform(-1.0, 1.0, (3, nz, n)))
    # adjoint through the taped, checkpointed steps
    x = torch.from_numpy(P["u0"]).cuda().requires_grad_(True)
    st = T.run(x, STEPS, checkpoint=5)
    J = 0.5 * T.mass_norm2(st[0] - torch.from_numpy(Ut).cuda())
    J.backward()
    adj = float((x.grad.cpu().numpy() * delta).sum())
    # finite differences of the library's DNS
    D = DNS(T.gu, T.gp, nz, P["beta"], P["kinvis"], P["dt"], const_force=P["const_force"],
            advection=SX_SKEW)

    def J_gpu(u_phys):
        u0 = torch.zeros((3, nz, D.stride), dtype=torch.float64)
        u0[:, :, :n] = torch.from_numpy(u_phys)
        D.set_state(u0.cuda())
        D.step(STEPS)
        U = np.stack([D.field(i)[:, :n].cpu().numpy() for i in range(3)])
        return _J_np(T, U, Ut)

    j0 = J_gpu(P["u0"])
    assert abs(j0 - float(J.detach())) <= 1e-10 * abs(j0), (j0, float(J.detach()))
    fd = (J_gpu(P["u0"] + H * delta) - J_gpu(P["u0"] - H * delta)) / (2 * H)
    print("adjoint", adj, "GPU fd", fd, flush=True)
    assert abs(adj - fd) <= 1e-6 * abs(fd), (adj, fd)

    # finite differences of the CPU oracle's DNS
    def J_cpu(u_phys):
        O = DnsOracle(P["mesh_u"], P["mesh_p"], nz, P["beta"], P["kinvis"], P["dt"],
                      const_force=P["const_force"], advection="skew")
        O.set_initial(u_phys)
        for _ in range(STEPS):
            O.step()
        return _J_np(T, np.stack(O.u), Ut)

    fc = (J_cpu(P["u0"] + H * delta) - J_cpu(P["u0"] - H * delta)) / (2 * H)
    print("adjoint", adj, "CPU fd", fc, flush=True)
    assert abs(adj - fc) <= 1e-6 * abs(fc), (adj, fc)
