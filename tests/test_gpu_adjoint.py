"""Gradients of J = 1/2 ||u - u_target||_M^2 through the batched Helmholtz
solve (helmholtz_solve, torch.autograd) with respect to the forcing field
and lambda2: the adjoint (one solve of the symmetric system) against
central finite differences, h = 1e-6, computed on the GPU and with the CPU
oracle (the same objective from oracle solves).  A seeded Gaussian target
and direction; agreement to 1e-6 relative (finite-difference truncation and
round-off bound it, not the adjoint)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 1e-6


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _weights(m):
    """Q4 (mass weight) of every element node, row-major per element."""
    from oracle.oracle import Q4, mapping
    return np.array([mapping(m.np, m.x[e], m.y[e])[1][Q4] for e in range(m.nel)]).ravel()


def _setup(small):
    from semtex_amd.mesh import rect_mesh
    from semtex_amd.problems import fourier_forcing, helmholtz_fourier
    if small:
        m = rect_mesh(6, 4, 7, 0.0, 2 * np.pi, -1.0, 1.0, bc=("D", "N", "D", "N"), periodic_x=True)
        nz, lam2, beta = 8, 2.0, 1.0
        F = fourier_forcing(m, nz, seed=21)
    else:
        P = helmholtz_fourier()
        m, nz, lam2, beta, F = P["mesh"], P["nz"], P["lambda2"], P["beta"], P["forcing"]
    rng = np.random.default_rng(20261202)
    n = m.nel * m.npnp
    T = np.zeros_like(F)
    D = np.zeros_like(F)
    for k in range(nz):
        if k == 1:
            continue
        T[k, :n] = 0.01 * rng.standard_normal(n)
        D[k, :n] = rng.standard_normal(n)
    return m, nz, lam2, beta, F, T, D


def _J_gpu(torch, gm, F, lam2, beta, nz, T, W):
    from semtex_amd.solver import helmholtz_solve
    n = W.numel()
    u = helmholtz_solve(F, lam2, gm, beta, nz // 2)
    d = (u - T)[:, :n]
    return 0.5 * (W * d * d).sum()


def _J_oracle(oracle, m, F, lam2, beta, T, W):
    n = m.nel * m.npnp
    U = oracle.solve_modes(m, lam2, beta, np.ascontiguousarray(F[:, :n]))
    d = U - T[:, :n]
    planes = [k for k in range(F.shape[0]) if k != 1]
    return 0.5 * sum(float((W * d[k] * d[k]).sum()) for k in planes)


@pytest.mark.parametrize("small", [True, False])
def test_adjoint_gradients_match_finite_differences(torch_cuda, oracle, small):
    import torch
    from semtex_amd.solver import GpuMesh
    print("case: adjoint", "small" if small else "16x16 np=10 nz=64", flush=True)
    m, nz, lam2, beta, F, T, D = _setup(small)
    W = _weights(m)
    gm = GpuMesh(m)
    Wt = torch.from_numpy(W).cuda()
    Tt = torch.from_numpy(T).cuda()
    f = torch.from_numpy(F).cuda().requires_grad_(True)
    l2 = torch.tensor(lam2, dtype=torch.float64, requires_grad=True)
    J = _J_gpu(torch, gm, f, l2, beta, nz, Tt, Wt)
    J.backward()
    Dt = torch.from_numpy(D).cuda()
    adj_f = float((f.grad * Dt).sum())
    adj_l = float(l2.grad)
    with torch.no_grad():
        fp, fm = torch.from_numpy(F + H * D).cuda(), torch.from_numpy(F - H * D).cuda()
        fd_f = float(_J_gpu(torch, gm, fp, torch.tensor(lam2, dtype=torch.float64), beta, nz, Tt, Wt)
                     - _J_gpu(torch, gm, fm, torch.tensor(lam2, dtype=torch.float64), beta, nz, Tt, Wt)) / (2 * H)
        f0 = torch.from_numpy(F).cuda()
        fd_l = float(_J_gpu(torch, gm, f0, torch.tensor(lam2 + H, dtype=torch.float64), beta, nz, Tt, Wt)
                     - _J_gpu(torch, gm, f0, torch.tensor(lam2 - H, dtype=torch.float64), beta, nz, Tt, Wt)) / (2 * H)
    print("adjoint f", adj_f, "fd", fd_f, "| adjoint lam2", adj_l, "fd", fd_l, flush=True)
    assert abs(adj_f - fd_f) <= 1e-6 * abs(fd_f)
    assert abs(adj_l - fd_l) <= 1e-6 * abs(fd_l)
    # the same finite differences of the CPU oracle's objective
    co_f = (_J_oracle(oracle, m, F + H * D, lam2, beta, T, W)
            - _J_oracle(oracle, m, F - H * D, lam2, beta, T, W)) / (2 * H)
    co_l = (_J_oracle(oracle, m, F, lam2 + H, beta, T, W)
            - _J_oracle(oracle, m, F, lam2 - H, beta, T, W)) / (2 * H)
    assert abs(adj_f - co_f) <= 1e-6 * abs(co_f)
    assert abs(adj_l - co_l) <= 1e-6 * abs(co_l)
