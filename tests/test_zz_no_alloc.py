"""sx_field_solve_mixed_noalloc (the mixed solve on the library's fixed
device array, no allocation) against sx_field_solve_mixed (the same solve
on the workspace the system allocates), both called through the C ABI:
two calls in a row, the second on a smaller input, each result identical
to the allocating entry point's; a workspace above the fixed size is
rejected with SX_ERR_LIMIT before any work is queued.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _solve(torch, fn, M, F, u, lines):
    from semtex_amd import _lib
    from semtex_amd.solver import _stream
    f = F.clone()
    it = C.c_int(-1)
    rc = fn(M.handle, u.nz, f.shape[1], C.c_void_p(f.data_ptr()), C.c_void_p(u.data.data_ptr()),
            lines, C.c_double(1e-6), 20, C.byref(it), _stream())
    _lib.check(rc, fn.__name__)
    torch.cuda.synchronize()
    return u.data.clone(), it.value


def test_noalloc_mixed_solve_matches_allocating_entry_point(torch_cuda):
    import torch
    from semtex_amd import _lib
    from semtex_amd.problems import helmholtz_fourier
    from semtex_amd.solver import SX_MIXED, Field, GpuMesh, ModalMatrixSys
    L = _lib.load()
    cases = [("16x16 np=10 nz=64", dict(nx=16, ny=16, np_=10, nz=64)),
             ("6x4 np=7 nz=8", dict(nx=6, ny=4, np_=7, nz=8))]
    for name, kw in cases:
        print("case:", name, flush=True)
        P = helmholtz_fourier(**kw)
        m = P["mesh"]
        gm = GpuMesh(m)
        M = ModalMatrixSys(P["lambda2"], P["beta"], 0, P["nmodes"], gm, precision=SX_MIXED)
        F = torch.from_numpy(P["forcing"]).cuda()
        u = Field(gm, P["nz"])
        lines = C.c_void_p(u.lines.data_ptr()) if m.nbound else None
        ua, ia = _solve(torch, L.sx_field_solve_mixed, M, F, u, lines)
        u.data.zero_()
        ub, ib = _solve(torch, L.sx_field_solve_mixed_noalloc, M, F, u, lines)
        n = m.nel * m.npnp
        planes = [k for k in range(P["nz"]) if k != 1]
        a = ua[planes, :n].cpu().numpy()
        b = ub[planes, :n].cpu().numpy()
        assert ia == ib and ia >= 1, (ia, ib)
        assert np.all(np.isfinite(b))
        rel = np.linalg.norm(a - b) / np.linalg.norm(a)
        print(f"  refinement passes {ia}, rel L2 noalloc vs allocating {rel:.3e}", flush=True)
        assert rel <= 1e-14, rel


def test_noalloc_rejects_workspace_above_limit(torch_cuda):
    import torch
    from semtex_amd import _lib
    from semtex_amd.problems import helmholtz_fourier
    from semtex_amd.solver import SX_MIXED, GpuMesh, ModalMatrixSys, _stream
    L = _lib.load()
    P = helmholtz_fourier(nx=4, ny=4, np_=6, nz=8)
    gm = GpuMesh(P["mesh"])
    M = ModalMatrixSys(P["lambda2"], P["beta"], 0, P["nmodes"], gm, precision=SX_MIXED)
    # a plane stride whose workspace (5 nz stride + 1024) exceeds the array;
    # the size check runs before any pointer is used
    stride = _lib.SX_NOALLOC_DOUBLES // (5 * 8) + 1
    f = torch.zeros(8, 64, dtype=torch.float64, device="cuda")
    rc = L.sx_field_solve_mixed_noalloc(M.handle, 8, stride, C.c_void_p(f.data_ptr()),
                                        C.c_void_p(f.data_ptr()), None, C.c_double(1e-6), 20,
                                        None, _stream())
    assert rc == _lib.SX_ERR_LIMIT
    assert b"SX_NOALLOC" in L.sx_last_error() or b"above the fixed" in L.sx_last_error()
