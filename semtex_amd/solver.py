"""Host-side mirror of semtex's Helmholtz solve interface over the C ABI.

    M = ModalMatrixSys(lambda2, beta, base_mode, num_modes, mesh)   # src/matrix.h:61-66
    u = Field(mesh, nz);  u.solve(f, M)                              # src/field.h:27

Field data are torch tensors on the GPU laid out as the reference's
AuxField: [nz][plane_size] doubles, element-major, row-major np x np per
element (src/auxfield.h:109-114).  torch supplies device memory and the
stream only; all arithmetic runs in the HIP library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .mesh import Mesh2D

SX_FP64, SX_MIXED, SX_JACPCG = _lib.SX_FP64, _lib.SX_MIXED, _lib.SX_JACPCG


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else None


class GpuMesh:
    """Element geometry + NumberSys + BoundarySys resident on the GPU."""

    def __init__(self, mesh: Mesh2D):
        L = _lib.load()
        self.mesh = mesh
        keep = [np.ascontiguousarray(mesh.x, np.float64),
                np.ascontiguousarray(mesh.y, np.float64),
                np.ascontiguousarray(mesh.btog, np.int32),
                np.ascontiguousarray(mesh.bmask, np.int32),
                np.ascontiguousarray(mesh.emask, np.int32),
                np.ascontiguousarray(mesh.bnd_elmt, np.int32),
                np.ascontiguousarray(mesh.bnd_side, np.int32),
                np.ascontiguousarray(mesh.bnd_kind, np.int32)]
        x, y, bt, bm, em, be, bs, bk = keep
        h = C.c_void_p()
        _lib.check(L.sx_mesh_create(mesh.np, mesh.nel, mesh.cyl, _ptr(x), _ptr(y),
                                    _ptr(bt), _ptr(bm), _ptr(em), mesh.nglobal,
                                    mesh.nsolve, mesh.nband, mesh.nbound, _ptr(be),
                                    _ptr(bs), _ptr(bk), C.byref(h)), "sx_mesh_create")
        self.handle = h
        self.nclass = L.sx_mesh_nclass(h)

    @property
    def plane_size(self):
        return self.mesh.plane_size

    def __del__(self):
        if getattr(self, "handle", None):
            _lib.load().sx_mesh_destroy(self.handle)
            self.handle = None


class ModalMatrixSys:
    """Per-Fourier-mode DIRECT systems (src/matrix.cpp:33-105), factorised on
    the GPU.  mode k uses betak2 = (beta * (base_mode + k))^2."""

    def __init__(self, lambda2, beta, base_mode, num_modes, gmesh: GpuMesh,
                 precision=SX_FP64, name="c"):
        L = _lib.load()
        self.gmesh = gmesh
        self.lambda2, self.beta = float(lambda2), float(beta)
        self.base_mode, self.num_modes = int(base_mode), int(num_modes)
        self.precision, self.name = precision, name
        h = C.c_void_p()
        _lib.check(L.sx_msys_create_named(gmesh.handle, self.lambda2, self.beta,
                                          self.base_mode, self.num_modes, name.encode(),
                                          precision, _stream(), C.byref(h)),
                   "sx_msys_create_named")
        self.handle = h

    def nsolve(self, mode):
        return _lib.load().sx_msys_nsolve(self.handle, mode)

    def layout(self):
        """{"panels", "D", "twisted", "nband", "unit", "kernel", "nr"} of the
        factor's solve copies and of the last band solve (kernel: 0 none yet,
        1 panel sweeps, 2 unit-diagonal sweeps, 3 split sweeps; nr:
        right-hand sides per workgroup)."""
        info = (C.c_int * 7)()
        _lib.check(_lib.load().sx_msys_layout(self.handle, info, 7), "sx_msys_layout")
        keys = ("panels", "D", "twisted", "nband", "unit", "kernel", "nr")
        return dict(zip(keys, info))

    def status(self):
        """Synchronises the current stream; raises SxError (code SX_ERR_SYNC)
        when a split band sweep's wait gave up in a solve since the last
        call (its solution holds NaN), as the reference's message(ERROR)
        would end the run."""
        _lib.check(_lib.load().sx_msys_status(self.handle, _stream()), "sx_msys_status")

    def __len__(self):
        return self.num_modes

    def __del__(self):
        if getattr(self, "handle", None):
            _lib.load().sx_msys_destroy(self.handle)
            self.handle = None


class Field:
    """A scalar field with boundary line data (src/field.h)."""

    def __init__(self, gmesh: GpuMesh, nz: int, device="cuda"):
        import torch
        m = gmesh.mesh
        self.gmesh, self.nz = gmesh, nz
        self.data = torch.zeros((nz, m.plane_size), dtype=torch.float64, device=device)
        self.lines = torch.zeros((nz, max(m.nbound * m.np, 1)), dtype=torch.float64,
                                 device=device)

    def solve(self, f, M: ModalMatrixSys, mixed_tol=None, max_iter=20, tol_rel=1e-8,
              step_max=500):
        """Field::solve (src/field.cpp:598-805): f ([nz][plane_size] CUDA
        float64 tensor) is destroyed; the solution overwrites self.data.
        JACPCG systems start from self.data and stop at tol_rel (TOL_REL) or
        step_max (STEP_MAX)."""
        import torch
        assert f.dtype == torch.float64 and f.is_cuda and f.is_contiguous()
        assert tuple(f.shape) == tuple(self.data.shape)
        L = _lib.load()
        lines = C.c_void_p(self.lines.data_ptr()) if self.gmesh.mesh.nbound else None
        if M.precision == SX_JACPCG:
            it = C.c_int(0)
            _lib.check(L.sx_field_solve_pcg(M.handle, self.nz, f.shape[1],
                                            C.c_void_p(f.data_ptr()),
                                            C.c_void_p(self.data.data_ptr()), lines,
                                            float(tol_rel), int(step_max), C.byref(it),
                                            _stream()), "sx_field_solve_pcg")
            self.last_iters = it.value
        elif M.precision == SX_MIXED:
            it = C.c_int(0)
            _lib.check(L.sx_field_solve_mixed(M.handle, self.nz, f.shape[1],
                                              C.c_void_p(f.data_ptr()),
                                              C.c_void_p(self.data.data_ptr()), lines,
                                              float(mixed_tol or 0.0), int(max_iter),
                                              C.byref(it), _stream()),
                       "sx_field_solve_mixed")
            self.last_iters = it.value
        else:
            _lib.check(L.sx_field_solve(M.handle, self.nz, f.shape[1],
                                        C.c_void_p(f.data_ptr()),
                                        C.c_void_p(self.data.data_ptr()), lines,
                                        _stream()), "sx_field_solve")
        return self


def helmholtz_op(M: ModalMatrixSys, u, out):
    """out = H u element by element (Element::HelmholtzOp), per plane."""
    _lib.check(_lib.load().sx_helmholtz_op(M.handle, u.shape[0], u.shape[1],
                                           C.c_void_p(u.data_ptr()),
                                           C.c_void_p(out.data_ptr()), _stream()),
               "sx_helmholtz_op")
    return out


def fourier_transform(x, sign, out=None):
    """AuxField::transform (src/auxfield.cpp:1157-1194): real z transform of
    a [nz][stride] CUDA float64 tensor; sign +1 forward (packed, 1/nz), -1
    inverse.  In place when `out` is None."""
    import torch
    assert x.dtype == torch.float64 and x.is_cuda and x.is_contiguous()
    y = x if out is None else out
    _lib.check(_lib.load().sx_fourier_transform(x.shape[0], x.shape[1], C.c_void_p(x.data_ptr()),
                                                C.c_void_p(y.data_ptr()), int(sign), _stream()),
               "sx_fourier_transform")
    return y


SX_FORWARD, SX_INVERSE = _lib.SX_FORWARD, _lib.SX_INVERSE


def couple(v, w, direction, n=None, base_mode=0):
    """AuxField::couple (src/auxfield.cpp:1376-1451) in place on two
    [nz][stride] CUDA float64 tensors (Field::coupleBCs FORWARD on boundary
    lines, n = stride): FORWARD v~ = v + i w, w~ = v - i w per mode; INVERSE
    back.  Mode 0 of base mode 0 is not coupled."""
    import torch
    assert v.dtype == torch.float64 and v.is_cuda and v.is_contiguous() and v.shape == w.shape
    n = v.shape[1] if n is None else n
    _lib.check(_lib.load().sx_couple(v.shape[0], v.shape[1], n, C.c_void_p(v.data_ptr()),
                                     C.c_void_p(w.data_ptr()), int(direction), int(base_mode),
                                     _stream()), "sx_couple")
    return v, w


class CylModalSys:
    """ModalMatrixSys of one field of a cylindrical 3D problem over its
    BoundarySys (src/matrix.cpp:33-105, src/bsys.cpp:42-123): the axis
    condition changes with the mode (BCmgr::getCondition,
    src/bcmgr.cpp:478-530), so `meshes` holds the numbering of mode 0,
    mode 1 and modes >= 2 (the same mesh object may repeat) and one system
    is built per class with the field's Fourier constant (Field::modeConstant:
    beta k, or beta k +- 1 for the coupled v~, w~).  solve() hands each class
    its planes ([0,2), [2,4), [4,nz)) and boundary lines."""

    def __init__(self, name, lambda2, beta, num_modes, meshes, precision=SX_FP64):
        assert len(meshes) == 3 and num_modes >= 1
        self.name, self.num_modes = name, int(num_modes)
        self.parts = []                   # (first plane, planes, system, mesh)
        for base, count in ((0, 1), (1, 1), (2, num_modes - 2)):
            if count <= 0 or base >= num_modes:
                continue
            gm = meshes[base]
            M = ModalMatrixSys(lambda2, beta, base, count, gm, precision, name=name)
            self.parts.append((2 * base, 2 * count, M, gm))

    def solve(self, f, u_data, lines=None):
        """Field::solve for the whole field: f, u_data [nz][stride] CUDA
        float64 (f destroyed); lines [nz][nbound*np] or None."""
        L = _lib.load()
        nz, S = f.shape
        for k0, npl, M, gm in self.parts:
            if k0 >= nz:
                continue
            npl = min(npl, nz - k0)
            lp = (C.c_void_p(lines[k0].data_ptr())
                  if lines is not None and gm.mesh.nbound else None)
            _lib.check(L.sx_field_solve(M.handle, npl, S, C.c_void_p(f[k0].data_ptr()),
                                        C.c_void_p(u_data[k0].data_ptr()), lp, _stream()),
                       "sx_field_solve")
        return u_data


SX_SKEW, SX_ALTSKEW = _lib.SX_SKEW, _lib.SX_ALTSKEW


class _FieldHeader(C.Structure):
    """sx_field_header (include/semtex_amd.h)."""
    _fields_ = [("session", C.c_char * 64), ("created", C.c_char * 64), ("nr", C.c_int),
                ("ns", C.c_int), ("nz", C.c_int), ("nel", C.c_int), ("step", C.c_int),
                ("time", C.c_double), ("dt", C.c_double), ("kinvis", C.c_double),
                ("beta", C.c_double), ("fields", C.c_char * 32), ("format", C.c_char * 64)]


class DNS:
    """integrate() (dns/integrate.cpp:66-201) on the GPU: the velocity and
    pressure fields and their time history live in HBM; each step runs the
    advection terms, the computed pressure BCs, the pressure Poisson solve,
    the projection and the viscous Helmholtz solves."""

    def __init__(self, gmesh_u: GpuMesh, gmesh_p: GpuMesh, nz, beta, kinvis, dt,
                 const_force=(0.0, 0.0, 0.0), advection=SX_SKEW, nord=2):
        L = _lib.load()
        self.gu, self.gp = gmesh_u, gmesh_p
        self.nz = nz
        self.ncom = 2 if nz == 1 else 3
        cf = (C.c_double * 3)(*const_force)
        h = C.c_void_p()
        _lib.check(L.sx_dns_create(gmesh_u.handle, gmesh_p.handle, nz, float(beta),
                                   float(kinvis), float(dt), C.cast(cf, C.c_void_p),
                                   advection, nord, _stream(), C.byref(h)), "sx_dns_create")
        self.handle = h
        self.stride = L.sx_dns_plane_stride(h)
        self.dt, self.kinvis, self.beta = float(dt), float(kinvis), float(beta)
        self.step_no = 0

    def set_state(self, u_phys):
        """u_phys: [ncom][nz][stride] CUDA float64, physical space."""
        import torch
        assert u_phys.dtype == torch.float64 and u_phys.is_cuda and u_phys.is_contiguous()
        assert tuple(u_phys.shape) == (self.ncom, self.nz, self.stride)
        _lib.check(_lib.load().sx_dns_set_state(self.handle, C.c_void_p(u_phys.data_ptr()),
                                                _stream()), "sx_dns_set_state")

    def step(self, n=1):
        _lib.check(_lib.load().sx_dns_step(self.handle, int(n), _stream()), "sx_dns_step")
        return self

    def set_dealias(self, on=True):
        """3/2-rule dealiasing of the skew-symmetric term's z products."""
        _lib.check(_lib.load().sx_dns_set_dealias(self.handle, int(bool(on)), _stream()),
                   "sx_dns_set_dealias")
        return self

    def set_velocity_bc(self, lines_phys):
        """Steady velocity boundary values, [ncom][nz][nbound*np] CUDA float64
        in physical space (Field::evaluateBoundaries + bTransform); None
        resets them to zero."""
        import torch
        if lines_phys is None:
            ptr = None
        else:
            assert lines_phys.dtype == torch.float64 and lines_phys.is_cuda
            assert lines_phys.is_contiguous()
            self._ulines = lines_phys
            ptr = C.c_void_p(lines_phys.data_ptr())
        _lib.check(_lib.load().sx_dns_set_velocity_bc(self.handle, ptr, _stream()),
                   "sx_dns_set_velocity_bc")
        return self

    def set_startup(self, method, tol_rel=1e-8, step_max=500):
        """Start-up velocity solves: SX_FP64 (direct) or SX_JACPCG."""
        _lib.check(_lib.load().sx_dns_set_startup(self.handle, int(method), float(tol_rel),
                                                  int(step_max), _stream()),
                   "sx_dns_set_startup")
        return self

    def dump(self, path, session="session", append=False):
        """Domain::dump (src/domain.cpp:160-199): the velocity and pressure,
        transformed to physical space, written as one dump of a field file
        (writeField, src/misc.cpp:106-180) with the step and time."""
        import torch
        nz, S = self.nz, self.stride
        mesh = self.gu.mesh
        names = "uv" + ("w" if self.ncom == 3 else "") + "p"
        host = []
        for i in range(self.ncom + 1):
            f = self.field(i)
            if nz > 1:
                f = fourier_transform(f, -1, out=torch.empty_like(f))
            host.append(np.ascontiguousarray(f.cpu().numpy()))
        h = _FieldHeader()
        h.session = session.encode()[:63]
        h.nr = h.ns = mesh.np
        h.nz, h.nel, h.step = nz, mesh.nel, self.step_no
        h.time, h.dt, h.kinvis, h.beta = self.step_no * self.dt, self.dt, self.kinvis, self.beta
        h.fields = names.encode()
        ptrs = (C.c_void_p * len(host))(*[a.ctypes.data for a in host])
        _lib.check(_lib.load().sx_field_write(str(path).encode(), C.byref(h), ptrs, S,
                                              int(bool(append))), "sx_field_write")

    def restart(self, path, dump=0):
        """Domain::restart (src/domain.cpp:123-157): the velocity of dump
        `dump` of a field file (physical space) becomes the state."""
        import torch
        nz, S = self.nz, self.stride
        u = np.zeros((self.ncom, nz, S))
        for i, c in enumerate("uvw"[:self.ncom]):
            _lib.check(_lib.load().sx_field_read(str(path).encode(), int(dump), c.encode(),
                                                 C.c_void_p(u[i].ctypes.data), S),
                       "sx_field_read")
        self.set_state(torch.from_numpy(u).cuda())

    def field(self, which):
        """Fourier-space copy of velocity component `which` (ncom: pressure)."""
        import torch
        out = torch.empty((self.nz, self.stride), dtype=torch.float64, device="cuda")
        st = C.c_int(0)
        _lib.check(_lib.load().sx_dns_get(self.handle, int(which), C.c_void_p(out.data_ptr()),
                                          C.byref(st), _stream()), "sx_dns_get")
        self.step_no = st.value
        return out

    def __del__(self):
        if getattr(self, "handle", None):
            _lib.load().sx_dns_destroy(self.handle)
            self.handle = None


def mass_inner(gmesh: GpuMesh, a, b):
    """Per-plane element-quadrature inner products sum Q4 a b (device)."""
    import torch
    out = torch.empty(a.shape[0], dtype=torch.float64, device=a.device)
    _lib.check(_lib.load().sx_mass_inner(gmesh.handle, a.shape[0], a.shape[1],
                                         C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()),
                                         C.c_void_p(out.data_ptr()), _stream()), "sx_mass_inner")
    return out


def mass_weight(gmesh: GpuMesh, x, scale=1.0):
    import torch
    out = torch.zeros_like(x)
    _lib.check(_lib.load().sx_mass_weight(gmesh.handle, x.shape[0], x.shape[1],
                                          C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()),
                                          float(scale), _stream()), "sx_mass_weight")
    return out


def solve_weak(M: ModalMatrixSys, w):
    """A u = L^T w (zero essential values); w is consumed."""
    import torch
    u = torch.zeros_like(w)
    wc = w.contiguous().clone()
    _lib.check(_lib.load().sx_field_solve_weak(M.handle, w.shape[0], w.shape[1],
                                               C.c_void_p(wc.data_ptr()),
                                               C.c_void_p(u.data_ptr()), _stream()),
               "sx_field_solve_weak")
    return u


def _make_helmholtz_fn():
    import torch

    class HelmholtzSolve(torch.autograd.Function):
        """u = Field::solve(f) for the modes of M, differentiable in f and
        lambda2.  backward: one adjoint solve of the symmetric system,
        dL/df = -Q8 lam, dL/dlambda2 = -sum_k <lam, u>_M."""

        @staticmethod
        def forward(ctx, f, lam2, gmesh, beta, nmodes, lines):
            # the last system built on this mesh is reused while lambda2,
            # beta and the mode count repeat (gradient loops over f)
            key = (float(lam2), float(beta), int(nmodes))
            cached = getattr(gmesh, "_adjoint_msys", None)
            if cached is not None and cached[0] == key:
                M = cached[1]
            else:
                M = ModalMatrixSys(key[0], beta, 0, nmodes, gmesh)
                gmesh._adjoint_msys = (key, M)
            u = Field(gmesh, f.shape[0])
            if lines is not None:
                u.lines.copy_(lines)
            u.solve(f.detach().contiguous().clone(), M)
            ctx.M, ctx.gmesh = M, gmesh
            ctx.save_for_backward(u.data)
            return u.data.clone()

        @staticmethod
        def backward(ctx, g):
            (u,) = ctx.saved_tensors
            lam = solve_weak(ctx.M, g)
            grad_f = mass_weight(ctx.gmesh, lam, -1.0)
            dlam2 = -mass_inner(ctx.gmesh, lam, u).sum()
            return grad_f, dlam2.reshape(()), None, None, None, None

    return HelmholtzSolve


_HELM_FN = None


def helmholtz_solve(f, lam2, gmesh: GpuMesh, beta, nmodes, lines=None):
    """Differentiable batched Helmholtz solve (torch.autograd): f is the
    [nz][plane] forcing (not destroyed), lam2 a 0-d float64 tensor."""
    global _HELM_FN
    if _HELM_FN is None:
        _HELM_FN = _make_helmholtz_fn()
    return _HELM_FN.apply(f, lam2, gmesh, beta, nmodes, lines)


class Batch:
    """Independent problems solved by one captured HIP graph (sx_batch_*).

    problems: list of (ModalMatrixSys, Field, forcing source tensor).  The
    forcing sources and the fields' line data are read in place at every
    run; solutions land in each Field's data."""

    def __init__(self, problems, refine_passes=3):
        import torch
        L = _lib.load()
        self.problems = problems
        n = len(problems)
        self.work = [torch.empty_like(src) for (_, _, src) in problems]
        arr = lambda vals, t: (t * n)(*vals)  # noqa: E731
        sys_ = arr([M.handle.value for (M, _, _) in problems], C.c_void_p)
        nz = arr([u.nz for (_, u, _) in problems], C.c_int)
        stride = arr([src.shape[1] for (_, _, src) in problems], C.c_int)
        fsrc = arr([src.data_ptr() for (_, _, src) in problems], C.c_void_p)
        fwork = arr([w.data_ptr() for w in self.work], C.c_void_p)
        uu = arr([u.data.data_ptr() for (_, u, _) in problems], C.c_void_p)
        lines = arr([u.lines.data_ptr() if u.gmesh.mesh.nbound else None
                     for (_, u, _) in problems], C.c_void_p)
        torch.cuda.synchronize()
        h = C.c_void_p()
        _lib.check(L.sx_batch_create(n, sys_, nz, stride, fsrc, fwork, uu, lines,
                                     int(refine_passes), C.byref(h)), "sx_batch_create")
        self.handle = h

    def run(self):
        _lib.check(_lib.load().sx_batch_run(self.handle, _stream()), "sx_batch_run")

    def tune(self, reps=10):
        """sx_batch_tune: time the sweep choices per band-system class and
        keep the fastest; returns info()."""
        _lib.check(_lib.load().sx_batch_tune(self.handle, int(reps), _stream()), "sx_batch_tune")
        return self.info()

    def info(self):
        buf = C.create_string_buffer(4096)
        _lib.check(_lib.load().sx_batch_info(self.handle, buf, 4096), "sx_batch_info")
        return buf.value.decode()

    def __del__(self):
        if getattr(self, "handle", None):
            _lib.load().sx_batch_destroy(self.handle)
            self.handle = None
