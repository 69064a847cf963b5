"""Gradients through the DNS time step (integrate(), dns/integrate.cpp:138-200).

The step is restated on the GPU as a composition whose expensive linear
parts are the library's HIP kernels, each a torch.autograd.Function with a
hand-written adjoint:

* ``solve``: Field::solve (DIRECT, src/field.cpp:598-709), u = A^-1 (L^T (-Q8 f)
  + N^T (area g)) with f the forcing and g the natural boundary data.  A is
  symmetric, so the adjoint is one weak-form solve lam = L A^-1 L^T w
  (sx_field_solve_weak): df = -Q8 lam, dg = area * lam on the boundary
  nodes (Natural::sum, src/condition.cpp:192).  The unsolved Nyquist plane
  passes its old value through.
* ``transform``: AuxField::transform (src/auxfield.cpp:1157-1194, packing
  of femlib/fourier.c:40-165) by the library's z FFT.  With the packing
  [Re0, ReNyq, Re1, Im1, ...] and the forward 1/N scaling, the transposes are
  again transforms: F^T g = INV(W g) / N and INV^T h = N W' FWD(h), with
  W = diag(1, 1, 1/2, 1/2, ...) and W' = diag(1, 1, 2, 2, ...).

* ``grad_xy``, ``grad_z``, ``smooth``: Element::grad (src/element.cpp:521-565),
  the z derivative of AuxField::gradient (src/auxfield.cpp:640-668) and
  Field::smooth (src/field.cpp:282-327) by the library's plane kernels
  (sx_grad_planes, sx_dz_planes, sx_smooth_planes), each backward the
  transpose kernel of the same entry point.

The products of the nonlinear term and the linear combinations of the time
levels and of the computed pressure boundary condition (src/bcmgr.cpp:1037-1335,
src/edge.cpp:98-220; boundary values gathered from the planes) are pointwise
torch arithmetic between these operators, whose adjoint torch forms.
Recomputation is checkpointed every ``checkpoint`` steps
(torch.utils.checkpoint): the backward pass of a segment re-runs its
forward from the stored state.

Only the skew-symmetric Cartesian step of a 3-component field (nz >= 2,
N_TIME = 2, direct start-up) is covered -- the channel configuration.
"""
from __future__ import annotations

import numpy as np

from . import solver as S
from .mesh import edgemaps, gll, side_node_index

SS = {1: [1.0, -1.0], 2: [1.5, -2.0, 0.5]}
EX = {1: [1.0], 2: [2.0, -1.0]}


def _geometry(mesh):
    """Per-node rx, ry, sx, sy, Q4 (= w_i w_j |J|) of every element
    (Element::mapping, src/element.cpp:1253-1408) and, per boundary segment,
    outward unit normals and quadrature weights |dx/dt| w (Element::sideGeom,
    src/element.cpp:1625-1742)."""
    np_ = mesh.np
    z, w, D = gll(np_)
    X = mesh.x.reshape(-1, np_, np_)
    Y = mesh.y.reshape(-1, np_, np_)
    xr, yr = X @ D.T, Y @ D.T                 # d/dr along a row (index j)
    xs, ys = np.einsum("ik,ekj->eij", D, X), np.einsum("ik,ekj->eij", D, Y)
    J = xr * ys - xs * yr
    if np.any(J <= 0):
        raise ValueError("non-positive Jacobian")
    geo = {"rx": ys / J, "ry": -xs / J, "sx": -yr / J, "sy": xr / J,
           "q4": (np.outer(w, w)[None] * J)}
    nb = mesh.nbound
    idx, nx, ny, area = [], [], [], []
    for b in range(nb):
        e, s = int(mesh.bnd_elmt[b]), int(mesh.bnd_side[b])
        k = side_node_index(np_, s)
        i, j = k // np_, k % np_
        if s in (0, 2):
            tx, ty, wt = xr[e, i, j], yr[e, i, j], w[j]
        else:
            tx, ty, wt = xs[e, i, j], ys[e, i, j], w[i]
        t = np.sqrt(tx * tx + ty * ty)
        sg = 1.0 if s in (0, 1) else -1.0     # sides 2, 3 run against r, s
        idx.append(e * np_ * np_ + k)
        nx.append(sg * ty / t)
        ny.append(-sg * tx / t)
        area.append(t * wt)
    cat = (lambda a: np.concatenate(a)) if nb else (lambda a: np.zeros(0))
    geo.update(bidx=cat(idx).astype(np.int64), bnx=cat(nx), bny=cat(ny), barea=cat(area))
    return geo


def _make_functions():
    import torch

    class Transform(torch.autograd.Function):
        """AuxField::transform of [nz, n] planes; sign +1 forward, -1 inverse."""

        @staticmethod
        def forward(ctx, x, sign):
            ctx.sign = sign
            out = torch.empty_like(x)
            return S.fourier_transform(x.detach().contiguous(), sign, out=out)

        @staticmethod
        def backward(ctx, g):
            g = g.contiguous()
            nz = g.shape[0]
            wt = torch.full((nz, 1), 0.5 if ctx.sign > 0 else 2.0, dtype=g.dtype,
                            device=g.device)
            wt[:2] = 1.0
            out = torch.empty_like(g)
            if ctx.sign > 0:                  # F^T g = INV(W g) / N
                S.fourier_transform((g * wt).contiguous(), -1, out=out)
                return out / nz, None
            S.fourier_transform(g, 1, out=out)  # INV^T h = N W' FWD(h)
            return out * wt * nz, None

    class Solve(torch.autograd.Function):
        """Field::solve of all planes of a [nz, n] forcing with natural
        boundary data `lines` [nz, nline]; plane 1 (Nyquist) is not solved
        and keeps keep[1]."""

        @staticmethod
        def forward(ctx, f, lines, keep, M, gm, geo):
            nz, n = f.shape
            ps = gm.mesh.plane_size
            fld = S.Field(gm, nz)
            fld.data[:, :n] = keep.detach()
            if lines is not None:
                fld.lines.copy_(lines.detach())
            fw = torch.zeros((nz, ps), dtype=f.dtype, device=f.device)
            fw[:, :n] = f.detach()
            fld.solve(fw, M)
            ctx.M, ctx.gm, ctx.geo, ctx.n = M, gm, geo, n
            ctx.has_lines = lines is not None
            return fld.data[:, :n].clone()

        @staticmethod
        def backward(ctx, g):
            nz, n = g.shape
            ps = ctx.gm.mesh.plane_size
            w = torch.zeros((nz, ps), dtype=g.dtype, device=g.device)
            w[:, :n] = g
            w[1] = 0.0
            lam = S.solve_weak(ctx.M, w)
            df = S.mass_weight(ctx.gm, lam, -1.0)[:, :n]
            df[1] = 0.0
            dkeep = torch.zeros_like(g)
            dkeep[1] = g[1]
            dlines = None
            if ctx.has_lines:
                geo = ctx.geo
                dlines = lam[:, geo["bidx_t"]] * geo["barea_t"] * geo["natural_t"]
                dlines[1] = 0.0
            return df, dlines, dkeep, None, None, None

    def _planes(x):
        import torch
        assert x.dtype == torch.float64 and x.is_cuda
        return x.contiguous()

    class GradXY(torch.autograd.Function):
        """Element::grad of [nz, n] planes, direction d (0: x, 1: y), by the
        library's kernel; backward: its transpose kernel."""

        @staticmethod
        def forward(ctx, x, d, gm):
            x = _planes(x)
            ctx.d, ctx.gm = d, gm
            out = torch.empty_like(x)
            _lib_check(_L().sx_grad_planes(gm.handle, x.shape[0], x.shape[1], _p(x), _p(out),
                                           d, 0, S._stream()), "sx_grad_planes")
            return out

        @staticmethod
        def backward(ctx, g):
            g = _planes(g)
            out = torch.empty_like(g)
            _lib_check(_L().sx_grad_planes(ctx.gm.handle, g.shape[0], g.shape[1], _p(g), _p(out),
                                           ctx.d, 1, S._stream()), "sx_grad_planes")
            return out, None, None

    class Smooth(torch.autograd.Function):
        """Field::smooth of [nz, n] planes by the library's kernel;
        backward: its transpose kernel."""

        @staticmethod
        def forward(ctx, x, gm):
            x = _planes(x)
            ctx.gm = gm
            out = torch.empty_like(x)
            _lib_check(_L().sx_smooth_planes(gm.handle, x.shape[0], x.shape[1], _p(x), _p(out),
                                             0, S._stream()), "sx_smooth_planes")
            return out

        @staticmethod
        def backward(ctx, g):
            g = _planes(g)
            out = torch.empty_like(g)
            _lib_check(_L().sx_smooth_planes(ctx.gm.handle, g.shape[0], g.shape[1], _p(g),
                                             _p(out), 1, S._stream()), "sx_smooth_planes")
            return out, None

    class GradZ(torch.autograd.Function):
        """d/dz of [nz, n] Fourier planes (i beta k) by the library's kernel;
        backward: the transpose kernel (-d/dz)."""

        @staticmethod
        def forward(ctx, x, beta):
            x = _planes(x)
            ctx.beta = beta
            out = torch.empty_like(x)
            _lib_check(_L().sx_dz_planes(x.shape[0], x.shape[1], x.shape[1], float(beta), _p(x),
                                         _p(out), 0, S._stream()), "sx_dz_planes")
            return out

        @staticmethod
        def backward(ctx, g):
            g = _planes(g)
            out = torch.empty_like(g)
            _lib_check(_L().sx_dz_planes(g.shape[0], g.shape[1], g.shape[1], float(ctx.beta),
                                         _p(g), _p(out), 1, S._stream()), "sx_dz_planes")
            return out, None

    return Transform, Solve, GradXY, Smooth, GradZ


def _L():
    from . import _lib
    return _lib.load()


def _lib_check(rc, what):
    from . import _lib
    _lib.check(rc, what)


def _p(t):
    import ctypes as C
    return C.c_void_p(t.data_ptr())


_FN = None


def _fns():
    global _FN
    if _FN is None:
        _FN = _make_functions()
    return _FN


class DnsTape:
    """The DNS step of one channel configuration as a differentiable map of
    the state (u, p, time levels, boundary histories).

    mesh_u / mesh_p: velocity (essential walls) and pressure (natural)
    numberings of the same elements, as for semtex_amd.solver.DNS."""

    def __init__(self, mesh_u, mesh_p, nz, beta, kinvis, dt, const_force=(0.0, 0.0, 0.0)):
        import torch
        if nz < 2:
            raise ValueError("DnsTape: 3-component fields (nz >= 2) only")
        self.mu, self.mp = mesh_u, mesh_p
        self.nz, self.beta, self.nu, self.dt = nz, float(beta), float(kinvis), float(dt)
        self.const = [float(c) for c in const_force]
        self.np_, self.nel = mesh_u.np, mesh_u.nel
        self.n = self.nel * self.np_ ** 2
        dev = "cuda"
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)  # noqa: E731
        g = _geometry(mesh_u)
        self.D = t(gll(self.np_)[2])
        self.rx, self.ry, self.sx, self.sy = t(g["rx"]), t(g["ry"]), t(g["sx"]), t(g["sy"])
        self.q4 = t(g["q4"].reshape(self.nel, -1))
        em, _ = edgemaps(self.np_)
        nb = mesh_u.next
        self.em_b = t(em[:nb].astype(np.int64))
        self.btog = t(mesh_u.btog.astype(np.int64))
        mass = np.zeros(mesh_u.nglobal)
        np.add.at(mass, mesh_u.btog, g["q4"].reshape(self.nel, -1)[:, em[:nb]])
        self.imass = t(1.0 / mass)
        self.nglobal = mesh_u.nglobal
        self.bidx, self.bnx, self.bny = t(g["bidx"]), t(g["bnx"]), t(g["bny"])
        gp = _geometry(mesh_p)
        kinds = np.repeat(np.asarray(mesh_p.bnd_kind) != 0, self.np_).astype(float)
        self.pgeo = {"bidx_t": t(gp["bidx"]), "barea_t": t(gp["barea"]), "natural_t": t(kinds)}
        self.nline = mesh_u.nbound * self.np_
        bk = np.zeros(nz)
        bk[2:] = self.beta * (np.arange(2, nz) >> 1)
        self.bk = t(bk).reshape(nz, 1)
        self.gu, self.gp = S.GpuMesh(mesh_u), S.GpuMesh(mesh_p)
        self.Mp = S.ModalMatrixSys(0.0, self.beta, 0, nz // 2, self.gp)
        self.Mu = {je: S.ModalMatrixSys(SS[je][0] / (self.dt * self.nu), self.beta, 0, nz // 2,
                                        self.gu) for je in (1, 2)}
        self.Transform, self.Solve, self.GradXY, self.Smooth, self.GradZ = _fns()

    # ------------------------------------------------------------ operators
    def fwd(self, x):
        return self.Transform.apply(x.contiguous(), 1)

    def inv(self, x):
        return self.Transform.apply(x.contiguous(), -1)

    def grad_xy(self, f, d):
        """Element::grad (src/element.cpp:521-565) of [nz, n] planes in
        direction d: the library's kernel and its transpose."""
        return self.GradXY.apply(f, d, self.gu)

    def grad_z(self, F):
        """Re_k <- -beta k Im_k, Im_k <- beta k Re_k; planes 0, 1 zero."""
        return self.GradZ.apply(F, self.beta)

    def smooth(self, F):
        """Field::smooth (src/field.cpp:282-327): the library's kernel and
        its transpose."""
        return self.Smooth.apply(F, self.gu)

    # the same three operators as torch arithmetic (Element::grad,
    # AuxField::gradient in z, Field::smooth), kept to check the kernels
    def grad_xy_torch(self, f, d):
        u = f.reshape(-1, self.nel, self.np_, self.np_)
        ur = u @ self.D.T
        us = self.D @ u
        out = ur * self.rx + us * self.sx if d == 0 else ur * self.ry + us * self.sy
        return out.reshape(f.shape)

    def grad_z_torch(self, F):
        import torch
        out = torch.zeros_like(F)
        out[2::2] = -self.bk[2::2] * F[3::2]
        out[3::2] = self.bk[3::2] * F[2::2]
        return out

    def smooth_torch(self, F):
        import torch
        nz = F.shape[0]
        Fe = F.reshape(nz, self.nel, -1)
        vals = self.q4[:, self.em_b] * Fe[:, :, self.em_b]
        ds = torch.zeros((nz, self.nglobal), dtype=F.dtype, device=F.device)
        ds = ds.index_add(1, self.btog.reshape(-1), vals.reshape(nz, -1))
        ds = ds * self.imass
        Fe = Fe.clone()
        Fe[:, :, self.em_b] = ds[:, self.btog]
        return Fe.reshape(F.shape)

    def side(self, F):
        return F[:, self.bidx]

    # -------------------------------------------------------------- pieces
    def advection(self, U):
        """skewSymmetric (dns/nonlinear.cpp:38-194, Cartesian branch)."""
        import torch
        phys = [self.inv(U[i]) for i in range(3)]
        N = []
        for i in range(3):
            n_i = torch.zeros_like(phys[0])
            for j in range(3):
                tmp = self.inv(self.grad_z(U[i])) if j == 2 else self.grad_xy(phys[i], j)
                n_i = n_i - phys[j] * tmp
                tmp = phys[i] * phys[j]
                tmp = self.inv(self.grad_z(self.fwd(tmp))) if j == 2 else self.grad_xy(tmp, j)
                n_i = n_i - tmp
            n_i = self.fwd(0.5 * n_i)
            if abs(self.const[i]) > 6e-14:
                n_i = n_i.clone()
                n_i[0] = n_i[0] + self.const[i]
            N.append(self.smooth(n_i))
        return torch.stack(N)

    def hopbc(self, U, N, un, hb, step_no):
        """maintainFourier + evaluateCNBCp (src/bcmgr.cpp:1037-1335): new
        un, hopbc histories and the pressure's boundary data."""
        nu, dt = self.nu, self.dt
        Je = min(step_no - 1, 2)
        un_new = self.side(U[0]) * self.bnx + self.side(U[1]) * self.bny
        mask = np.ones(self.nz)
        mask[1] = 0.0
        import torch
        mk = torch.as_tensor(mask, device=U.device).reshape(-1, 1)
        un_new = un_new * mk
        h = self.side(N[0]) * self.bnx + self.side(N[1]) * self.bny
        if Je:
            a = SS[Je]
            work = a[0] * un_new
            # as src/bcmgr.cpp:1124-1130 forms it: every a[q+1] multiplies the
            # previous step's u.n (un[0]), not level q
            for q in range(Je):
                work = work + a[q + 1] * un[0]
            h = h - work / dt
        # nu n.curl curl u on every plane
        w = self.grad_xy(U[1], 0) - self.grad_xy(U[0], 1)
        Wz = self.grad_z(U[2])
        b2 = self.bk * self.bk
        cx = self.grad_xy(w, 1) + self.grad_xy(Wz, 0) + b2 * U[0]
        cy = -self.grad_xy(w, 0) + self.grad_xy(Wz, 1) + b2 * U[1]
        h = (h - nu * (self.side(cx) * self.bnx + self.side(cy) * self.bny)) * mk
        hb = torch.stack([h, hb[0]])
        un = torch.stack([un_new, un[0]])
        Jl = min(step_no, 2)
        lines = sum(EX[Jl][q] * hb[q] for q in range(Jl)) * mk
        return un, hb, lines

    # ---------------------------------------------------------------- step
    def initial(self, u_phys):
        """Domain::restart of a physical [3, nz, n] field: the state tuple
        (u, p, Us1, Uf0, Uf1, un, hopbc) before step 1."""
        import torch
        u = torch.stack([self.fwd(u_phys[i]) for i in range(3)])
        mask = torch.ones((1, self.nz, 1), dtype=u.dtype, device=u.device)
        mask[0, 1] = 0.0
        u = u * mask
        z = torch.zeros_like(u)
        zl = torch.zeros((2, self.nz, self.nline), dtype=u.dtype, device=u.device)
        return (u, torch.zeros_like(u[0]), z, z, z, zl, zl)

    def step(self, state, step_no):
        """integrate() loop body for step `step_no` (1-based) of state
        (u, p, Us1, Uf0, Uf1, un, hopbc); Us1 = u^{n-1}, Uf0/Uf1 the last
        two advection terms."""
        import torch
        u, p, Us1, Uf0, Uf1, un, hb = state
        dt, nu = self.dt, self.nu
        U = u
        N = self.advection(U)
        un, hb, plines = self.hopbc(U, N, un, hb, step_no)
        Je = min(step_no, 2)
        a, b = SS[Je], [c * dt for c in EX[Je]]
        H = -a[1] * U + b[0] * N
        if Je == 2:
            H = H - a[2] * Us1 + b[1] * Uf0
        fp = (self.grad_xy(H[0], 0) + self.grad_xy(H[1], 1) + self.grad_z(H[2])) / dt
        p = self.Solve.apply(fp, plines, p, self.Mp, self.gp, self.pgeo)
        gp = [self.grad_xy(p, 0), self.grad_xy(p, 1), self.grad_z(p)]
        F = torch.stack([H[i] * (-1.0 / (dt * nu)) + gp[i] * (1.0 / nu) for i in range(3)])
        unew = torch.stack([self.Solve.apply(F[i], None, U[i], self.Mu[Je], self.gu, None)
                            for i in range(3)])
        return (unew, p, U, N, Uf0, un, hb)

    def run(self, u_phys, steps, checkpoint=5):
        """Velocity after `steps` steps from the physical field u_phys, with
        the forward stored only every `checkpoint` steps."""
        from torch.utils.checkpoint import checkpoint as ckpt
        state = self.initial(u_phys)
        s = 0
        while s < steps:
            seg = min(checkpoint, steps - s)

            def segment(*st, s0=s, k=seg):
                for q in range(k):
                    st = self.step(st, s0 + q + 1)
                return st

            state = ckpt(segment, *state, use_reentrant=False)
            s += seg
        return state

    def mass_norm2(self, X):
        """sum over components and planes (Nyquist excluded) of sum Q4 X^2."""
        import torch
        mask = torch.ones((1, self.nz, 1), dtype=X.dtype, device=X.device)
        mask[0, 1] = 0.0
        Xe = (X * mask).reshape(X.shape[0], self.nz, self.nel, -1)
        return (self.q4 * Xe * Xe).sum()
