"""CPU oracle of semtex's DNS time step (Cartesian, 3 velocity components,
skew-symmetric advection, stiffly-stable velocity-correction splitting).

TEST INFRASTRUCTURE ONLY (the checker of the GPU DNS step).  Field algebra
is numpy; every Helmholtz plane solve goes through the C oracle
(oracle/sem_oracle.c or_solve_plane).  The routine structure follows the
reference:

  integrate            dns/integrate.cpp:66-201
  waveProp             dns/integrate.cpp:204-239
  setPForce            dns/integrate.cpp:242-257
  project              dns/integrate.cpp:260-289
  Solve / preSolve     dns/integrate.cpp:292-356
  skewSymmetric        dns/nonlinear.cpp:38-194 (Cartesian branch :162-193)
  altSkewSymmetric     dns/nonlinear.cpp:198-381 (Cartesian branch :342-380)
  AuxField::gradient   src/auxfield.cpp:529-680
  AuxField::transform  src/auxfield.cpp:1157-1194 (femlib/fourier.c packing)
  Field::smooth        src/field.cpp:282-327
  BCmgr::maintainFourier / evaluateCNBCp   src/bcmgr.cpp:1037-1335
  Edge::curlCurl       src/edge.cpp:98-220
  ConstForce::fourier  dns/fieldforce.cpp:268-276
  Integration          src/integration.cpp:99-164

Documented departure: the reference solves the velocity Helmholtz problems
of the first N_TIME-1 steps with a temporary Jacobi-PCG system to TOL_REL
(dns/integrate.cpp:340-353); here (and on the GPU) those steps use the
direct solve of the same first-order system, i.e. the PCG limit
TOL_REL -> 0.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

SS = {1: [1.0, -1.0], 2: [1.5, -2.0, 0.5], 3: [11.0 / 6.0, -3.0, 1.5, -1.0 / 3.0]}
EX = {1: [1.0], 2: [2.0, -1.0], 3: [3.0, -3.0, 1.0]}


def fwd(x):
    """AuxField::transform(FORWARD): real DFT along z of [nz, n] planes,
    scaled 1/nz, packed [Re0, ReNyq, Re1, Im1, ...] (femlib/fourier.c)."""
    nz = x.shape[0]
    if nz < 2:
        return x.copy()
    F = np.fft.rfft(x, axis=0) / nz
    out = np.empty_like(x)
    out[0] = F[0].real
    out[1] = F[nz // 2].real
    out[2::2] = F[1:nz // 2].real
    out[3::2] = F[1:nz // 2].imag
    return out


def inv(X):
    """AuxField::transform(INVERSE)."""
    nz = X.shape[0]
    if nz < 2:
        return X.copy()
    F = np.zeros((nz // 2 + 1,) + X.shape[1:], dtype=complex)
    F[0] = X[0]
    F[nz // 2] = X[1]
    F[1:nz // 2] = X[2::2] + 1j * X[3::2]
    return np.fft.irfft(F, n=nz, axis=0) * nz


def inv32(X, M):
    """AuxField::transform32(INVERSE) (src/auxfield.cpp:1197-1258) with the
    3/2 rule: the N packed Fourier planes, modes 0..N/2-1 (the Nyquist plane
    dropped), zero-padded and inverse transformed onto M = 3N/2 z points."""
    N = X.shape[0]
    F = np.zeros((M // 2 + 1,) + X.shape[1:], dtype=complex)
    F[0] = X[0]
    F[1:N // 2] = X[2::2] + 1j * X[3::2]
    return np.fft.irfft(F, n=M, axis=0) * M


def fwd32(x, N):
    """AuxField::transform32(FORWARD) with the 3/2 rule: the real DFT of M
    physical planes (scaled 1/M), truncated to modes 0..N/2-1 in the N-plane
    packing; the Nyquist plane is zero (the truncation drops mode N/2)."""
    M = x.shape[0]
    F = np.fft.rfft(x, axis=0) / M
    out = np.zeros((N,) + x.shape[1:])
    out[0] = F[0].real
    out[2::2] = F[1:N // 2].real
    out[3::2] = F[1:N // 2].imag
    return out


class DnsOracle:
    """State and step of integrate() for a 2D mesh x N_Z Fourier planes.

    mesh_u: velocity numbering (essential walls), mesh_p: pressure numbering
    (natural / computed BCs on the same walls).  Boundary segments of the
    two meshes are the same list in the same order.
    """

    def __init__(self, mesh_u, mesh_p, nz, beta, kinvis, dt, const_force=(0.0, 0.0, 0.0),
                 nord=2, advection="skew", startup="direct", tol_rel=1e-8, step_max=500):
        self.mu, self.mp = mesh_u, mesh_p
        # 2D (nz == 1): u, v only (NCOM = NDIM = 2); 3D: u, v, w
        self.ncom = 2 if nz == 1 else 3
        self.advection = advection      # "skew" (-S) or "alt" (default ADVECTION=1)
        # velocity solves of the start-up steps (step < NORD): "direct", or
        # "pcg" as the reference does (dns/integrate.cpp:340-353, JACPCG at
        # TOL_REL / STEP_MAX from the current field)
        self.startup, self.tol_rel, self.step_max = startup, tol_rel, step_max
        self.toggle = 1                 # altSkewSymmetric's static toggle
        self.np_ = mesh_u.np
        self.nel = mesh_u.nel
        self.nz, self.beta, self.nu, self.dt = nz, beta, kinvis, dt
        self.nord = nord
        self.const = np.asarray(const_force, float)
        npnp = self.np_ ** 2
        self.n = self.nel * npnp
        z, w, D, _ = O.gll(self.np_)
        self.D = D
        self.geo = np.array([O.mapping(self.np_, mesh_u.x[e], mesh_u.y[e])[1]
                             for e in range(self.nel)])          # [nel, NGEO, npnp]
        g = self.geo
        sh = (self.nel, self.np_, self.np_)
        self.rx, self.ry = g[:, O.DRDX].reshape(sh), g[:, O.DRDY].reshape(sh)
        self.sx, self.sy = g[:, O.DSDX].reshape(sh), g[:, O.DSDY].reshape(sh)
        self.q4 = g[:, O.Q4]
        # inverse global mass on element boundaries for smoothing (BCmgr imass)
        em, _ = O.edgemaps(self.np_)
        self.em = em
        nb = mesh_u.next
        mass = np.zeros(mesh_u.nglobal)
        np.add.at(mass, mesh_u.btog, self.q4[:, em[:nb]])
        self.imass = 1.0 / mass
        # boundary segments: node indices and unit outward normals
        self.nb = mesh_u.nbound
        self.seg_nodes, self.seg_nx, self.seg_ny = [], [], []
        from semtex_amd.mesh import side_node_index
        for b in range(self.nb):
            e, s = int(mesh_u.bnd_elmt[b]), int(mesh_u.bnd_side[b])
            area, nx, ny = np.zeros(self.np_), np.zeros(self.np_), np.zeros(self.np_)
            O.lib().or_side_geom(self.np_, np.ascontiguousarray(self.geo[e].ravel()),
                                 np.ascontiguousarray(mesh_u.x[e]),
                                 np.ascontiguousarray(mesh_u.y[e]), s, area, nx, ny)
            self.seg_nodes.append((e, side_node_index(self.np_, s)))
            self.seg_nx.append(nx)
            self.seg_ny.append(ny)
        self.nline = self.nb * self.np_
        # Helmholtz systems per mode (preSolve, Solve)
        self.om_u = O.OracleMesh(mesh_u)
        self.om_p = O.OracleMesh(mesh_p)
        self._sys = {}
        self.step_no = 0
        nc = self.ncom
        self.u = np.zeros((nc, nz, self.n))
        self.p = np.zeros((nz, self.n))
        # time levels as lists of arrays: rolling rotates the list (rollm)
        self.Us = [np.zeros((nc, nz, self.n)) for _ in range(nord)]
        self.Uf = [np.zeros((nc, nz, self.n)) for _ in range(nord)]
        self.un = np.zeros((nord, nz, self.nline))
        self.hopbc = np.zeros((nord, nz, self.nline))
        # velocity boundary values in Fourier space, [ncom][nz][nline], or
        # None for zero data (the essential segments are then no-slip walls)
        self.ulines = None
        # 3/2-rule dealiasing of the z products (advection "skew" only)
        self.dealias = False

    # ---------------------------------------------------------------- ops
    def grad_xy(self, f, d):
        """AuxField::gradient(0|1) per plane (Element::grad)."""
        sh = f.shape
        u = f.reshape(-1, self.nel, self.np_, self.np_)
        ur = u @ self.D.T
        us = np.einsum("ik,pekj->peij", self.D, u)
        if d == 0:
            out = ur * self.rx + us * self.sx
        else:
            out = ur * self.ry + us * self.sy
        return out.reshape(sh)

    def grad_z(self, F):
        """AuxField::gradient(2) on Fourier data: planes 0, 1 zeroed,
        Re_k <- -beta k Im_k, Im_k <- beta k Re_k."""
        out = np.zeros_like(F)
        for k in range(1, self.nz // 2):
            bk = self.beta * k
            out[2 * k] = -bk * F[2 * k + 1]
            out[2 * k + 1] = bk * F[2 * k]
        return out

    def smooth(self, F):
        """Field::smooth: mass-averaged direct stiffness sum on element
        boundaries, every plane."""
        nb = self.mu.next
        em = self.em[:nb]
        out = F.copy().reshape(self.nz, self.nel, -1)
        for k in range(self.nz):
            ds = np.zeros(self.mu.nglobal)
            np.add.at(ds, self.mu.btog, self.q4[:, em] * out[k][:, em])
            ds *= self.imass
            out[k][:, em] = ds[self.mu.btog]
        return out.reshape(F.shape)

    def system(self, field, lam2, mode):
        key = (field, lam2, mode)
        if key not in self._sys:
            om = self.om_p if field == "p" else self.om_u
            self._sys[key] = O.OracleMsys(om, lam2, (self.beta * mode) ** 2)
        return self._sys[key]

    def solve(self, field, lam2, F, lines, keep):
        """Field::solve over planes; the Nyquist plane keeps its old data."""
        out = keep.copy()
        for k in range(self.nz):
            if self.nz > 1 and k == 1:
                continue
            S = self.system(field, lam2, k >> 1)
            out[k] = S.solve_plane(F[k], None if lines is None else lines[k])
        return out

    # ---------------------------------------------------------- advection
    def skew_symmetric(self):
        """dns/nonlinear.cpp:38-194 (Cartesian) or, for advection "alt",
        altSkewSymmetric (:198-381): convective form u.grad(u) on odd
        calls, conservative div(uu) on even calls, no 1/2 factor.
        Returns (U, N): the Fourier copy of the velocity and N."""
        nc = self.ncom
        U = self.u.copy()
        if self.dealias:
            return U, self.skew_symmetric_32(U)
        phys = np.array([inv(U[i]) for i in range(nc)])
        N = np.zeros_like(U)
        conv = self.advection == "skew" or self.toggle == 1
        cons = self.advection == "skew" or self.toggle == 0
        for i in range(nc):
            n_i = np.zeros((self.nz, self.n))
            for j in range(nc):
                if conv:
                    if j == 2:
                        tmp = inv(self.grad_z(U[i]))
                    else:
                        tmp = self.grad_xy(phys[i], j)
                    n_i -= phys[j] * tmp
                if cons:
                    tmp = phys[i] * phys[j]
                    if j == 2:
                        tmp = inv(self.grad_z(fwd(tmp)))
                    else:
                        tmp = self.grad_xy(tmp, j)
                    n_i -= tmp
            if self.advection == "skew":
                n_i *= 0.5
            n_i = fwd(n_i)
            if abs(self.const[i]) > 6e-14:
                n_i[0] += self.const[i]
            N[i] = self.smooth(n_i)
        if self.advection == "alt":
            self.toggle = 1 - self.toggle
        return U, N

    def skew_symmetric_32(self, U):
        """The skew-symmetric nonlinear term with every z product formed on
        M = 3N/2 physical planes (transform32, src/auxfield.cpp:1197-1258):
        the product of two fields with modes |k| <= N/2-1 has modes |k| <=
        N-2, which on M points alias onto |k| >= M-(N-2) = N/2+2, so the
        retained modes 0..N/2-1 are free of aliasing.  Same term by term as
        skewSymmetric (dns/nonlinear.cpp:162-193): the physical-space parts
        are summed on M planes and truncated once; d_z(u_i w) is formed in
        Fourier space from the truncated product."""
        nz, nc = self.nz, self.ncom
        Mz = 3 * nz // 2
        physM = np.array([inv32(U[i], Mz) for i in range(nc)])
        N = np.zeros_like(U)
        for i in range(nc):
            n_i = np.zeros((Mz, self.n))
            for j in range(nc):
                # convective u_j d_j u_i
                tmp = inv32(self.grad_z(U[i]), Mz) if j == 2 else self.grad_xy(physM[i], j)
                n_i -= physM[j] * tmp
                # conservative d_j (u_i u_j), x and y on the M planes
                if j < 2:
                    n_i -= self.grad_xy(physM[i] * physM[j], j)
            Ni = fwd32(n_i, nz)
            Ni -= self.grad_z(fwd32(physM[i] * physM[2], nz))
            Ni *= 0.5
            if abs(self.const[i]) > 6e-14:
                Ni[0] += self.const[i]
            N[i] = self.smooth(Ni)
        return N

    def set_initial(self, u_phys):
        """Domain::restart: physical IC, forward transform, Nyquist plane
        zeroed (src/domain.cpp:123-157)."""
        for i in range(self.ncom):
            self.u[i] = fwd(np.asarray(u_phys[i], float))
            if self.nz > 1:
                self.u[i][1] = 0.0
        self.p[:] = 0.0

    def physical(self, F):
        return inv(F)

    def set_velocity_bc(self, lines_phys):
        """Steady velocity boundary values given per z plane in physical
        space, [ncom][nz][nline] (segment b's np values in CCW traverse
        order): evaluated, then forward transformed as Field::bTransform
        does (src/field.cpp:174-203: for nz == 2 the first plane keeps its
        physical values and the second is zeroed).  Only the essential
        segments' values are used."""
        L = np.asarray(lines_phys, float)
        out = np.zeros_like(L)
        for i in range(self.ncom):
            if self.nz == 2:
                out[i][0] = L[i][0]
            else:
                out[i] = fwd(L[i])
        self.ulines = out

    # -------------------------------------------------------------- HOPBC
    def side_vals(self, F, b):
        e, idx = self.seg_nodes[b]
        return F.reshape(self.nel, -1)[e, idx]

    def curl_curl(self, Ur, Ui, Vr, Vi, Wr, Wi, mode):
        """Edge::curlCurl (src/edge.cpp:98-220), Cartesian, on every
        boundary segment: (x, y) components of curl curl u, Re and Im."""
        def wz(Uf, Vf):
            return self.grad_xy(Vf[None], 0)[0] - self.grad_xy(Uf[None], 1)[0]
        wr = wz(Ur, Vr)
        dwx_r, dwy_r = self.grad_xy(wr[None], 0)[0], self.grad_xy(wr[None], 1)[0]
        res = []
        if mode == 0:
            for b in range(self.nb):
                res.append((self.side_vals(dwy_r, b), -self.side_vals(dwx_r, b), None, None))
            return res
        bk = self.beta * mode
        wi = wz(Ui, Vi)
        dwx_i, dwy_i = self.grad_xy(wi[None], 0)[0], self.grad_xy(wi[None], 1)[0]
        gWr = (self.grad_xy(Wr[None], 0)[0], self.grad_xy(Wr[None], 1)[0])
        gWi = (self.grad_xy(Wi[None], 0)[0], self.grad_xy(Wi[None], 1)[0])
        for b in range(self.nb):
            sv = lambda A: self.side_vals(A, b)  # noqa: E731
            xr = sv(dwy_r) - bk * sv(gWi[0]) + bk * bk * sv(Ur)
            yr = -sv(dwx_r) - bk * sv(gWi[1]) + bk * bk * sv(Vr)
            xi = sv(dwy_i) + bk * sv(gWr[0]) + bk * bk * sv(Ui)
            yi = -sv(dwx_i) + bk * sv(gWr[1]) + bk * bk * sv(Vi)
            res.append((xr, yr, xi, yi))
        return res

    def maintain_fourier(self, U, N):
        """BCmgr::maintainFourier (src/bcmgr.cpp:1037-1296), the parts that
        feed the computed pressure BC: -d(u.n)/dt + n.(N+f) - nu n.curlcurl u."""
        step, nz, npn = self.step_no, self.nz, self.np_
        Je = min(step - 1, self.nord)
        alpha = SS[Je] if Je else None
        self.hopbc = np.roll(self.hopbc, 1, axis=0)
        self.hopbc[0] = 0.0
        un_new = np.zeros((nz, self.nline))
        for b in range(self.nb):
            sl = slice(b * npn, (b + 1) * npn)
            nx, ny = self.seg_nx[b], self.seg_ny[b]
            for k in range(nz):
                if nz > 1 and k == 1:
                    continue
                un_new[k, sl] = self.side_vals(U[0][k], b) * nx + self.side_vals(U[1][k], b) * ny
                if Je:
                    # as src/bcmgr.cpp:1124-1130 forms it: every alpha[q+1]
                    # multiplies the previous step's u.n (_un[0], before the
                    # roll), not level q
                    work = alpha[0] * un_new[k, sl]
                    for q in range(Je):
                        work = work + alpha[q + 1] * self.un[0, k, sl]
                    self.hopbc[0, k, sl] = -work / self.dt
                self.hopbc[0, k, sl] += (self.side_vals(N[0][k], b) * nx +
                                         self.side_vals(N[1][k], b) * ny)
        self.un = np.roll(self.un, 1, axis=0)
        self.un[0] = un_new
        nu = self.nu
        cc = self.curl_curl(U[0][0], None, U[1][0], None, None, None, 0)
        for b, (xr, yr, _, _) in enumerate(cc):
            sl = slice(b * npn, (b + 1) * npn)
            self.hopbc[0, 0, sl] += -nu * xr * self.seg_nx[b] - nu * yr * self.seg_ny[b]
        for m in range(1, nz // 2):
            kr, ki = 2 * m, 2 * m + 1
            cc = self.curl_curl(U[0][kr], U[0][ki], U[1][kr], U[1][ki], U[2][kr], U[2][ki], m)
            for b, (xr, yr, xi, yi) in enumerate(cc):
                sl = slice(b * npn, (b + 1) * npn)
                nx, ny = self.seg_nx[b], self.seg_ny[b]
                self.hopbc[0, kr, sl] += -nu * xr * nx - nu * yr * ny
                self.hopbc[0, ki, sl] += -nu * xi * nx - nu * yi * ny

    def pressure_lines(self):
        """BCmgr::evaluateCNBCp (src/bcmgr.cpp:1299-1335): extrapolated
        dp/dn on every boundary segment, every plane but Nyquist."""
        Je = min(self.step_no, self.nord)
        beta = EX[Je]
        lines = np.zeros((self.nz, self.nline))
        for q in range(Je):
            lines += beta[q] * self.hopbc[q]
        if self.nz > 1:
            lines[1] = 0.0
        return lines

    # ---------------------------------------------------------------- step
    def step(self):
        """One pass of the integrate() loop body (dns/integrate.cpp:138-200)."""
        dt, nu, nord = self.dt, self.nu, self.nord
        U, N = self.skew_symmetric()
        # Us[0] <- U, Uf[0] <- N (the first time levels)
        self.Us[0] = U
        self.Uf[0] = N
        self.step_no += 1
        if self.nb:
            self.maintain_fourier(U, N)
            plines = self.pressure_lines()
        else:
            plines = None
        # waveProp
        Je = min(self.step_no, nord)
        alpha, beta = SS[Je], [b * dt for b in EX[Je]]
        H = np.zeros_like(U)
        for q in range(Je):
            H += -alpha[q + 1] * self.Us[q] + beta[q] * self.Uf[q]
        u_old = self.Us[0].copy()                 # D->u after swapData
        self.Uf = self.Uf[-1:] + self.Uf[:-1]     # rollm(Uf)
        # setPForce: div(u^) / dt
        fp = self.grad_xy(H[0], 0) + self.grad_xy(H[1], 1)
        if self.ncom == 3:
            fp = fp + self.grad_z(H[2])
        fp = fp / dt
        self.p = self.solve("p", 0.0, fp, plines, self.p)
        # project: forcing -u^/(dt nu) + grad p / nu
        F = np.empty_like(H)
        gp = [self.grad_xy(self.p, 0), self.grad_xy(self.p, 1)]
        if self.ncom == 3:
            gp.append(self.grad_z(self.p))
        for i in range(self.ncom):
            F[i] = H[i] * (-1.0 / (dt * nu)) + gp[i] * (1.0 / nu)
        self.Uf[0] = F
        # Us: store u^n and roll
        self.Us[0] = u_old
        self.Us = self.Us[-1:] + self.Us[:-1]
        # viscous substep (first-order constant until the history is full)
        lam2 = SS[Je][0] / (dt * nu)
        for i in range(self.ncom):
            li = None if self.ulines is None else self.ulines[i]
            if self.startup == "pcg" and self.step_no < nord:
                self.u[i] = self.solve_pcg(lam2, F[i], u_old[i], li)
            else:
                self.u[i] = self.solve("u", lam2, F[i], li, u_old[i])
        return self

    def solve_pcg(self, lam2, F, u0, lines=None):
        """JACPCG velocity solve of every plane from the current field u0
        (Field::solve, src/field.cpp:711-805); the Nyquist plane keeps u0."""
        out = u0.copy()
        for k in range(self.nz):
            if self.nz > 1 and k == 1:
                continue
            key = ("upcg", lam2, k >> 1)
            if key not in self._sys:
                self._sys[key] = O.OracleMsys(self.om_u, lam2, (self.beta * (k >> 1)) ** 2)
            out[k], _ = self._sys[key].solve_plane_pcg(F[k], None if lines is None else lines[k],
                                                       u0[k], self.tol_rel, self.step_max)
        return out
