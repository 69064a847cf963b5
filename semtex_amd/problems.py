"""Workload builders for the Helmholtz solve: meshes, forcing and boundary
data of the benchmark configurations (synthetic, seeded).

  helmholtz_2d_square   4x4 elements on the unit square, N_P=8, lambda2=1,
                        u = sin(pi x) sin(pi y), f = -(2 pi^2 + 1) u.
  helmholtz_fourier     16x16 elements on [0,2pi]x[-1,1], N_P=10, N_Z=64
                        (32 modes, beta=1), lambda2=1 so mode k solves with
                        1+k^2; u=0 on y=+-1, periodic in x; forcing per mode
                        i.i.d. N(0,1) x (1+k)^-2 (Re and Im; mode 0 real;
                        Nyquist plane zero).
"""
from __future__ import annotations

import numpy as np

from .mesh import rect_mesh

SEED = 20261202


def helmholtz_2d_square(np_=8, nel_side=4, ordering="auto"):
    m = rect_mesh(nel_side, nel_side, np_, ordering=ordering)
    exact = np.sin(np.pi * m.x) * np.sin(np.pi * m.y)
    f = -(2 * np.pi ** 2 + 1.0) * exact
    return dict(mesh=m, lambda2=1.0, beta=0.0, nz=1, nmodes=1,
                forcing=f.reshape(1, -1), exact=exact.reshape(1, -1))


def fourier_forcing(mesh, nz, seed=SEED, amp_pow=2.0):
    """[nz][plane_size] Fourier planes, amplitude (1+k)^-amp_pow."""
    rng = np.random.default_rng(seed)
    ps = mesh.plane_size
    n = mesh.nel * mesh.npnp
    F = np.zeros((nz, ps))
    for k in range(nz):
        if k == 1 and nz > 1:
            continue                       # Nyquist plane stays zero
        mode = k >> 1
        F[k, :n] = rng.standard_normal(n) * (1.0 + mode) ** (-amp_pow)
    return F


def helmholtz_fourier(nx=16, ny=16, np_=10, nz=64, seed=SEED, ordering="auto"):
    m = rect_mesh(nx, ny, np_, 0.0, 2 * np.pi, -1.0, 1.0,
                  bc=("D", "N", "D", "N"), periodic_x=True, ordering=ordering)
    return dict(mesh=m, lambda2=1.0, beta=1.0, nz=nz, nmodes=nz // 2,
                forcing=fourier_forcing(m, nz, seed))


def dns_channel(nx=12, ny=24, np_=8, nz=128, Re=2000.0, dt=1e-3, Lz=2 * np.pi,
                amp=1e-2, seed=SEED, ordering="auto"):
    """Forced periodic channel for the DNS step: nx x ny elements on
    [0,4pi]x[-1,1], N_P, N_Z planes with Lz = 2pi (beta = 2pi/Lz = 1),
    KINVIS = 1/Re, CONST_X = 2 KINVIS (holds u = 1 - y^2), no-slip walls
    on y = +-1 with computed pressure BCs, periodic x.  Initial field:
    laminar Poiseuille plus i.i.d. U(-amp, amp) in u, v, w at every node and
    plane (seeded), made single-valued on shared element-boundary nodes.
    Returns the velocity and pressure meshes and the physical-space
    initial state [3][nz][nel*np*np]."""
    from .mesh import edgemaps
    bcu, bcp = ("D", "N", "D", "N"), ("N", "N", "N", "N")
    mu = rect_mesh(nx, ny, np_, 0.0, 4 * np.pi, -1.0, 1.0, bc=bcu, periodic_x=True,
                   ordering=ordering)
    mp = rect_mesh(nx, ny, np_, 0.0, 4 * np.pi, -1.0, 1.0, bc=bcp, periodic_x=True,
                   ordering=ordering)
    nu = 1.0 / Re
    n = mu.nel * mu.npnp
    rng = np.random.default_rng(seed)
    y = mu.y.ravel()
    u0 = np.zeros((3, nz, n))
    em, _ = edgemaps(np_)
    ext = em[:mu.next]
    for i in range(3):
        u0[i] = rng.uniform(-amp, amp, (nz, n))
        if i == 0:
            u0[i] += 1.0 - y ** 2
        for k in range(nz):
            f = u0[i, k].reshape(mu.nel, -1)
            g = np.zeros(mu.nglobal)
            g[mu.btog] = f[:, ext]            # last writer wins on shared nodes
            f[:, ext] = g[mu.btog]
    return dict(mesh_u=mu, mesh_p=mp, nz=nz, beta=2 * np.pi / Lz, kinvis=nu, dt=dt,
                const_force=(2 * nu, 0.0, 0.0), u0=u0, Lz=Lz)


def uneven_batch(count=64, seed=SEED):
    """Independent 2D Helmholtz problems with seeded parameters: N_P uniform
    in {4..12}, nx, ny uniform in {4..20}, rectangles [0, a] x [0, 1] with
    aspect a uniform in [1, 4], lambda2 log-uniform in [1e-2, 1e2], N(0,1)
    nodal forcing.  The first half are Cartesian with u = 0 on all four
    sides; the second half are cylindrical (axisymmetric mode 0) with the
    axis on y = 0, u = 0 on y = 1 and natural ends, their forcing weighted
    by r as the reference's callers supply it (Element::weight,
    src/element.cpp:800-806, does not multiply by r)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        np_ = int(rng.integers(4, 13))
        nx, ny = int(rng.integers(4, 21)), int(rng.integers(4, 21))
        a = float(rng.uniform(1.0, 4.0))
        lam2 = float(10.0 ** rng.uniform(-2.0, 2.0))
        cyl = i >= count // 2
        if cyl:
            m = rect_mesh(nx, ny, np_, 0.0, a, 0.0, 1.0, bc=("A", "N", "D", "N"), cyl=1)
        else:
            m = rect_mesh(nx, ny, np_, 0.0, a, 0.0, 1.0, bc=("D", "D", "D", "D"))
        f = rng.standard_normal((m.nel, m.npnp))
        if cyl:
            f = f * m.y
        out.append(dict(mesh=m, lambda2=lam2, forcing=f.reshape(1, -1), cyl=cyl,
                        np=np_, nx=nx, ny=ny, aspect=a))
    return out


# axis condition per field and mode class (BCmgr::getCondition,
# src/bcmgr.cpp:478-530): "A" natural, "D" essential (zero on the axis)
AXIS_KIND = {"u": ("A", "D", "D"), "p": ("A", "D", "D"), "c": ("A", "D", "D"),
             "v": ("D", "D", "D"), "w": ("D", "A", "D")}


def cylindrical_meshes(name, nx=4, ny=3, np_=8, length=2.0, radius=1.0, wall="D",
                       ordering="auto"):
    """The numbering of field `name` of a cylindrical 3D problem for mode 0,
    mode 1 and modes >= 2 (a BoundarySys's three NumberSys,
    src/bsys.cpp:42-123): nx x ny elements on [0, length] x [0, radius],
    axis on y = 0 with the mode-dependent axis condition, `wall` on
    y = radius, periodic in x.  Classes that share an axis kind share one
    mesh object."""
    made = {}
    out = []
    for kind in AXIS_KIND[name]:
        if kind not in made:
            made[kind] = rect_mesh(nx, ny, np_, 0.0, length, 0.0, radius,
                                   bc=(kind, "N", wall, "N"), periodic_x=True, cyl=1,
                                   ordering=ordering)
        out.append(made[kind])
    return out
